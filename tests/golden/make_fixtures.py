"""Write the small golden fixtures under tests/golden/ from the data files the
reference ships in src/SAIGE/extdata/input/.

Run once in a container that has the reference mounted:
    python tests/golden/make_fixtures.py /root/reference

What is kept (data only: genotype bytes, marker ids, known answers):
  * random1000.{bed.gz,bim}   nfam_100_..._22chr_random1000.{bed,bim}  (N=1000, M=1000)
  * forcate_vr.{bed.gz,bim}   nfam_100_..._22chr.forCate_vr.{bed,bim}  (N=1000, M=2000)
  * acount_subset.tsv         the .acount rows (PLINK2 ALT_CTS, OBS_CT) for
                              exactly the marker ids of the two .bim files
  * vmiss_subset.tsv          the .vmiss rows (MISSING_CT) for the same ids
  * grm1000x10k.{bed.gz,bim}  plinkforGRM_1000samples_10kMarkers.{bed,bim}
  * grm1000x10k.frq.gz        its PLINK 1.9 .frq (MAF to 4 decimals)
  * fam1000.fam               the 1000-sample .fam of the two nfam_100 files
  * grm1000x10k.fam           the 1000-sample .fam of plinkforGRM
  * pheno_1000samples.txt     phenotype + covariates of the same 1000 samples
"""
import gzip
import os
import shutil
import sys

OUT = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    inp = os.path.join(ref_root, "src", "SAIGE", "extdata", "input")
    pre = "nfam_100_nindep_0_step1_includeMoreRareVariants_poly"

    def gz_copy(src, dst):
        with open(src, "rb") as f, open(dst, "wb") as raw:
            with gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0) as g:
                g.write(f.read())

    gz_copy(os.path.join(inp, pre + "_22chr_random1000.bed"), os.path.join(OUT, "random1000.bed.gz"))
    shutil.copyfile(os.path.join(inp, pre + "_22chr_random1000.bim"), os.path.join(OUT, "random1000.bim"))
    gz_copy(os.path.join(inp, pre + "_22chr.forCate_vr.bed"), os.path.join(OUT, "forcate_vr.bed.gz"))
    shutil.copyfile(os.path.join(inp, pre + "_22chr.forCate_vr.bim"), os.path.join(OUT, "forcate_vr.bim"))
    shutil.copyfile(os.path.join(inp, pre + "_22chr_random1000.fam"), os.path.join(OUT, "fam1000.fam"))
    gz_copy(os.path.join(inp, "plinkforGRM_1000samples_10kMarkers.bed"), os.path.join(OUT, "grm1000x10k.bed.gz"))
    shutil.copyfile(os.path.join(inp, "plinkforGRM_1000samples_10kMarkers.bim"), os.path.join(OUT, "grm1000x10k.bim"))
    gz_copy(os.path.join(inp, "plinkforGRM_1000samples_10kMarkers.frq"), os.path.join(OUT, "grm1000x10k.frq.gz"))
    shutil.copyfile(os.path.join(inp, "plinkforGRM_1000samples_10kMarkers.fam"), os.path.join(OUT, "grm1000x10k.fam"))
    shutil.copyfile(os.path.join(inp, "pheno_1000samples.txt"), os.path.join(OUT, "pheno_1000samples.txt"))

    ids = set()
    for name in ("random1000.bim", "forcate_vr.bim"):
        with open(os.path.join(OUT, name)) as f:
            for line in f:
                ids.add(line.split()[1])

    def subset(src, dst):
        with open(src) as f, open(dst, "w") as g:
            header = f.readline()
            g.write(header)
            for line in f:
                if line.split("\t")[1] in ids:
                    g.write(line)

    subset(os.path.join(inp, pre + "_22chr.acount"), os.path.join(OUT, "acount_subset.tsv"))
    subset(os.path.join(inp, pre + ".vmiss"), os.path.join(OUT, "vmiss_subset.tsv"))

    # null models saved by the reference's fitNULLGLMM (gzip'd R data files,
    # read by the tests with a data-only parser) and its variance-ratio file
    out = os.path.join(ref_root, "src", "SAIGE", "extdata", "output")
    for name in ("example_binary_fullGRM.rda", "example_quantitative_fullGRM.rda",
                 "example_binary_fullGRM.varianceRatio.txt"):
        shutil.copyfile(os.path.join(out, name), os.path.join(OUT, name))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
