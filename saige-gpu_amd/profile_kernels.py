"""Profiling driver: runs the hot kernels at the benchmark workload's size
(N = 100,000, M = 150,000 unless overridden) a fixed number of times, next
to a device-to-device copy of known size, so that rocprofv3 kernel traces
and HBM counters (FETCH_SIZE, WRITE_SIZE) of each can be read per dispatch:

    rocprofv3 --kernel-trace --stats -d OUT -o kernels -- python3 saige-gpu_amd/profile_kernels.py
    rocprofv3 --pmc FETCH_SIZE -d OUT -o fetch -- python3 saige-gpu_amd/profile_kernels.py
    rocprofv3 --pmc WRITE_SIZE -d OUT -o write -- python3 saige-gpu_amd/profile_kernels.py

Kernels exercised: decode/QC (k_decode_stats), repack, diag, the GRM SYRK
(2-bit input k_syrk_pk256 for --grm bf16; k_syrk_pk256_split, hi.hi + hi.lo
+ lo.hi from two tables, and k_mirror_diag for bf16_split; --grm takes a
comma-separated list, built in turn, the last one resident), the
autotuner's block-GEMM shapes, the tuned 64-column block GEMM (k_split_p +
k_gemm_t16 + k_tri_reduce + k_gemm_epi in the default --gemm-mode
bf16x3_tri, k_split_p + k_gemm_x3w/x3 + k_gemm_epi in bf16x3, k_gemm_part +
k_gemm_epi in fp32, through saige_grm_matvec) and its 16-column form (in
bf16x3_tri a 16-column launch, else a one-half launch), the single-column
GEMV (k_gemv_tri), one LOCO range build (copy + range SYRK in subtract mode
over --loco-frac of the markers), and the runtime's copy kernel on
COPY_BYTES.
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

COPY_BYTES = 1 << 30


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=100_000)
    ap.add_argument("--M", type=int, default=150_000)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--grm", default="bf16")
    ap.add_argument("--loco-frac", type=float, default=0.08)
    ap.add_argument("--gemm-mode", default="bf16x3_tri", choices=["bf16x3_tri", "bf16x3", "fp32"])
    args = ap.parse_args()

    import numpy as np
    from saige_gpu import Session

    N, M = args.N, args.M
    s = Session(0)
    bpm = (N + 3) // 4
    bed = s.alloc_device(M * bpm)
    s.synth_bed_device(bed, N, M, 0.01, 0.5, 0.01, 0.0, 20261202 + 2000)
    s.setminMAFforGRM(0.01)
    s.setmaxMissingRateforGRM(0.15)
    s.setgeno_device(bed, N, M)
    for prec in args.grm.split(","):
        t = s.build_grm(prec)
        print(f"GRM {prec} {t:.3f}s", flush=True)
    s.set_gemm_mode(args.gemm_mode)
    tune = s.autotune(reps=2)
    print(f"block GEMM shape rt{tune['rt']}/pd{tune['pd']}/s{tune['nsplit']} {tune['ms']:.3f} ms", flush=True)
    rng = np.random.default_rng(0)
    B = rng.standard_normal((N, 64)).astype(np.float32)
    for _ in range(args.reps):
        s.grm_matvec(B)
    for _ in range(args.reps):
        s.grm_matvec(B[:, :16])
    for _ in range(args.reps):
        s.getCrossprodMatAndKin(B[:, 0])
    if args.loco_frac > 0:
        mqc = s.getMsub_MAFge_minMAFtoConstructGRM()
        n_chr = max(1, int(args.loco_frac * mqc))
        s.setStartEndIndex(0, n_chr - 1, 0)
        print(f"LOCO range of {n_chr} markers built", flush=True)
    a = s.alloc_device(COPY_BYTES)
    b = s.alloc_device(COPY_BYTES)
    for _ in range(args.reps):
        s.copy_device(b, a, COPY_BYTES)
    s.free_device(a)
    s.free_device(b)
    s.free_device(bed)
    s.close()
    print("profile driver done", flush=True)


if __name__ == "__main__":
    main()
