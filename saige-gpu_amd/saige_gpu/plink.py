"""PLINK-1 binary fileset helpers (.bed/.bim/.fam) and the covariate
transform of fitNULLGLMM."""
import numpy as np

BED_MAGIC = bytes([0x6C, 0x1B, 0x01])

# PLINK-1 2-bit codes for an A1 allele count (missing = 0b01)
_CODE_OF_COUNT = np.array([3, 2, 0], np.uint8)   # 0 -> 0b11, 1 -> 0b10, 2 -> 0b00


def pack_bed(G, missing=None):
    """Pack an N x M matrix of A1 counts {0,1,2} (missing where `missing` is
    True) into a SNP-major .bed body (M x ceil(N/4) bytes, LSB-first)."""
    G = np.asarray(G)
    N, M = G.shape
    codes = _CODE_OF_COUNT[np.clip(G, 0, 2).astype(np.int64)]
    if missing is not None:
        codes = np.where(missing, np.uint8(1), codes)
    bpm = (N + 3) // 4
    pad = np.zeros((bpm * 4 - N, M), np.uint8)
    c = np.vstack([codes, pad]).T.reshape(M, bpm, 4).astype(np.uint8)
    body = c[:, :, 0] | (c[:, :, 1] << 2) | (c[:, :, 2] << 4) | (c[:, :, 3] << 6)
    return np.ascontiguousarray(body, dtype=np.uint8)


def write_bed(prefix, G, missing=None, chrom=None):
    """Write prefix.{bed,bim,fam} for an N x M genotype matrix."""
    N, M = np.asarray(G).shape
    body = pack_bed(G, missing)
    with open(prefix + ".bed", "wb") as f:
        f.write(BED_MAGIC)
        f.write(body.tobytes())
    chrom = np.ones(M, np.int64) if chrom is None else np.asarray(chrom)
    with open(prefix + ".bim", "w") as f:
        for j in range(M):
            f.write(f"{chrom[j]}\trs{j + 1}\t0\t{j + 1}\t2\t1\n")
    with open(prefix + ".fam", "w") as f:
        for i in range(N):
            f.write(f"s{i + 1} s{i + 1} 0 0 0 -9\n")


def read_bed_body(path, N, M):
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:3] != BED_MAGIC:
        raise ValueError(f"{path}: not a SNP-major PLINK-1 .bed")
    body = np.frombuffer(raw[3:], np.uint8)
    need = M * ((N + 3) // 4)
    if body.size < need:
        raise ValueError(f"{path}: {body.size} genotype bytes, expected {need}")
    return body[:need]


def read_bim(path):
    ids, chrom = [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            chrom.append(p[0])
            ids.append(p[1])
    return ids, chrom


def read_fam(path):
    with open(path) as f:
        return [line.split()[1] for line in f]


def covariate_transform(X):
    """Covariate_Transform (FG.R:1612-1647): X -> Q sqrt(N) of its QR
    decomposition, after dropping columns that are linear combinations of
    earlier ones.  Returns (X_new, R, kept_columns)."""
    X = np.asarray(X, np.float64)
    N = X.shape[0]
    Q, R = np.linalg.qr(X)
    keep = np.abs(np.diag(R)) > 1e-10 * max(1.0, np.abs(np.diag(R)).max())
    if not keep.all():
        X = X[:, keep]
        Q, R = np.linalg.qr(X)
    return Q * np.sqrt(N), R, np.nonzero(keep)[0]
