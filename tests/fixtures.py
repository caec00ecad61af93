"""Loaders for the golden fixtures under tests/golden/ and small synthetic
inputs shared by the CPU and GPU tests."""
import gzip
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_bed(name, N):
    raw = gzip.open(os.path.join(GOLDEN, name + ".bed.gz")).read()
    assert raw[:3] == bytes([0x6C, 0x1B, 0x01])
    ids = [l.split()[1] for l in open(os.path.join(GOLDEN, name + ".bim"))]
    M = len(ids)
    body = np.frombuffer(raw[3:], np.uint8)
    assert body.size == M * ((N + 3) // 4)
    return body, ids


def load_acount():
    out = {}
    with open(os.path.join(GOLDEN, "acount_subset.tsv")) as f:
        f.readline()
        for line in f:
            p = line.rstrip("\n").split("\t")
            out[p[1]] = (int(p[4]), int(p[5]))   # ALT_CTS, OBS_CT
    return out


def load_vmiss():
    out = {}
    with open(os.path.join(GOLDEN, "vmiss_subset.tsv")) as f:
        f.readline()
        for line in f:
            p = line.rstrip("\n").split("\t")
            out[p[1]] = int(p[2])
    return out


def load_frq():
    out = {}
    with gzip.open(os.path.join(GOLDEN, "grm1000x10k.frq.gz"), "rt") as f:
        f.readline()
        for line in f:
            p = line.split()
            out[p[1]] = float(p[4])
    return out


def load_pheno():
    rows = [l.split() for l in open(os.path.join(GOLDEN, "pheno_1000samples.txt"))][1:]
    y = np.array([float(r[0]) for r in rows])
    x1 = np.array([float(r[1]) for r in rows])
    x2 = np.array([float(r[2]) for r in rows])
    iid = [r[3] for r in rows]
    return y, x1, x2, iid


def brute_decode(body, N, M):
    """Per-sample decoder written from the PLINK-1 .bed specification:
    sample s of marker j sits in byte j*ceil(N/4) + s//4 at bits 2*(s%4),
    2*(s%4)+1 (low bit first); 00 = two A1 alleles, 01 = missing,
    10 = heterozygous, 11 = zero A1 alleles."""
    bpm = (N + 3) // 4
    body = np.asarray(body, np.uint8).reshape(-1)
    G = np.zeros((N, M), np.int64)
    miss = np.zeros((N, M), bool)
    table = {0: 2, 2: 1, 3: 0}
    for j in range(M):
        for s in range(N):
            byte = int(body[j * bpm + s // 4])
            code = (byte >> (2 * (s % 4))) & 3
            if code == 1:
                miss[s, j] = True
            else:
                G[s, j] = table[code]
    return G, miss


def random_bed(N, M, miss_rate, seed, maf_lo=0.0, maf_hi=0.5):
    rng = np.random.default_rng(seed)
    p = rng.uniform(maf_lo, maf_hi, M)
    G = rng.binomial(2, p[None, :], size=(N, M))
    miss = rng.random((N, M)) < miss_rate
    from saige_gpu.plink import pack_bed
    return pack_bed(G, miss), G, miss


def round16(S):
    """FP32 values rounded to nearest (ties to even) at 16 significant bits:
    the split GRM's storage."""
    b = np.ascontiguousarray(S, np.float32).view(np.uint32)
    return ((b + np.uint32(0x7F) + ((b >> np.uint32(8)) & np.uint32(1))) & np.uint32(0xFFFFFF00)).view(np.float32)


def on_16bit_grid(K, m):
    """True when K, as the GRM block accessor returns it (FP32 S / m), comes
    from an S whose values all have at most 16 significant bits: K * m lies
    within an FP32 ulp of S, which round16 then recovers exactly."""
    K = np.asarray(K, np.float32)
    S = round16((K.astype(np.float64) * m).astype(np.float32))
    return bool(np.array_equal(S / np.float32(m), K))
