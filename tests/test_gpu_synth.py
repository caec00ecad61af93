"""The device .bed generator (saige_synth_bed_device) that the benchmark
cohorts come from, against a numpy restatement of its counter-based draws,
including bytes past the 2^32nd (N = 200k, M = 500k is 2.5e10 bytes: one
launch cannot address them work-item by work-item, so the kernel strides)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z):
    z = np.uint64(z)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M64
    return z ^ (z >> np.uint64(31))


def _unif(seed, a, b, k):
    with np.errstate(over="ignore"):
        h = _mix64(np.uint64(seed) ^ _mix64((np.uint64(a) * np.uint64(0x9E3779B97F4A7C15)
                                             + _mix64((np.uint64(b) * np.uint64(0xC2B2AE3D27D4EB4F)
                                                       + np.uint64(k)) & M64)) & M64))
    return np.float32(int(h >> np.uint64(40))) * np.float32(1.0 / 16777216.0)


def _geno(seed, m, s, pj):
    return int(_unif(seed, m, s, 1) < pj) + int(_unif(seed, m, s, 2) < pj)


def expected_row(seed, m, N, maf_lo, maf_hi, miss, npairs):
    lo, hi = np.float32(maf_lo), np.float32(maf_hi)
    pj = np.float32(lo + np.float32(np.float32(hi - lo) * _unif(seed, m, 0x7FFFFFFF, 0)))
    out = np.zeros((N + 3) // 4, np.uint8)
    for s in range(N):
        g = _geno(seed, m, s, pj)
        if s < 2 * npairs and (s & 1) and _unif(seed, m, s, 4) < np.float32(0.5):
            g = _geno(seed, m, s - 1, pj)
        code = 1 if _unif(seed, m, s, 3) < np.float32(miss) else (0 if g == 2 else (2 if g == 1 else 3))
        out[s // 4] |= np.uint8(code << (2 * (s % 4)))
    return out


def test_synth_bed_matches_restatement_past_2_32_bytes(session):
    N, M = 4001, 4_300_000                       # M * ceil(N/4) = 4.304e9 bytes > 2^32
    bpm = (N + 3) // 4
    assert M * bpm > 2 ** 32
    seed, lo, hi, miss, frac = 987654321, 0.01, 0.5, 0.02, 0.1
    npairs = int(np.floor(frac * N / 2.0))
    print(f"case: synth N={N} M={M} ({M * bpm / 1e9:.2f} GB)", flush=True)
    dev = session.alloc_device(M * bpm)
    try:
        session.synth_bed_device(dev, N, M, lo, hi, miss, frac, seed)
        for m in (0, 1, M // 2, M - 2, M - 1):
            got = session.copy_device_to_host(dev, bpm, offset=m * bpm)
            assert np.array_equal(got, expected_row(seed, m, N, lo, hi, miss, npairs)), m
    finally:
        session.free_device(dev)
