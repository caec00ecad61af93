"""saige_dots_device_noalloc against saige_dots_device: the allocation-free
variant computes the same FP64 column dots from its fixed per-device
scratch, bit for bit, on two calls in a row (the second with a smaller
input, so stale scratch from the first would show), and refuses a cohort
larger than its scratch."""
import ctypes

import numpy as np
import pytest

from saige_gpu import _lib, synth

pytestmark = pytest.mark.gpu


def _device_block(s, A):
    A = np.asfortranarray(A, np.float32)
    buf = s.alloc_device(A.nbytes)
    s.copy_host_to_device(buf, A)
    return buf


def test_noalloc_dots_match_allocating_entry_point():
    from saige_gpu import Session
    print("case: saige_dots_device_noalloc, N=1500", flush=True)
    with Session(0) as s:
        c = synth.small_cohort(1500, 600, seed=synth.config_seed(0))
        s.setgeno_bytes(c["bed"], c["N"], c["M"])
        N = s.getNnomissingOut()
        rng = np.random.default_rng(3)
        bufs = []
        try:
            for k, m in ((64, 16), (5, 3)):                     # the second call is smaller
                A = _device_block(s, rng.standard_normal((N, k)))
                Q = _device_block(s, rng.standard_normal((N, m)))
                bufs += [A, Q]
                ref = s.dots_device(A, k, Q, m)
                got = s.dots_device_noalloc(A, k, Q, m)
                print(f"  k={k} m={m}: max |diff| {np.max(np.abs(got - ref)):.3e}", flush=True)
                assert got.shape == (k, m)
                assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
                assert np.all(np.isfinite(got)) and np.any(got != 0)
        finally:
            for b in bufs:
                s.free_device(b)


def test_noalloc_rejects_more_samples_than_its_scratch():
    from saige_gpu import Session
    print("case: saige_dots_device_noalloc, N=270000 (over the limit)", flush=True)
    with Session(0) as s:
        N, M = 270_000, 8
        rng = np.random.default_rng(4)
        G = rng.integers(0, 3, (N, M))
        from saige_gpu.plink import pack_bed
        s.setgeno_bytes(pack_bed(G, np.zeros((N, M), bool)), N, M)
        A = s.alloc_device(4 * N)
        try:
            out = np.zeros(1, np.float64)
            rc = s._L.saige_dots_device_noalloc(s.ctx, A, N, 1, A, N, 1, out.ctypes.data_as(ctypes.c_void_p))
            assert rc == _lib.SAIGE_E_ARG
            assert b"exceed" in s._L.saige_last_error()
        finally:
            s.free_device(A)
