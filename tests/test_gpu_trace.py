"""GPU parity of the device-resident trace terms of GetTrace / getAIScore
(FG.cpp:3113-3160, 3279-3295): probe generation, FP64 column dots, and slots
taken into device buffers with their dots, against host restatements."""
import numpy as np
import pytest

from fixtures import load_bed
from oracle import oracle as O
from saige_gpu import driver as D

pytestmark = pytest.mark.gpu


@pytest.fixture
def grm1000(session):
    body, ids = load_bed("grm1000x10k", 1000)
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setminMAC_VarianceRatio(20, -1, False)
    session.setgeno_bytes(body, 1000, len(ids))
    session.build_grm("fp32")
    og = O.setgeno(body, 1000, len(ids), min_maf=0.01, max_missing=0.15)
    return og, O.grm_f64(og) / og["Mqc"]


def test_probes_and_dots(session, grm1000):
    print("case: device probes / dots", flush=True)
    N = 1000
    U = session.alloc_device(4 * N * 12)
    Q = session.alloc_device(4 * N * 3)
    try:
        session.probes_device(U, 2, 200, 7, 10)          # probes 7..16 into columns 2..11
        got = session.copy_device_to_host(U, 4 * N * 12).view(np.float32).reshape(12, N).T
        for j in range(10):
            assert np.array_equal(got[:, 2 + j], D.probe_vector(200, 7 + j, N).astype(np.float32))
        rng = np.random.default_rng(3)
        q = rng.standard_normal((N, 3)).astype(np.float32)
        session.copy_host_to_device(Q, q)
        dots = session.dots_device(U, 12, Q, 3)
        ref = got.astype(np.float64).T @ q.astype(np.float64)
        assert np.allclose(dots, ref, rtol=1e-12, atol=1e-12 * np.max(np.abs(ref)))
    finally:
        session.free_device(U)
        session.free_device(Q)


def test_subset_probes(session, grm1000):
    """Probes of a fit on a sample subset, made on the device: entry i of
    probe i0 + j is the hash at position i of the subset, at row rows[i];
    every other row of the column is zero (whatever the buffer held)."""
    print("case: device subset probes", flush=True)
    N = 1000
    rng = np.random.default_rng(8)
    U = session.alloc_device(4 * N * 9)
    R = session.alloc_device(8 * N)
    try:
        for n in (1000, 617, 1):
            rows = np.sort(rng.choice(N, n, replace=False)).astype(np.int64)
            session.copy_host_to_device(U, np.full((N, 9), 7.0, np.float32))
            session.copy_bytes_to_device(R, rows)
            session.probes_device_rows(U, 1, 200, 5, 8, R, n)
            got = session.copy_device_to_host(U, 4 * N * 9).view(np.float32).reshape(9, N).T
            assert np.all(got[:, 0] == 7.0)                      # column before col0 untouched
            for j in range(8):
                want = np.zeros(N, np.float32)
                want[rows] = D.probe_vector(200, 5 + j, n)
                assert np.array_equal(got[:, 1 + j], want), (n, j)
    finally:
        session.free_device(U)
        session.free_device(R)


def test_take_with_dots(session, grm1000):
    """Product slots stashed on the device with dots against shared columns,
    then solve slots returning pairwise dots, equal the host computation."""
    og, K = grm1000
    print("case: take_batch_ex", flush=True)
    N = 1000
    rng = np.random.default_rng(4)
    w = rng.uniform(0.05, 0.25, N).astype(np.float32)
    tau = np.array([1.0, 0.3], np.float32)
    k = 6
    U = session.alloc_device(4 * N * k)
    AU = session.alloc_device(4 * N * 8)
    Up = session.alloc_device(4 * N * 8)
    SX = session.alloc_device(4 * N * 2)
    pc = [7, 0, 1, 2, 3, 4]                      # stash / partner column of each slot
    try:
        session.probes_device(U, 0, 200, 0, k)
        u = np.column_stack([D.probe_vector(200, j, N) for j in range(k)]).astype(np.float32)
        sx = rng.standard_normal((N, 2)).astype(np.float32)
        session.copy_host_to_device(SX, sx)
        session.engine_reserve(64, 2)
        session.engine_set_system(0, w, tau)
        slots = [3, 9, 10, 11, 40, 41]
        session.engine_add_products_device(slots, [0] * k, U, 0)
        session.engine_step(1)
        X, h = session.engine_take_batch_ex(slots, download=True, X_dev=AU, xcols=pc, Q=SX, m=2)
        au = K @ u.astype(np.float64)
        assert np.max(np.abs(X - au)) <= 2e-5 * np.max(np.abs(au))
        assert np.allclose(h, X.astype(np.float64).T @ sx.astype(np.float64), rtol=1e-12, atol=1e-9)
        stash = session.copy_device_to_host(AU, 4 * N * 8).view(np.float32).reshape(8, N).T
        assert np.array_equal(stash[:, pc], X)
        up = np.zeros((N, 8), np.float32)
        up[:, pc] = u
        session.copy_host_to_device(Up, up)
        # solves from the device probes, returning d = (K u)^T x and e = u^T x only
        session.engine_add_batch_device(slots, [0] * k, U, 0, 500, 1e-5)
        while not all(session.engine_step(1)[0][slots]):
            pass
        got, de = session.engine_take_batch_ex(slots, download=True, P1=AU, P2=Up, pcols=pc)
        xf = got.astype(np.float64)
        assert np.allclose(de[:, 0], np.einsum("ij,ij->j", X.astype(np.float64), xf), rtol=1e-12)
        assert np.allclose(de[:, 1], np.einsum("ij,ij->j", u.astype(np.float64), xf), rtol=1e-12)
        Sig = tau[0] * np.diag(1.0 / w.astype(np.float64)) + tau[1] * K
        x64 = np.linalg.solve(Sig, u.astype(np.float64))
        assert np.max(np.abs(xf - x64)) <= 2e-3 * np.max(np.abs(x64))
    finally:
        for b in (U, AU, Up, SX):
            session.free_device(b)
