"""GPU parity of the resident GRM (MFMA SYRK), its products, the Sigma
operators and the PCG solvers, against the oracle (FG.cpp:1576-1708,
2322-2425, 2593-2809) and dense FP64 linear algebra.

Tolerances (stated here, derived in DESIGN.md):
  GRM FP32 SYRK   : max |K - K64| <= 4 sqrt(Mqc) 2^-24 max|K64|  (FP32 sum of Mqc terms)
  GRM BF16 SYRK   : max |K - K64| <= 4e-3 max|K64|               (BF16 z, 2^-9 per operand)
  GRM split BF16  : as FP32 (z = hi + lo within 2^-16 |z|, lo.lo dropped),
                    plus 2^-16 for its storage at 16 significant bits
  K b (FP32)      : max |Kb - Kb64| <= 2e-5 * max|Kb64|
  PCG (tol 1e-5)  : max |x - x64| <= 2e-3 * max|x64|, ||Sigma x - b||^2 <= 1.5 tol
"""
import numpy as np
import pytest

from fixtures import load_bed, on_16bit_grid, random_bed
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grm1000_ref():
    body, ids = load_bed("grm1000x10k", 1000)
    og = O.setgeno(body, 1000, len(ids), min_maf=0.01, max_missing=0.15)
    return body, len(ids), og, O.grm_f64(og)


@pytest.fixture
def grm1000(session, grm1000_ref):
    """(Re)loads the 1000 x 10k fixture into the shared session: other tests
    load other genotypes."""
    body, M, og, S = grm1000_ref
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setminMAC_VarianceRatio(20, -1, False)
    session.setgeno_bytes(body, 1000, M)
    return og, S


def _k64(og, S):
    return S / og["Mqc"]


def fp32_tol(mqc):
    return 4.0 * np.sqrt(mqc) * 2.0 ** -24


@pytest.mark.parametrize("prec,tol,panel", [("fp32", None, None), ("fp32", None, "2048"),
                                            ("bf16", 4e-3, None), ("bf16_panel", 4e-3, None),
                                            ("bf16_panel", 4e-3, "4096"), ("bf16_split", None, None)])
def test_grm_syrk(session, grm1000, monkeypatch, prec, tol, panel):
    og, S = grm1000
    print(f"case: SYRK {prec} panel={panel}", flush=True)
    if panel:
        monkeypatch.setenv("SAIGE_GRM_PANEL_MARKERS", panel)
    else:
        monkeypatch.delenv("SAIGE_GRM_PANEL_MARKERS", raising=False)
    session.build_grm(prec)
    K = session.grm_block(0, 0, 1000, 1000)
    K64 = _k64(og, S)
    tol = tol or fp32_tol(og["Mqc"])
    if prec == "bf16_split":
        tol += 2.0 ** -16                 # stored at 16 significant bits
        assert on_16bit_grid(K, og["Mqc"])
    err = np.max(np.abs(K - K64))
    print(f"  max abs err {err:.3e}  max|K| {np.max(np.abs(K64)):.3f}  tol {tol:.2e}")
    assert err <= tol * np.max(np.abs(K64))
    assert np.array_equal(K, K.T)          # mirrored tiles are exact copies
    # diagonal against the sequential FP32 diag (independent computation)
    assert np.allclose(np.diag(K), session.get_DiagofKin(), rtol=10 * tol)


def test_bf16_two_bit_input_equals_panels(session, grm1000, monkeypatch):
    """The 2-bit-input SYRK decodes to the same BF16 operands as the
    materialised panel path and runs the same MFMA loop: bitwise equal."""
    print("case: SYRK bf16 2-bit vs panel", flush=True)
    monkeypatch.delenv("SAIGE_GRM_PANEL_MARKERS", raising=False)
    session.build_grm("bf16_panel")
    Kp = session.grm_block(0, 0, 1000, 1000)
    session.build_grm("bf16")
    Kb = session.grm_block(0, 0, 1000, 1000)
    assert np.array_equal(Kp.view(np.uint32), Kb.view(np.uint32))


def test_two_bit_syrk_tiles_agree(session, grm1000, monkeypatch):
    """The 256x256-tile and 128x128-tile 2-bit SYRKs accumulate every element
    in the same order: bitwise equal GRMs, and LOCO ranges built by each."""
    print("case: SYRK 2-bit 128 vs 256 tiles", flush=True)
    monkeypatch.delenv("SAIGE_GRM_PANEL_MARKERS", raising=False)
    out = {}
    for tile in ("128", "256"):
        monkeypatch.setenv("SAIGE_SYRK_TILE", tile)
        session.build_grm("bf16")
        K = session.grm_block(0, 0, 1000, 1000)
        session.setStartEndIndex(1000, 2999, 0)
        e = np.zeros(1000, np.float32)
        e[17] = 1
        out[tile] = (K, session.getCrossprodMatAndKin_LOCO(e))
    assert np.array_equal(out["128"][0].view(np.uint32), out["256"][0].view(np.uint32))
    assert np.array_equal(out["128"][1].view(np.uint32), out["256"][1].view(np.uint32))


def test_grm_products(session, grm1000):
    og, S = grm1000
    print("case: GRM products", flush=True)
    session.build_grm("fp32")
    K64 = _k64(og, S)
    rng = np.random.default_rng(7)
    b = rng.standard_normal(1000).astype(np.float32)
    kb = session.getCrossprodMatAndKin(b)
    ref = K64 @ b.astype(np.float64)
    assert np.max(np.abs(kb - ref)) <= 2e-5 * np.max(np.abs(ref))
    # the oracle's per-marker FP32 product (the reference's CPU path)
    kbo = O.crossprod(og, b)
    assert np.max(np.abs(kb - kbo)) <= 2e-5 * np.max(np.abs(ref))
    B = rng.standard_normal((1000, 70)).astype(np.float32)      # two GEMM launches
    KB = session.grm_matvec(B)
    RB = K64 @ B.astype(np.float64)
    assert np.max(np.abs(KB - RB)) <= 2e-5 * np.max(np.abs(RB))
    # GEMM column j equals the GEMV of column j to FP32 rounding
    assert np.allclose(KB[:, 3], session.getCrossprodMatAndKin(B[:, 3]), rtol=1e-4, atol=1e-5)


def test_sigma_operators(session, grm1000):
    og, S = grm1000
    print("case: Sigma operators", flush=True)
    session.build_grm("fp32")
    rng = np.random.default_rng(8)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    tau = np.array([1.0, 0.37], np.float32)
    d = session.getDiagOfSigma(w, tau)
    do = O.diag_sigma_f32(O.diag_std(og), og["Mqc"], w, tau)
    assert np.array_equal(d.view(np.uint32), do.view(np.uint32))
    b = rng.standard_normal(1000).astype(np.float32)
    s = session.getCrossprod(b, w, tau)
    so = O.sigma_mul_f32(b, O.crossprod(og, b), w, tau)
    assert np.max(np.abs(s - so)) <= 1e-5 * np.max(np.abs(so))
    tau0 = np.array([1.0, 0.0], np.float32)     # tau1 == 0 shortcut: bit-exact
    s0 = session.getCrossprod(b, w, tau0)
    assert np.array_equal(s0, O.sigma_mul_f32(b, np.zeros(1000, np.float32), w, tau0))


def _sigma64(og, S, w, tau):
    return tau[0] * np.diag(1.0 / w.astype(np.float64)) + tau[1] * _k64(og, S)


def test_pcg1(session, grm1000):
    og, S = grm1000
    print("case: pcg1", flush=True)
    session.build_grm("fp32")
    rng = np.random.default_rng(9)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    tau = np.array([1.0, 0.3], np.float32)
    b = rng.standard_normal(1000).astype(np.float32)
    x, it = session.getPCG1ofSigmaAndVector(w, tau, b, 500, 1e-5, return_iters=True)
    Sig = _sigma64(og, S, w, tau)
    x64 = np.linalg.solve(Sig, b.astype(np.float64))
    assert np.max(np.abs(x - x64)) <= 2e-3 * np.max(np.abs(x64))
    r = Sig @ x.astype(np.float64) - b
    assert r @ r <= 1.5e-5
    # the oracle's FP32 PCG on the same problem takes about as many steps
    xo, ito = O.pcg_f32(og, O.diag_std(og), w, tau, b, 500, 1e-5)
    assert abs(it - ito) <= max(3, ito // 10)
    assert np.max(np.abs(x - xo)) <= 2e-3 * np.max(np.abs(x64))
    # maxiter reached: x is returned, the status is "not converged"
    from saige_gpu import _lib
    import ctypes
    L = _lib.load()
    xx = np.zeros(1000, np.float32)
    itc = ctypes.c_int32(0)
    rc = L.saige_pcg1(session.ctx, _lib.ptr(w), _lib.ptr(tau), _lib.ptr(b), 2, 1e-12, _lib.ptr(xx),
                      ctypes.byref(itc))
    assert rc == _lib.SAIGE_E_NOTCONV and itc.value == 2 and np.any(xx != 0)


def test_pcg_block_and_engine(session, grm1000):
    og, S = grm1000
    print("case: block PCG / engine", flush=True)
    session.build_grm("fp32")
    rng = np.random.default_rng(10)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    tau = np.array([1.0, 0.5], np.float32)
    B = rng.standard_normal((1000, 5)).astype(np.float32)
    X, iters = session.pcg_block(w, tau, B, 500, 1e-5)
    Sig = _sigma64(og, S, w, tau)
    X64 = np.linalg.solve(Sig, B.astype(np.float64))
    assert np.max(np.abs(X - X64)) <= 2e-3 * np.max(np.abs(X64))
    for j in range(5):
        x1 = session.getPCG1ofSigmaAndVector(w, tau, B[:, j], 500, 1e-5)
        assert np.max(np.abs(X[:, j] - x1)) <= 2e-3 * np.max(np.abs(X64[:, j]))

    # engine: a masked system solves the restricted problem
    session.engine_reserve(64, 4)
    sub = np.sort(rng.choice(1000, 640, replace=False))
    mask = np.zeros(1000, np.uint8)
    mask[sub] = 1
    session.engine_set_system(0, w, tau)
    session.engine_set_system(1, w, np.array([0.8, 0.2], np.float32), mask)
    bsub = np.zeros(1000, np.float32)
    bsub[sub] = rng.standard_normal(len(sub))
    session.engine_add(0, 0, B[:, 0], 500, 1e-5)
    session.engine_add(7, 1, bsub, 500, 1e-5)
    for _ in range(1000):
        done, it = session.engine_step(1)
        if done[0] and done[7]:
            break
    x0 = session.engine_take(0)
    x7 = session.engine_take(7)
    assert np.max(np.abs(x0 - X64[:, 0])) <= 2e-3 * np.max(np.abs(X64[:, 0]))
    Ss = _sigma64(og, S, w, np.array([0.8, 0.2]))[np.ix_(sub, sub)]
    xs = np.linalg.solve(Ss, bsub[sub].astype(np.float64))
    assert np.all(x7[mask == 0] == 0)
    assert np.max(np.abs(x7[sub] - xs)) <= 2e-3 * np.max(np.abs(xs))

    # alone vs in a mix: a column's result does not depend on its neighbours
    session.engine_add(3, 0, B[:, 2], 500, 1e-5)
    while not session.engine_step(50)[0][3]:
        pass
    alone = session.engine_take(3)
    for slot in range(64):
        session.engine_add(slot, slot % 2, B[:, slot % 5] * (1 + slot) if slot != 40 else B[:, 2], 500, 1e-5)
    for _ in range(2000):
        done, _ = session.engine_step(1)
        if done.all():
            break
    mixed = session.engine_take(40)
    for slot in range(64):
        if slot != 40:
            session.engine_take(slot)
    assert np.array_equal(alone.view(np.uint32), mixed.view(np.uint32))
    # alone in the upper 32-column half: the GEMM computes only that half
    session.engine_add(50, 0, B[:, 2], 500, 1e-5)
    while not session.engine_step(50)[0][50]:
        pass
    upper = session.engine_take(50)
    assert np.array_equal(alone.view(np.uint32), upper.view(np.uint32))


def test_odd_sizes(session):
    """N not a multiple of the tiles, a subset, several panels."""
    N, M = 333, 250
    print("case: odd sizes", flush=True)
    body, G, mm = random_bed(N, M, 0.03, seed=33, maf_lo=0.02, maf_hi=0.5)
    sub = np.arange(3, N + 1, dtype=np.int32)
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setgeno_bytes(body, N, M, sub)
    og = O.setgeno(body, N, M, subidx=sub, min_maf=0.01, max_missing=0.15)
    session.build_grm("fp32")
    K = session.grm_block(0, 0, len(sub), len(sub))
    K64 = O.grm_f64(og) / og["Mqc"]
    assert np.max(np.abs(K - K64)) <= fp32_tol(og["Mqc"]) * np.max(np.abs(K64))
    w = np.full(len(sub), 0.2, np.float32)
    b = np.linspace(-1, 1, len(sub)).astype(np.float32)
    x = session.getPCG1ofSigmaAndVector(w, np.array([1.0, 0.4], np.float32), b)
    x64 = np.linalg.solve(np.diag(np.full(len(sub), 5.0)) + 0.4 * K64, b)
    assert np.max(np.abs(x - x64)) <= 2e-3 * np.max(np.abs(x64))


@pytest.mark.parametrize("N,M", [(5, 70), (257, 129), (333, 250)])
@pytest.mark.parametrize("prec", ["fp32", "bf16", "bf16_split"])
def test_grm_tiny_and_odd_sizes(session, monkeypatch, N, M, prec):
    """Cohorts far below one 256-sample tile and just past it, marker counts
    off the 64-marker chunk: every GRM precision against FP64 (the split
    within its worst case, 3 2^-16 per product, plus 2^-16 for its storage
    at 16 significant bits and the FP32 sum bound; one
    BF16 pass within 2^-7 per product), exactly symmetric; then a 64-column
    block solve in the default lower-triangle mode against a dense solve
    (at N=257 the second row set runs a short item left of its diagonal)."""
    print(f"case: tiny/odd GRM N={N} M={M} {prec}", flush=True)
    monkeypatch.delenv("SAIGE_GRM_PANEL_MARKERS", raising=False)
    body, G, mm = random_bed(N, M, 0.02, seed=N + M, maf_lo=0.05, maf_hi=0.5)
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setgeno_bytes(body, N, M)
    og = O.setgeno(body, N, M, min_maf=0.01, max_missing=0.15)
    session.build_grm(prec)
    K = session.grm_block(0, 0, N, N)
    K64 = O.grm_f64(og) / og["Mqc"]
    tol = {"fp32": fp32_tol(og["Mqc"]), "bf16": 2.0 ** -7 + fp32_tol(og["Mqc"]),
           "bf16_split": 4 * 2.0 ** -16 + fp32_tol(og["Mqc"])}[prec]     # 3 per product, 1 storage
    err = np.max(np.abs(K - K64)) / np.max(np.abs(K64))
    print(f"  max|K - K64| / max|K64| = {err:.2e} (tol {tol:.2e})", flush=True)
    assert err <= tol
    assert np.array_equal(K, K.T)
    mode = session.gemm_mode()
    try:
        session.set_gemm_mode("bf16x3_tri")
        rng = np.random.default_rng(N)
        w = rng.uniform(0.1, 0.25, N)
        tau = np.array([1.0, 0.5])
        B = rng.standard_normal((N, 64))
        X, it = session.pcg_block(w, tau, B, 500, 1e-6)
        Kd = K.astype(np.float64)
        exact = np.linalg.solve(tau[0] * np.diag(1.0 / w.astype(np.float32).astype(np.float64)) + tau[1] * Kd, B)
        rel = np.max(np.abs(X - exact)) / np.max(np.abs(exact))
        print(f"  block solve rel error {rel:.2e}, iterations {it.min()}-{it.max()}", flush=True)
        assert rel < 2e-3
    finally:
        session.set_gemm_mode(mode)


def test_engine_product_slots(session, grm1000):
    """GRM products ride along in the engine's GEMM next to live solves."""
    og, S = grm1000
    print("case: engine product slots", flush=True)
    session.build_grm("fp32")
    K64 = _k64(og, S)
    rng = np.random.default_rng(12)
    session.engine_reserve(64, 2)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    session.engine_set_system(0, w, np.array([1.0, 0.3], np.float32))
    sub = np.sort(rng.choice(1000, 700, replace=False))
    mask = np.zeros(1000, np.uint8)
    mask[sub] = 1
    session.engine_set_system(1, w, np.array([1.0, 0.3], np.float32), mask)
    B = rng.standard_normal((1000, 6)).astype(np.float32)
    session.engine_add_batch([0, 1], [0, 0], B[:, :2], 500, 1e-5)        # two solves in flight
    session.engine_add_products([5, 9, 11], [0, 0, 1], B[:, 2:5])
    done, it = session.engine_step(1)
    assert done[5] and done[9] and done[11] and it[5] == 1
    P = session.engine_take_batch([5, 9, 11])
    ref = K64 @ B[:, 2:4].astype(np.float64)
    assert np.max(np.abs(P[:, :2] - ref)) <= 2e-5 * np.max(np.abs(ref))
    refm = np.zeros(1000)
    refm[sub] = K64[np.ix_(sub, sub)] @ B[sub, 4].astype(np.float64)
    assert np.all(P[mask == 0, 2] == 0)
    assert np.max(np.abs(P[:, 2] - refm)) <= 2e-5 * np.max(np.abs(refm))
    # the solves that shared the GEMM are unaffected
    for _ in range(1000):
        done, _ = session.engine_step(1)
        if done[0] and done[1]:
            break
    X = session.engine_take_batch([0, 1])
    X64 = np.linalg.solve(_sigma64(og, S, w, np.array([1.0, 0.3])), B[:, :2].astype(np.float64))
    assert np.max(np.abs(X - X64)) <= 2e-3 * np.max(np.abs(X64))


def test_gemm_bf16x3_tri_mode(session, grm1000):
    """The split GEMM over the lower triangle (each 32x32 block used for
    S_IJ P_J and, transposed on the MFMA, S_IJ^T P_I): exact on unit vectors
    (the transpose and every partial carry one exact value), within 2^-21 of
    s*b for one nonzero entry, K B within the FP32 product tolerance, the
    same bits for both prefetch depths at one wave count and for 16-column,
    one-half and full launches, PCG solves as in FP32 mode, and an engine solve with the
    same bits alone and in a full mix.  N = 1000 pads to 1024: the last
    row set is partial (waves past the padded size read a clamped row and
    store nothing)."""
    og, S = grm1000
    print("case: GEMM bf16x3 lower-triangle mode", flush=True)
    session.build_grm("fp32")
    K64 = _k64(og, S)
    rng = np.random.default_rng(31)
    E = np.zeros((1000, 64), np.float32)
    cols = rng.choice(1000, 64, replace=False)
    E[cols, np.arange(64)] = 1.0
    session.set_gemm_config(1, 1, 4)
    ref_e = session.grm_matvec(E)
    Bv = np.zeros((1000, 64), np.float32)
    vals = (rng.uniform(0.5, 2.0, 64) / 3.0).astype(np.float32)
    Bv[cols, np.arange(64)] = vals
    ref_v = session.grm_matvec(Bv)
    B = rng.standard_normal((1000, 70)).astype(np.float32)
    RB = K64 @ B.astype(np.float64)
    KB32 = session.grm_matvec(B)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    tau = np.array([1.0, 0.5], np.float32)
    X64 = np.linalg.solve(_sigma64(og, S, w, tau), B[:, :3].astype(np.float64))
    try:
        session.set_gemm_mode("bf16x3_tri")
        by_rt = {}
        for rt, pd in [(1, 1), (1, 3), (2, 1), (2, 3)]:
            session.set_gemm_config(rt, pd, 8)
            assert session.gemm_config()["mode"] == "bf16x3_tri"
            print(f"  rt{rt}/pd{pd}", flush=True)
            got_e = session.grm_matvec(E)
            assert np.array_equal(got_e.view(np.uint32), ref_e.view(np.uint32)), (rt, pd)
            got_v = session.grm_matvec(Bv)
            assert np.all(np.abs(got_v.astype(np.float64) - ref_v) <= 2.0 ** -21 * np.abs(ref_v.astype(np.float64)))
            KB = session.grm_matvec(B)
            err, err32 = np.max(np.abs(KB - RB)), np.max(np.abs(KB32 - RB))
            print(f"  rt{rt}/pd{pd}: max|KB-K64 B| {err:.3e} (FP32 MFMA {err32:.3e})", flush=True)
            assert err <= 2e-5 * np.max(np.abs(RB)), (rt, pd)
            assert np.array_equal(KB, session.grm_matvec(B))
            if rt in by_rt:
                assert np.array_equal(KB, by_rt[rt]), (rt, pd)
            by_rt[rt] = KB
            narrow = session.grm_matvec(B[:, :5])           # a 16-column launch
            assert np.array_equal(narrow.view(np.uint32), KB[:, :5].view(np.uint32))
            half = session.grm_matvec(B[:, :20])            # a 32-column launch
            assert np.array_equal(half.view(np.uint32), KB[:, :20].view(np.uint32))
            X, _ = session.pcg_block(w, tau, B[:, :3], 500, 1e-5)
            assert np.max(np.abs(X - X64)) <= 2e-3 * np.max(np.abs(X64)), (rt, pd)
        session.set_gemm_config(1, 1, 8)
        session.engine_reserve(64, 2)
        session.engine_set_system(0, w, tau)
        session.engine_add(3, 0, B[:, 2], 500, 1e-5)
        while not session.engine_step(50)[0][3]:
            pass
        alone = session.engine_take(3)
        assert np.max(np.abs(alone - X64[:, 2])) <= 2e-3 * np.max(np.abs(X64[:, 2]))
        for slot in range(64):
            session.engine_add(slot, 0, B[:, slot % 5] * (1 + slot) if slot != 40 else B[:, 2], 500, 1e-5)
        for _ in range(2000):
            done, _ = session.engine_step(1)
            if done.all():
                break
        mixed = session.engine_take(40)
        for slot in range(64):
            if slot != 40:
                session.engine_take(slot)
        assert np.array_equal(alone.view(np.uint32), mixed.view(np.uint32))
        # the upper half alone: a one-half launch of the other tile
        session.engine_add(50, 0, B[:, 2], 500, 1e-5)
        while not session.engine_step(50)[0][50]:
            pass
        upper = session.engine_take(50)
        assert np.array_equal(alone.view(np.uint32), upper.view(np.uint32))
    finally:
        session.set_gemm_mode("fp32")
        session.set_gemm_config(1, 1, 4)


def test_scattered_slots_run_narrow_launches(session, grm1000):
    """At most 16 (32) occupied slots anywhere in the block run as one 16-
    (32-) column launch of the lower-triangle product (launch column k ->
    block column of the k-th occupied slot, free slots after them): four
    columns in slots of four different 16-column tiles take only 16-column
    launches and give the bits each gives alone in slot 0 (launch column 0,
    block column 0); 20 occupied slots spread over both halves (done slots
    stay occupied until taken) take only 32-column launches, and 40 only
    64-column ones, with the same bits again."""
    og, S = grm1000
    print("case: scattered slots in 16-column launches", flush=True)
    session.build_grm("fp32")
    rng = np.random.default_rng(77)
    B = rng.standard_normal((1000, 20)).astype(np.float32)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    tau = np.array([1.0, 0.5], np.float32)
    try:
        session.set_gemm_mode("bf16x3_tri")
        session.set_gemm_config(1, 1, 8)
        session.engine_reserve(64, 1)
        session.engine_set_system(0, w, tau)
        alone = []
        for j in range(4):
            session.engine_add(0, 0, B[:, j], 500, 1e-5)
            while not session.engine_step(50)[0][0]:
                pass
            alone.append(session.engine_take(0))
        slots = [5, 21, 38, 60]
        n_full0 = session.engine_stats("gemm_full")[1] + session.engine_stats("gemm_half")[1]
        n_q0 = session.engine_stats("gemm_quarter")[1]
        for j, sl in enumerate(slots):
            session.engine_add(sl, 0, B[:, j], 500, 1e-5)
        steps = 0
        while True:
            done, _ = session.engine_step(1)
            steps += 1
            if all(done[sl] for sl in slots):
                break
        assert session.engine_stats("gemm_quarter")[1] - n_q0 == steps
        assert session.engine_stats("gemm_full")[1] + session.engine_stats("gemm_half")[1] == n_full0
        for j, sl in enumerate(slots):
            got = session.engine_take(sl)
            assert np.array_equal(got.view(np.uint32), alone[j].view(np.uint32)), sl
        # 20 occupied slots spread over both halves: 32-column launches; 40:
        # 64-column ones
        for many, kind in ((list(range(1, 64, 3))[:20], "gemm_half"), (list(range(0, 64))[3::3][:20] +
                                                                        list(range(1, 64, 3))[:20], "gemm_full")):
            counts0 = {k: session.engine_stats(k)[1] for k in ("gemm_quarter", "gemm_half", "gemm_full")}
            for j, sl in enumerate(many):
                session.engine_add(sl, 0, B[:, j % 20], 500, 1e-5)
            steps = 0
            for _ in range(2000):
                done, _ = session.engine_step(1)
                steps += 1
                if all(done[sl] for sl in many):
                    break
            counts = {k: session.engine_stats(k)[1] - counts0[k] for k in counts0}
            assert counts[kind] == steps and sum(counts.values()) == steps, (kind, counts, steps)
            for j, sl in enumerate(many):
                got = session.engine_take(sl)
                if j < 4:
                    assert np.array_equal(got.view(np.uint32), alone[j].view(np.uint32)), sl
    finally:
        session.set_gemm_mode("fp32")
        session.set_gemm_config(1, 1, 4)


def test_gemm_tri_mode_many_items(session):
    """The lower-triangle product where row sets span several items (N =
    4700: 13 row sets of 384 rows, up to 144 chunks left of the diagonal
    cut into items of at most 96), against the FP64 product of the GPU's own
    FP32 GRM, for both wave counts, full, one-half and 16-column launches."""
    N, M = 4700, 300
    print("case: GEMM lower-triangle mode, N=4700", flush=True)
    body, G, mm = random_bed(N, M, 0.01, seed=47, maf_lo=0.05, maf_hi=0.5)
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setgeno_bytes(body, N, M)
    session.build_grm("fp32")
    K = session.grm_block(0, 0, N, N).astype(np.float64)
    rng = np.random.default_rng(5)
    B = rng.standard_normal((N, 64)).astype(np.float32)
    RB = K @ B.astype(np.float64)
    try:
        session.set_gemm_mode("bf16x3")
        session.set_gemm_config(1, 3, 8)
        KBx = session.grm_matvec(B)
        session.set_gemm_mode("bf16x3_tri")
        for rt in (1, 2):
            session.set_gemm_config(rt, 3, 8)
            KB = session.grm_matvec(B)
            err = np.max(np.abs(KB - RB))
            print(f"  rt{rt}: max|KB-K B| {err:.3e}, split mode {np.max(np.abs(KBx - RB)):.3e}", flush=True)
            assert err <= 2e-5 * np.max(np.abs(RB)), rt
            half = session.grm_matvec(B[:, :20])
            assert np.array_equal(half.view(np.uint32), KB[:, :20].view(np.uint32)), rt
            quarter = session.grm_matvec(B[:, :12])
            assert np.array_equal(quarter.view(np.uint32), KB[:, :12].view(np.uint32)), rt
        # the autotuner in this mode: one wave count for all three widths,
        # the fastest pd of each at that count
        res = session.autotune(reps=1)
        assert len(res["table"]) == 4 and len(res["half_table"]) == 4 and len(res["quarter_table"]) == 4
        assert res["half"]["rt"] == res["rt"] == res["quarter"]["rt"]
        for key, tab in (("half", res["half_table"]), ("quarter", res["quarter_table"])):
            assert res[key]["ms"] == min(t[3] for t in tab if t[0] == res["rt"]), key
        KBt = session.grm_matvec(B)
        assert np.array_equal(session.grm_matvec(B[:, :7]).view(np.uint32), KBt[:, :7].view(np.uint32))
        assert np.array_equal(session.grm_matvec(B[:, :30]).view(np.uint32), KBt[:, :30].view(np.uint32))
    finally:
        session.set_gemm_mode("fp32")
        session.set_gemm_config(1, 1, 4)


def test_split_grm_two_piece_gemm(session, monkeypatch):
    """The split GRM is stored at 16 significant bits, so the lower-triangle
    GEMM splits S into two BF16 pieces and leaves out the products of the
    third, which would be zero: K B (64 columns, a full launch; 20, one
    half) has bitwise the bits of the three-piece form
    (SAIGE_GEMM_S_PIECES=3), for both wave counts, at N = 4700 (row sets
    over several items); and the same for engine products on a LOCO range
    of the split GRM (subtracted in place, stored the same way)."""
    N, M = 4700, 300
    print("case: split GRM, two-piece GEMM, N=4700", flush=True)
    body, G, mm = random_bed(N, M, 0.01, seed=48, maf_lo=0.05, maf_hi=0.5)
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setgeno_bytes(body, N, M)
    session.build_grm("bf16_split")
    K = session.grm_block(0, 0, N, N)
    assert on_16bit_grid(K, session.getMsub_MAFge_minMAFtoConstructGRM())
    rng = np.random.default_rng(6)
    B = rng.standard_normal((N, 64)).astype(np.float32)
    RB = K.astype(np.float64) @ B.astype(np.float64)
    try:
        session.set_gemm_mode("bf16x3_tri")
        for rt in (1, 2):
            session.set_gemm_config(rt, 3, 8)
            monkeypatch.delenv("SAIGE_GEMM_S_PIECES", raising=False)
            two = session.grm_matvec(B), session.grm_matvec(B[:, :20]), session.grm_matvec(B[:, :9])
            monkeypatch.setenv("SAIGE_GEMM_S_PIECES", "3")
            three = session.grm_matvec(B), session.grm_matvec(B[:, :20]), session.grm_matvec(B[:, :9])
            for a, b in zip(two, three):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), rt
            assert np.max(np.abs(two[0] - RB)) <= 2e-5 * np.max(np.abs(RB)), rt
        # a LOCO range, subtracted in place: products through the engine
        monkeypatch.setenv("SAIGE_LOCO_INPLACE", "1")
        session.setStartEndIndex(40, 139, 0)
        session.engine_reserve(64, 1)
        session.engine_use_loco(True)
        session.engine_set_system(0, rng.uniform(0.05, 0.25, N).astype(np.float32),
                                  np.array([1.0, 0.3], np.float32))
        out = {}
        for pieces in (None, "3"):
            if pieces:
                monkeypatch.setenv("SAIGE_GEMM_S_PIECES", pieces)
            else:
                monkeypatch.delenv("SAIGE_GEMM_S_PIECES", raising=False)
            for k in (64, 20, 9):
                session.engine_add_products(list(range(k)), [0] * k, B[:, :k])
                done, _ = session.engine_step(1)
                assert all(done[:k])
                out[pieces, k] = session.engine_take_batch(list(range(k)))
        for k in (64, 20, 9):
            assert np.array_equal(out[None, k].view(np.uint32), out["3", k].view(np.uint32)), k
        session.engine_use_loco(False)
    finally:
        session.set_gemm_mode("fp32")
        session.set_gemm_config(1, 1, 4)


def test_gemv_tri_mode(session, grm1000):
    """Single-column products over the lower triangle (the default) against
    all of S and the FP64 product: K x, getCrossprod and the single-RHS PCG
    within FP32 tolerances, reproducible bits, and the mode switch."""
    og, S = grm1000
    print("case: GEMV lower-triangle mode", flush=True)
    session.build_grm("fp32")
    K64 = _k64(og, S)
    rng = np.random.default_rng(41)
    x = rng.standard_normal(1000).astype(np.float32)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    tau = np.array([1.0, 0.5], np.float32)
    ref = K64 @ x.astype(np.float64)
    assert session.gemv_mode() == "tri"
    try:
        out = {}
        for mode in ("full", "tri"):
            session.set_gemv_mode(mode)
            kx = session.getCrossprodMatAndKin(x)
            # the FP32 GRM's own error against FP64 bounds both modes
            assert np.max(np.abs(kx - ref)) <= 2e-5 * np.max(np.abs(ref)), mode
            assert np.array_equal(kx, session.getCrossprodMatAndKin(x))
            cp = session.getCrossprod(x, w, tau)
            cp64 = _sigma64(og, S, w, tau) @ x.astype(np.float64)
            assert np.max(np.abs(cp - cp64)) <= 2e-5 * np.max(np.abs(cp64)), mode
            sol = session.getPCG1ofSigmaAndVector(w, tau, x)
            x64 = np.linalg.solve(_sigma64(og, S, w, tau), x.astype(np.float64))
            assert np.max(np.abs(sol - x64)) <= 2e-3 * np.max(np.abs(x64)), mode
            out[mode] = kx
        print(f"  max|tri - full| {np.max(np.abs(out['tri'] - out['full'])):.3e}", flush=True)
        with pytest.raises(Exception):
            session.set_gemv_mode("half")
    finally:
        session.set_gemv_mode("tri")


def test_gemv_tri_many_items(session):
    """N = 7000: 28 row sets of 256 rows, the last ones with their chunks
    left of the diagonal cut into two items, against the FP64 product of the
    GPU's own FP32 GRM, elementwise within an FP32 summation bound
    (8 sqrt(N) 2^-24 sum_j |K_ij x_j|)."""
    N, M = 7000, 200
    print("case: GEMV lower-triangle mode, N=7000", flush=True)
    body, G, mm = random_bed(N, M, 0.01, seed=57, maf_lo=0.05, maf_hi=0.5)
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setgeno_bytes(body, N, M)
    session.build_grm("fp32")
    K = session.grm_block(0, 0, N, N).astype(np.float64)
    rng = np.random.default_rng(6)
    for _ in range(2):
        x = rng.standard_normal(N).astype(np.float32)
        ref = K @ x.astype(np.float64)
        bound = 8.0 * np.sqrt(N) * 2.0 ** -24 * (np.abs(K) @ np.abs(x.astype(np.float64)))
        got = session.getCrossprodMatAndKin(x)
        assert np.all(np.abs(got - ref) <= bound)


def test_autotune_and_every_gemm_shape(session, grm1000):
    """saige_autotune times the sixteen block-GEMM shapes and keeps the
    fastest; every shape gives K B within the FP32 product tolerance, and a
    fixed shape repeats bitwise."""
    og, S = grm1000
    print("case: autotune / GEMM shapes", flush=True)
    session.build_grm("fp32")
    res = session.autotune(reps=2)
    table = res["table"]
    assert len(table) == 16 and all(ms > 0 for *_, ms in table)
    best = min(table, key=lambda t: t[3])
    assert (res["rt"], res["pd"], res["nsplit"]) == best[:3] and res["ms"] == best[3]
    half = res["half_table"]
    assert len(half) == 4 and all(ns == res["nsplit"] and ms > 0 for _, _, ns, ms in half)
    hbest = min(half, key=lambda t: t[3])
    assert (res["half"]["rt"], res["half"]["pd"]) == hbest[:2] and res["half"]["ms"] == hbest[3]
    print("  " + ", ".join(f"rt{a}/pd{b}/s{c}: {t:.3f} ms" for a, b, c, t in table))
    # with the tuned shapes (full and one-half may differ in rt/pd), a column
    # computed in a one-half launch equals the same column in a full launch
    Bm = np.random.default_rng(14).standard_normal((1000, 70)).astype(np.float32)
    narrow = session.grm_matvec(Bm[:, :5])           # 5 columns: one-half launch
    wide = session.grm_matvec(Bm)                    # 70 columns: full launches
    assert np.array_equal(narrow.view(np.uint32), wide[:, :5].view(np.uint32))
    K64 = _k64(og, S)
    rng = np.random.default_rng(13)
    B = rng.standard_normal((1000, 70)).astype(np.float32)
    RB = K64 @ B.astype(np.float64)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    tau = np.array([1.0, 0.5], np.float32)
    X64 = np.linalg.solve(_sigma64(og, S, w, tau), B[:, :3].astype(np.float64))
    by_split = {}
    for rt, pd, ns, _ in table:
        session.set_gemm_config(rt, pd, ns)
        assert session.gemm_config() == {"rt": rt, "pd": pd, "nsplit": ns, "ms": 0.0, "mode": "fp32",
                                         "half": {"rt": rt, "pd": pd, "ms": 0.0},
                                         "quarter": {"rt": rt, "pd": pd, "ms": 0.0}}
        KB = session.grm_matvec(B)
        assert np.max(np.abs(KB - RB)) <= 2e-5 * np.max(np.abs(RB)), (rt, pd, ns)
        X, _ = session.pcg_block(w, tau, B[:, :3], 500, 1e-5)
        assert np.max(np.abs(X - X64)) <= 2e-3 * np.max(np.abs(X64)), (rt, pd, ns)
        assert np.array_equal(KB, session.grm_matvec(B))          # a fixed shape repeats bitwise
        # rt and pd change the schedule, not the arithmetic: same bits for one nsplit
        if ns in by_split:
            assert np.array_equal(KB, by_split[ns]), (rt, pd, ns)
        by_split[ns] = KB
    from saige_gpu import SaigeError
    with pytest.raises(SaigeError, match="shape"):
        session.set_gemm_config(3, 1, 4)
    with pytest.raises(SaigeError, match="shape"):
        session.set_gemm_config(1, 2, 4)
    session.set_gemm_config(1, 1, 4)


def test_gemm_bf16x3_mode(session, grm1000):
    """The three-piece BF16 split GEMM: exact on unit vectors (the split
    loses nothing: s = hi + mid + lo), within 2^-22 of s*b for one nonzero
    entry (only products below 2^-23 |s b| are dropped), K B within the FP32
    product tolerance, the same bits for every rt/pd at one nsplit and for
    one-half vs full launches, and PCG/engine solves as in FP32 mode."""
    og, S = grm1000
    print("case: GEMM bf16x3 mode", flush=True)
    session.build_grm("fp32")
    K64 = _k64(og, S)
    rng = np.random.default_rng(21)
    E = np.zeros((1000, 64), np.float32)
    cols = rng.choice(1000, 64, replace=False)
    E[cols, np.arange(64)] = 1.0
    session.set_gemm_config(1, 1, 4)
    ref_e = session.grm_matvec(E)                       # FP32 MFMA: s * 1 exactly
    Bv = np.zeros((1000, 64), np.float32)
    vals = (rng.uniform(0.5, 2.0, 64) / 3.0).astype(np.float32)
    Bv[cols, np.arange(64)] = vals
    ref_v = session.grm_matvec(Bv)                      # FP32 MFMA: fl(s * b) / Mqc
    B = rng.standard_normal((1000, 70)).astype(np.float32)
    RB = K64 @ B.astype(np.float64)
    KB32 = session.grm_matvec(B)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    tau = np.array([1.0, 0.5], np.float32)
    X64 = np.linalg.solve(_sigma64(og, S, w, tau), B[:, :3].astype(np.float64))
    try:
        session.set_gemm_mode("bf16x3")
        assert session.gemm_config()["mode"] == "bf16x3"
        got_e = session.grm_matvec(E)
        assert np.array_equal(got_e.view(np.uint32), ref_e.view(np.uint32))
        # one nonzero b per column: the six kept products give s * b to
        # 2^-23 before the division by Mqc; a dropped 2^-8 or 2^-16 term fails
        got_v = session.grm_matvec(Bv)
        assert np.all(np.abs(got_v.astype(np.float64) - ref_v) <= 2.0 ** -21 * np.abs(ref_v.astype(np.float64)))
        by_split = {}
        for rt, pd, ns in [(1, 1, 4), (2, 1, 4), (1, 3, 4), (2, 3, 4), (1, 1, 1), (2, 3, 8)]:
            session.set_gemm_config(rt, pd, ns)
            assert session.gemm_config()["mode"] == "bf16x3"
            KB = session.grm_matvec(B)
            err, err32 = np.max(np.abs(KB - RB)), np.max(np.abs(KB32 - RB))
            print(f"  rt{rt}/pd{pd}/s{ns}: max|KB-K64 B| {err:.3e} (FP32 MFMA {err32:.3e})", flush=True)
            assert err <= 2e-5 * np.max(np.abs(RB)), (rt, pd, ns)
            assert np.max(np.abs(KB - KB32)) <= 2e-5 * np.max(np.abs(RB))
            assert np.array_equal(KB, session.grm_matvec(B))
            if ns in by_split:
                assert np.array_equal(KB, by_split[ns]), (rt, pd, ns)
            by_split[ns] = KB
            narrow = session.grm_matvec(B[:, :5])
            assert np.array_equal(narrow.view(np.uint32), KB[:, :5].view(np.uint32))
            X, _ = session.pcg_block(w, tau, B[:, :3], 500, 1e-5)
            assert np.max(np.abs(X - X64)) <= 2e-3 * np.max(np.abs(X64)), (rt, pd, ns)
        # engine: alone vs in a full mix, bitwise
        session.set_gemm_config(1, 3, 4)
        session.engine_reserve(64, 2)
        session.engine_set_system(0, w, tau)
        session.engine_add(3, 0, B[:, 2], 500, 1e-5)
        while not session.engine_step(50)[0][3]:
            pass
        alone = session.engine_take(3)
        assert np.max(np.abs(alone - X64[:, 2])) <= 2e-3 * np.max(np.abs(X64[:, 2]))
        for slot in range(64):
            session.engine_add(slot, 0, B[:, slot % 5] * (1 + slot) if slot != 40 else B[:, 2], 500, 1e-5)
        for _ in range(2000):
            done, _ = session.engine_step(1)
            if done.all():
                break
        mixed = session.engine_take(40)
        for slot in range(64):
            if slot != 40:
                session.engine_take(slot)
        assert np.array_equal(alone.view(np.uint32), mixed.view(np.uint32))
        from saige_gpu import SaigeError
        with pytest.raises(ValueError):
            session.set_gemm_mode("fp16")
    finally:
        session.set_gemm_mode("fp32")
        session.set_gemm_config(1, 1, 4)
