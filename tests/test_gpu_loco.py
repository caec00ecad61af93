"""GPU parity of leave-one-chromosome-out (setStartEndIndex FG.cpp:3067,
set_Diagof_StdGeno_LOCO FG.cpp:4934-4958, parallelCrossProd_LOCO
FG.cpp:1790-1851, getDiagOfSigma_LOCO FG.cpp:2368-2392, getCrossprod_LOCO
FG.cpp:2430-2460, getPCG1ofSigmaAndVector_LOCO FG.cpp:2943-3030) and of the
LOCO loop (FG.R:1240-1272, 1486-1512) against the oracle's complement-marker
FP64 GRM.

Tolerances (M = Mqc, M_loco = M - range size):
  K_loco columns, FP32 SYRK : 2 * 4 sqrt(M) 2^-24 (M / M_loco) max|K64|
                              (S and the range product each carry the FP32
                              SYRK bound; both are divided by M_loco)
  K_loco columns, BF16      : 2 * 4e-3 (M / M_loco) max|K64|
  K_loco columns, split BF16: as FP32 (three BF16 piece products)
  diag(Sigma_loco)          : bit-exact (same FP32 operations in the same order)
  Sigma_loco b              : 4e-5 max|ref|
  PCG (tol 1e-5)            : 2e-3 max|x64|, ||Sigma x - b||^2 <= 1.5 tol
  LOCO loop (tolPCG 1e-5)   : coefficients rtol 2e-3, eta / mu atol 2e-3 max|.|
"""
import numpy as np
import pytest

from fixtures import load_bed
from oracle import oracle as O
from saige_gpu import SaigeError
from saige_gpu import driver as D
from saige_gpu import plink

pytestmark = pytest.mark.gpu

RANGE = (2000, 5499)


@pytest.fixture(scope="module")
def ref():
    body, ids = load_bed("grm1000x10k", 1000)
    og = O.setgeno(body, 1000, len(ids), min_maf=0.01, max_missing=0.15)
    return body, len(ids), og, O.loco_grm(og, *RANGE)


@pytest.fixture
def loaded(session, ref):
    body, M, og, L = ref
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setminMAC_VarianceRatio(20, -1, False)
    session.setgeno_bytes(body, 1000, M)
    return og, L


def _columns(session, idx):
    cols = []
    for i in idx:
        e = np.zeros(1000, np.float32)
        e[i] = 1
        cols.append(session.getCrossprodMatAndKin_LOCO(e))
    return np.column_stack(cols)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "bf16_panel", "bf16_split"])
def test_loco_grm_and_diag(session, loaded, monkeypatch, prec):
    og, L = loaded
    print(f"case: LOCO GRM {prec} range {RANGE}", flush=True)
    monkeypatch.delenv("SAIGE_GRM_PANEL_MARKERS", raising=False)
    session.build_grm(prec)
    with pytest.raises(SaigeError, match="LOCO"):
        session.getCrossprodMatAndKin_LOCO(np.ones(1000, np.float32))
    session.setStartEndIndex(*RANGE, 0)
    M = og["Mqc"]
    m_loco = M - (RANGE[1] - RANGE[0] + 1)
    assert session.loco_range() == (RANGE[0], RANGE[1], 0, m_loco)
    idx = [0, 1, 17, 333, 500, 998, 999]
    C = _columns(session, idx)
    ref = L.K[:, idx]
    base = 4.0 * np.sqrt(M) * 2.0 ** -24 if prec in ("fp32", "bf16_split") else 4e-3
    if prec == "bf16_split":
        base += 2.0 ** -16                # stored at 16 significant bits
    tol = 2 * base * (M / m_loco) * np.max(np.abs(L.K))
    err = np.max(np.abs(C - ref))
    print(f"  max abs err {err:.3e}  tol {tol:.3e}")
    assert err <= tol
    # and it is not the full GRM: the range moves K by far more than the error
    full = np.column_stack([session.getCrossprodMatAndKin(np.eye(1000, dtype=np.float32)[i]) for i in idx[:2]])
    assert np.max(np.abs(full - ref[:, :2])) > 10 * err
    # diag(Sigma_loco): bit-exact against the oracle's FP32 restatement
    rng = np.random.default_rng(4)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    tau = np.array([0.9, 0.41], np.float32)
    d = session.getDiagOfSigma_LOCO(w, tau)
    dl, _ = O.loco_diag_std(og, *RANGE)
    do = O.diag_sigma_f32(dl, m_loco, w, tau)
    assert np.array_equal(d.view(np.uint32), do.view(np.uint32))


def test_loco_sigma_and_pcg(session, loaded):
    og, L = loaded
    print("case: LOCO Sigma / pcg1", flush=True)
    session.build_grm("fp32")
    session.setStartEndIndex(*RANGE, 0)
    rng = np.random.default_rng(5)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    tau = np.array([1.0, 0.35], np.float32)
    b = rng.standard_normal(1000).astype(np.float32)
    s = session.getCrossprod_LOCO(b, w, tau)
    so = L.sigma_mul(b.astype(np.float64), w.astype(np.float64), tau.astype(np.float64))
    assert np.max(np.abs(s - so)) <= 4e-5 * np.max(np.abs(so))
    x, it = session.getPCG1ofSigmaAndVector_LOCO(w, tau, b, 500, 1e-5, return_iters=True)
    Sig = L.sigma(w.astype(np.float64), tau.astype(np.float64))
    x64 = np.linalg.solve(Sig, b.astype(np.float64))
    assert np.max(np.abs(x - x64)) <= 2e-3 * np.max(np.abs(x64))
    r = Sig @ x.astype(np.float64) - b
    assert r @ r <= 1.5e-5
    assert 0 < it < 500


@pytest.mark.parametrize("gemm", ["fp32", "bf16x3_tri"])
def test_loco_engine(session, loaded, gemm):
    """The block engine on the LOCO GRM, with the FP32-MFMA product and with
    the split product over the lower triangle (which reads the LOCO GRM's
    lower half and its diagonal blocks' upper halves)."""
    og, L = loaded
    print(f"case: LOCO engine, GEMM {gemm}", flush=True)
    session.build_grm("fp32")
    session.set_gemm_mode(gemm)
    try:
        _loco_engine_body(session, og, L)
    finally:
        session.set_gemm_mode("fp32")


def _loco_engine_body(session, og, L):
    rng = np.random.default_rng(6)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    tau = np.array([1.0, 0.5], np.float32)
    B = rng.standard_normal((1000, 3)).astype(np.float32)
    Xfull, _ = session.pcg_block(w, tau, B, 500, 1e-5)
    session.setStartEndIndex(*RANGE, 0)
    session.engine_reserve(64, 2)
    session.engine_use_loco(True)
    # getSigma_X keeps using the full GRM while the engine points at LOCO
    Xfull2, _ = session.pcg_block(w, tau, B, 500, 1e-5)
    assert np.array_equal(Xfull, Xfull2)
    sub = np.sort(rng.choice(1000, 700, replace=False))
    mask = np.zeros(1000, np.uint8)
    mask[sub] = 1
    session.engine_set_system(0, w, tau)
    session.engine_set_system(1, w, np.array([0.8, 0.3], np.float32), mask)
    bs = np.zeros(1000, np.float32)
    bs[sub] = B[sub, 1]
    session.engine_add(0, 0, B[:, 0], 500, 1e-5)
    session.engine_add(5, 1, bs, 500, 1e-5)
    # no range change and no GRM switch while LOCO slots are live
    with pytest.raises(SaigeError, match="live"):
        session.setStartEndIndex(0, 99, 1)
    with pytest.raises(SaigeError, match="live"):
        session.engine_use_loco(False)
    for _ in range(1000):
        done, _ = session.engine_step(1)
        if done[0] and done[5]:
            break
    x0 = session.engine_take(0)
    x5 = session.engine_take(5)
    session.engine_use_loco(False)
    w64 = w.astype(np.float64)
    x64 = np.linalg.solve(L.sigma(w64, tau.astype(np.float64)), B[:, 0].astype(np.float64))
    assert np.max(np.abs(x0 - x64)) <= 2e-3 * np.max(np.abs(x64))
    Ss = L.sigma(w64, np.array([0.8, 0.3]))[np.ix_(sub, sub)]
    xs = np.linalg.solve(Ss, bs[sub].astype(np.float64))
    assert np.all(x5[mask == 0] == 0)
    assert np.max(np.abs(x5[sub] - xs)) <= 2e-3 * np.max(np.abs(xs))


def test_loco_errors(session, loaded):
    og, L = loaded
    print("case: LOCO errors", flush=True)
    session.build_grm("fp32")
    M = og["Mqc"]
    for s, e in [(-1, 10), (10, 9), (0, M), (0, M - 1)]:
        with pytest.raises(SaigeError):
            session.setStartEndIndex(s, e, 0)
    session.setStartEndIndex(0, 9, 0)
    assert session.loco_range()[3] == M - 10
    session.build_grm("fp32")                     # a new GRM drops the LOCO one
    assert session.loco_range() == (-1, -1, -1, -1)
    with pytest.raises(SaigeError, match="LOCO"):
        session.getPCG1ofSigmaAndVector_LOCO(np.ones(1000), np.array([1.0, 0.1]), np.ones(1000))
    with pytest.raises(SaigeError, match="LOCO"):
        session.engine_use_loco(True)


@pytest.mark.parametrize("trait", ["binary", "quantitative"])
def test_loco_loop_vs_oracle(session, loaded, trait):
    """The LOCO loop over three chromosomes (the fixture's QC'd markers cut
    in three), batched and plain, against the oracle's Get_Coef_LOCO."""
    og, _ = loaded
    print(f"case: LOCO loop {trait}", flush=True)
    session.build_grm("fp32")
    M = og["Mqc"]
    chr_vec = ["1"] * 3000 + ["2"] * 4000 + ["3"] * (M - 7000)
    loco, starts, ends = D.update_chr_start_end_index(chr_vec)
    assert loco
    N = 1000
    rng = np.random.default_rng(12)
    X = np.column_stack([np.ones(N), rng.normal(0, 1, N)])
    Xt, _, _ = plink.covariate_transform(X)
    gz = O.std_matrix(og).astype(np.float64)[:, :200] @ rng.normal(0, np.sqrt(0.5 / 200), 200)
    if trait == "binary":
        y = (rng.random(N) < 1 / (1 + np.exp(-(-0.5 + gz)))).astype(float)
        fam, theta = D.Binomial, np.array([1.0, 0.4])
    else:
        y = gz + rng.normal(0, 0.7, N)
        fam, theta = D.Gaussian, np.array([0.5, 0.45])
    st = D.glm_start(y, Xt, fam, np.zeros(N))
    fit = {"theta": theta, "coefficients": st["coef"], "linear_predictors": st["eta"]}
    jobs = [{"fit": fit, "trait": trait, "y": y, "X": Xt}]
    opt = D.Options()
    batched = D.loco(session, jobs, starts, ends, opt, batched=True)[0]
    plain = D.loco(session, jobs, starts, ends, opt, batched=False)[0]
    assert [r["isLOCO"] for r in batched] == [True, True, True] + [False] * 19
    for j in range(3):
        Lj = O.loco_grm(og, starts[j], ends[j])
        ref = O.loco_chromosome(Lj, fit, y, Xt, trait)
        for got in (batched[j], plain[j]):
            assert np.allclose(got["coefficients"], ref["coefficients"], rtol=2e-3, atol=1e-6)
            for k in ("linear_predictors", "fitted_values"):
                assert np.max(np.abs(got[k] - ref[k])) <= 2e-3 * np.max(np.abs(ref[k])), (j, k)


@pytest.mark.parametrize("gemm", ["fp32", "bf16x3_tri"])
def test_loco_in_place(session, loaded, monkeypatch, gemm):
    """In-place mode (no second N^2 buffer): the range is subtracted from S
    itself, a new range first adds the old one back, and the next full-GRM
    use restores S to within a few FP32 ulps.  The engine's solve on the
    in-place LOCO GRM runs with each block-GEMM mode."""
    og, L = loaded
    print(f"case: LOCO in place, GEMM {gemm}", flush=True)
    session.set_gemm_mode(gemm)
    try:
        _loco_in_place_body(session, og, L, monkeypatch)
    finally:
        session.set_gemm_mode("fp32")


def _loco_in_place_body(session, og, L, monkeypatch):
    monkeypatch.setenv("SAIGE_LOCO_INPLACE", "1")
    session.build_grm("fp32")
    K0 = session.grm_block(0, 0, 1000, 1000)
    M = og["Mqc"]
    idx = [0, 17, 500, 999]
    for rng_ in [(4000, 6499), RANGE]:
        session.setStartEndIndex(*rng_, 0)
        C = _columns(session, idx)
        Lr = L if rng_ == RANGE else O.loco_grm(og, *rng_)
        m_loco = M - (rng_[1] - rng_[0] + 1)
        tol = 2 * 4.0 * np.sqrt(M) * 2.0 ** -24 * (M / m_loco) * np.max(np.abs(Lr.K))
        assert np.max(np.abs(C - Lr.K[:, idx])) <= tol, rng_
    # the batched engine runs on the in-place LOCO GRM too
    rng = np.random.default_rng(9)
    w = rng.uniform(0.05, 0.25, 1000).astype(np.float32)
    tau = np.array([1.0, 0.4], np.float32)
    b = rng.standard_normal(1000).astype(np.float32)
    session.engine_reserve(64, 1)
    session.engine_use_loco(True)
    session.engine_set_system(0, w, tau)
    session.engine_add(0, 0, b, 500, 1e-5)
    # a full-GRM call would have to restore S under a live LOCO slot: refused
    with pytest.raises(SaigeError, match="in-place"):
        session.getCrossprodMatAndKin(b)
    while not session.engine_step(20)[0][0]:
        pass
    x = session.engine_take(0)
    session.engine_use_loco(False)
    x64 = np.linalg.solve(L.sigma(w.astype(np.float64), tau.astype(np.float64)), b.astype(np.float64))
    assert np.max(np.abs(x - x64)) <= 2e-3 * np.max(np.abs(x64))
    # the next full-GRM use adds the range back
    K1 = session.grm_block(0, 0, 1000, 1000)
    assert session.loco_range() == (-1, -1, -1, -1)
    assert np.max(np.abs(K1 - K0)) <= 8 * 2.0 ** -24 * np.max(np.abs(K0))
