"""C1 at its full size: N = 50,000 samples, M = 100,000 markers, MAF
U(0.01, 0.5), 1% missing calls, FP32 GRM (FP32 MFMA SYRK) and the plain
single-RHS FP32 PCG of getPCG1ofSigmaAndVector (FG.cpp:2593-2809) -- the
reference's own call pattern.  A dense FP64 oracle does not fit this size,
so the checks are:

* the GRM of a 1,536-sample principal subblock against FP64 from those
  samples' genotypes over all 100,000 markers (FP32 sum bound
  4 sqrt(Mqc) 2^-24 max|K64|);
* a full-N plain solve: ||Sigma x - b||^2 <= 1.5 tolPCG in FP64 on the
  GPU's own GRM;
* the subblock's solve (the block engine in FP32-MFMA mode with the
  subblock's sample mask, i.e. K_SS of the resident GRM) against a dense
  FP64 solve and against the oracle's FP32 PCG (FG.cpp:2593-2809 restated)
  on the FP64 subblock GRM;
* a quantitative (h2 = 0.4) null fit at full N through the plain path, the
  batched scheduler and the C++ drivers: the same iterations and tau within
  1e-3 of each other (GEMV in FP32 against the block GEMM's FP32-accurate
  BF16 splits);
* a null fit on the subblock's samples against the FP64 oracle driver on the
  FP64 subblock GRM: tau within 1e-3, the same iteration count."""
import numpy as np
import pytest

from configs import (SEED, causal_std, causal_std_device, device_cohort, engine_solve, fp64_grm_subblock,
                     fp64_residual2, grm_rel_error, subblock_genotypes)
from oracle import oracle as O
from saige_gpu import driver as D
from saige_gpu import plink, synth

pytestmark = pytest.mark.gpu

N, M, N0 = 50_000, 100_000, 1536


@pytest.fixture(scope="module")
def c1():
    from saige_gpu import Session
    print(f"case: C1 cohort N={N} M={M}", flush=True)
    s = Session(0)
    try:
        dev = device_cohort(s, N, M, 0.01, 0.5, 0.01, 0.0, SEED[1])
        s.setminMAFforGRM(0.01)
        s.setmaxMissingRateforGRM(0.15)
        s.setminMAC_VarianceRatio(20, -1, False)
        s.setgeno_device(dev, N, M)
        G, miss = subblock_genotypes(s, dev, N, M, N0)
        Zc = causal_std_device(s, dev, N, M, 1000, SEED[1])
        s.free_device(dev)
        stats, mask = s.marker_stats(), s.getQCdMarkerIndex().astype(bool)
        Mqc = s.getMsub_MAFge_minMAFtoConstructGRM()
        assert Mqc == int(mask.sum()) and Mqc > 0.95 * M
        grm_s = s.build_grm("fp32")
        print(f"  FP32 GRM {grm_s:.2f} s ({N * (N + 1) * Mqc / grm_s / 1e12:.0f} TFLOP/s)", flush=True)
        K64 = fp64_grm_subblock(G, miss, stats, mask)
        Xraw = synth.covariates(N, SEED[1])
        Xt, _, _ = plink.covariate_transform(Xraw)
        y = synth.quantitative_phenotype(Xraw, synth.polygenic(Zc, 0.4, SEED[1]), 0.4, SEED[1])
        yield {"s": s, "G": G, "miss": miss, "K64": K64, "Mqc": Mqc, "Xraw": Xraw, "Xt": Xt, "y": y}
    finally:
        s.close()


def test_c1_grm_subblock_fp32(c1):
    s, K64, Mqc = c1["s"], c1["K64"], c1["Mqc"]
    K = s.grm_block(0, 0, N0, N0)
    err = grm_rel_error(K, K64)
    print(f"  C1 GRM subblock max|K - K64| / max|K64| = {err:.2e}", flush=True)
    assert err <= 4.0 * np.sqrt(Mqc) * 2.0 ** -24
    # the diagonal the preconditioner uses (sequential FP32 sum, FG.cpp:665-704)
    d = s.get_DiagofKin()[:N0].astype(np.float64)
    assert np.allclose(d, np.diag(K64), rtol=1e-4)


def test_c1_plain_pcg(c1):
    s, K64 = c1["s"], c1["K64"]
    rng = np.random.default_rng(11)
    w = np.ones(N)
    tau = np.array([0.6, 0.4])
    b = rng.standard_normal(N)
    x = s.getPCG1ofSigmaAndVector(w, tau, b, 500, 1e-5)
    r2 = fp64_residual2(s, w, tau, x, b)[0]
    print(f"  C1 plain PCG at full N: ||Sigma x - b||^2 = {r2:.3e} (tol 1e-5)", flush=True)
    assert r2 <= 1.5e-5
    # the subblock: K_SS through the sample mask, FP32 MFMA products
    assert s.gemm_mode() == "fp32"
    mask = np.zeros(N, np.uint8)
    mask[:N0] = 1
    bf = np.zeros(N)
    bf[:N0] = b[:N0]
    xs, it = engine_solve(s, w, tau, bf[:, None], mask, 500, 1e-5)
    xs = xs[:N0, 0]
    S64 = tau[0] * np.eye(N0) + tau[1] * K64
    exact = np.linalg.solve(S64, b[:N0])
    e64 = float(np.max(np.abs(xs - exact)) / np.max(np.abs(exact)))
    K32 = K64.astype(np.float32)
    kd = np.diag(K64).astype(np.float32)
    x32, it32 = O.pcg_f32({"packed": None, "Mqc": 1, "freq": None, "inv_std": None}, kd, w[:N0].astype(np.float32),
                          tau.astype(np.float32), b[:N0].astype(np.float32), 500, 1e-5, Kdense=K32)
    e32 = float(np.max(np.abs(xs - x32)) / np.max(np.abs(x32)))
    print(f"  C1 subblock solve: {it[0]} iterations (oracle FP32 PCG {it32}); rel error vs FP64 {e64:.2e}, "
          f"vs oracle FP32 PCG {e32:.2e}", flush=True)
    assert e64 < 2e-3 and e32 < 2e-3
    assert abs(int(it[0]) - int(it32)) <= max(2, 0.1 * it32)


def test_c1_quantitative_fit_three_paths(c1):
    s, y, Xt = c1["s"], c1["y"], c1["Xt"]
    opt = D.Options(nrun=30)
    plain = D.fit_plain(s, "quantitative", y, Xt, opt)
    hs, _ = D.fit_batched(s, [{"trait": "quantitative", "y": y, "X": Xt}], opt)
    assert hs[0].error is None, hs[0].error
    nat = s.fit_null_batch([{"trait": "quantitative", "y": y, "X": Xt}], opt)[0]
    assert nat["status"] == 0
    print(f"  C1 quantitative fit: plain {plain['theta']} ({plain['iterations']} it), batched "
          f"{hs[0].result['theta']}, C++ {nat['theta']}", flush=True)
    assert plain["theta"][1] > 0.05
    for other in (hs[0].result, nat):
        assert other["iterations"] == plain["iterations"]
        assert np.allclose(other["theta"], plain["theta"], rtol=1e-3, atol=1e-6)
        assert np.allclose(other["coefficients"], plain["coefficients"], rtol=1e-3, atol=1e-5)


def test_c1_subset_fit_matches_fp64_oracle(c1):
    s, K64, G, miss = c1["s"], c1["K64"], c1["G"], c1["miss"]
    sub = np.arange(N0)
    Xs, _, _ = plink.covariate_transform(c1["Xraw"][sub])
    Zc = causal_std(G, miss, 500, SEED[1])
    y = synth.quantitative_phenotype(c1["Xraw"][sub], synth.polygenic(Zc, 0.6, SEED[1] + 1), 0.6, SEED[1] + 1)
    opt = D.Options(nrun=30)
    hs, _ = D.fit_batched(s, [{"trait": "quantitative", "y": y, "X": Xs, "subset": sub}], opt)
    assert hs[0].error is None, hs[0].error
    grm = O.DenseGRM(K64, 1)
    ref = O.glmmkin_quantitative(grm, y, Xs, D.RademacherProbes(N0, 200), nrun=opt.nrun, tolPCG=opt.tolPCG,
                                 maxiterPCG=opt.maxiterPCG, cv_cutoff=opt.traceCVcutoff, tol=opt.tol,
                                 maxiter=opt.maxiter)
    res = hs[0].result
    rel = np.abs(res["theta"] - ref["theta"]) / np.abs(ref["theta"])
    print(f"  C1 subset fit: theta {res['theta']} FP64 oracle {ref['theta']} (rel error {rel}), iterations "
          f"{res['iterations']} / {ref['iterations']}", flush=True)
    assert ref["theta"][1] > 0.02
    assert res["iterations"] == ref["iterations"]
    assert np.allclose(res["theta"], ref["theta"], rtol=1e-3, atol=1e-6)
