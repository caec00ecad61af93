"""C4 at its full size: N = 200,000 samples, M = 500,000 markers, MAF
U(0.01, 0.5), 2% missing calls, 10% of the samples in sibling pairs that
share half their genotypes; the GRM (built from BF16 standardised genotypes
on the MFMA, held as FP32: 160 GB at Npad = 200,192) resident; a binary
trait at prevalence 0.01 with logit-scale h2 = 0.3; variance ratios on 100
markers in each of four MAF bins, held out of the GRM (FG.cpp:545-547) and
solved as block-PCG right-hand sides (FG.R:2309-2368).

Checks (no dense FP64 oracle fits this size):
* a 512-sample principal subblock (256 sibling pairs) against FP64 from the
  samples' genotypes over all GRM markers (BF16 bound), sibling entries near
  0.5 and unrelated ones near 0;
* the null fit; at tau1 = 0 every variance ratio is exactly 1 (G adjusted
  for X is W-orthogonal to X, so var1 = var2);
* variance ratios at tau = (1, 0.3), where the block PCG on K does the work:
  each ratio in (0, 1) (P <= W), and two markers' ratios recomputed through
  the single-column GEMV path agree within 1e-3."""
import numpy as np
import pytest

from configs import SEED, device_cohort, fp64_grm_subblock, grm_rel_error, subblock_genotypes
from saige_gpu import driver as D
from saige_gpu import plink, synth

pytestmark = pytest.mark.gpu

N, M, N0 = 200_000, 500_000, 512
BINS = [(0.01, 0.05), (0.05, 0.1), (0.1, 0.25), (0.25, 0.5001)]


@pytest.fixture(scope="module")
def c4():
    from saige_gpu import Session
    print(f"case: C4 cohort N={N} M={M}", flush=True)
    s = Session(0)
    try:
        seed = SEED[4]
        bpm = (N + 3) // 4
        dev = device_cohort(s, N, M, 0.01, 0.5, 0.02, 0.10, seed)
        s.setminMAFforGRM(0.01)
        s.setmaxMissingRateforGRM(0.15)
        s.setminMAC_VarianceRatio(20, -1, False)
        s.setgeno_device(dev, N, M)
        st = s.marker_stats()
        qc = (st["flags"] & 1).astype(bool)
        maf = np.minimum(st["freq"], 1 - st["freq"])
        rng = np.random.default_rng(seed)
        pool = []
        for lo, hi in BINS:
            cand = np.nonzero(qc & (maf >= lo) & (maf < hi) & (st["mac"] >= 20))[0]
            pool.append(np.sort(rng.choice(cand, 100, replace=False)))
        s.setminMAC_VarianceRatio(20, -1, True)
        s.set_vr_random_markers(np.sort(np.concatenate(pool)).astype(np.int32))
        s.setgeno_device(dev, N, M)                  # the VR markers leave the GRM
        vr_index = s.getIndexVec_forVarRatio()
        assert len(vr_index) == 400
        pos = {int(m): i for i, m in enumerate(vr_index)}
        cats = [[pos[int(m)] for m in rng.permutation(c)] for c in pool]
        G, miss = subblock_genotypes(s, dev, N, M, N0)
        causal = np.sort(rng.choice(M, 1000, replace=False))
        rows = np.stack([s.copy_device_to_host(dev, bpm, offset=int(j) * bpm) for j in causal])
        s.free_device(dev)
        Gc, mc = synth.decode_bed_rows(rows, N)
        Zc = synth.standardized(Gc, mc)
        del Gc, mc, rows
        stats, mask = s.marker_stats(), s.getQCdMarkerIndex().astype(bool)
        grm_s = s.build_grm("bf16")
        Mqc = s.getMsub_MAFge_minMAFtoConstructGRM()
        print(f"  BF16 GRM {grm_s:.2f} s ({N * (N + 1) * Mqc / grm_s / 1e12:.0f} TFLOP/s), Mqc {Mqc}", flush=True)
        s.set_gemm_mode("bf16x3_tri")
        tune = s.autotune(reps=1)
        print(f"  autotune: rt{tune['rt']}/pd{tune['pd']}/s{tune['nsplit']} {tune['ms']:.2f} ms", flush=True)
        K64 = fp64_grm_subblock(G, miss, stats, mask)
        Xraw = synth.covariates(N, seed)
        X, _, _ = plink.covariate_transform(Xraw)
        y = synth.binary_phenotype(Xraw, synth.polygenic(Zc, 0.3, seed), 0.01, seed)
        yield {"s": s, "K64": K64, "X": X, "y": y, "cats": cats}
    finally:
        s.close()


def test_c4_grm_subblock_with_sibling_pairs(c4):
    s, K64 = c4["s"], c4["K64"]
    K = s.grm_block(0, 0, N0, N0)
    err = grm_rel_error(K, K64)
    i = np.arange(0, N0, 2)
    sib = K[i, i + 1].astype(np.float64)
    off = K64.copy()
    off[np.diag_indices(N0)] = 0
    off[i, i + 1] = off[i + 1, i] = 0
    print(f"  C4 GRM subblock rel error {err:.2e}; sibling entries mean {sib.mean():.3f} "
          f"(sd {sib.std():.3f}); unrelated |K| max {np.abs(off).max():.3f}", flush=True)
    assert err <= 4e-3
    assert abs(sib.mean() - 0.5) < 0.02 and np.abs(off).max() < 0.1


def test_c4_null_fit_and_variance_ratios(c4):
    s, X, y, cats = c4["s"], c4["X"], c4["y"], c4["cats"]
    opt = D.Options(nrun=30)
    hs, _ = D.fit_batched(s, [{"trait": "binary", "y": y, "X": X}], opt)
    assert hs[0].error is None, hs[0].error
    fit = hs[0].result
    print(f"  C4 null fit: theta {fit['theta']}, {fit['iterations']} iterations, prevalence {y.mean():.4f}",
          flush=True)
    vr = D.variance_ratio(s, fit, y, X, "binary", cats, opt, vr_pool=True, num_markers=100)
    r = np.concatenate(vr["ratios"])
    print(f"  C4 variance ratios at the fitted tau: {[round(t[0], 6) for t in vr['table']]}", flush=True)
    assert [len(x) for x in vr["ratios"]] == [100] * 4
    if fit["theta"][1] == 0.0:
        # the reference zeroes tau1 below tol (FG.cpp:3318-3335); then VR = 1 exactly
        assert np.all(np.abs(r - 1.0) < 1e-6)
    else:
        assert np.all((r > 0) & (r < 1))
    # the block PCG on K at tau1 > 0
    fixed = dict(fit, theta=np.array([1.0, 0.3]))
    vr3 = D.variance_ratio(s, fixed, y, X, "binary", cats, opt, vr_pool=True, num_markers=100)
    r3 = np.concatenate(vr3["ratios"])
    print(f"  C4 variance ratios at tau (1, 0.3): {[round(t[0], 6) for t in vr3['table']]}, range "
          f"{r3.min():.4f}-{r3.max():.4f}", flush=True)
    assert np.all((r3 > 0) & (r3 < 1))
    assert np.std(r3) > 0
    # two markers through the single-column path (getSigma_G: GEMV per iteration)
    mu = fit["fitted_values"]
    W = mu * (1 - mu)
    Sigma_iX = np.column_stack([s.getSigma_G(W, fixed["theta"], X[:, j], opt.maxiterPCG, opt.tolPCG)
                                for j in range(X.shape[1])]).astype(np.float64)
    noK = D.score_test_null(mu, W, y, X)
    for k in (0, 3):
        G0 = s.Get_OneSNP_Geno_forVarRatio(cats[k][0]).astype(np.float64)
        if G0.sum() / (2 * N) > 0.5:
            G0 = 2.0 - G0
        AC = G0.sum()
        G = G0 - noK["XXVX_inv"] @ (noK["XV"] @ G0)
        SiG = s.getSigma_G(W, fixed["theta"], G, opt.maxiterPCG, opt.tolPCG).astype(np.float64)
        var1 = (G @ SiG - (G @ Sigma_iX) @ np.linalg.solve(X.T @ Sigma_iX, X.T @ SiG)) / AC
        var2 = np.sum(W * G * G) / AC
        print(f"  C4 marker {cats[k][0]}: single-column ratio {var1 / var2:.6f}, block {vr3['ratios'][k][0]:.6f}",
              flush=True)
        assert abs(var1 / var2 - vr3["ratios"][k][0]) < 1e-3 * vr3["ratios"][k][0]
