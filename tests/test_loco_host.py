"""Leave-one-chromosome-out on the CPU: the chromosome ranges of
updateChrStartEndIndexVec (Util.R:29-68), the oracle's LOCO GRM and diagonal
(setStartEndIndex FG.cpp:3067, set_Diagof_StdGeno_LOCO FG.cpp:4934-4958,
parallelCrossProd_LOCO FG.cpp:1790-1851) pinned against brute-force dense
algebra, and the product's LOCO driver (saige_gpu.driver.loco_co) against
the oracle's restatement of FG.R:1240-1272 / 1486-1512."""
import numpy as np
import pytest

from fixtures import random_bed
from oracle import oracle as O
from saige_gpu import driver as D
from saige_gpu import plink
from test_host_driver import DenseExecutor


def test_chr_ranges_follow_util_r():
    chr_vec = ["1", "1", "1", "2", "2", "4", "4", "4", "X", "Y"]
    loco, s, e = D.update_chr_start_end_index(chr_vec)
    assert loco
    assert len(s) == len(e) == 22
    assert (s[0], e[0]) == (0, 2) and (s[1], e[1]) == (3, 4) and (s[3], e[3]) == (5, 7)
    assert s[2] is None and e[2] is None and all(x is None for x in s[4:])
    # one autosome only: LOCO is switched off
    loco1, s1, _ = D.update_chr_start_end_index(["3"] * 5 + ["X"])
    assert not loco1 and s1[2] == 0
    # integer chromosome codes as read by fread
    loco2, s2, e2 = D.update_chr_start_end_index([1, 1, 2, 22])
    assert loco2 and (s2[21], e2[21]) == (3, 3)
    # chromosome 2 before chromosome 1
    with pytest.raises(ValueError, match="ordered"):
        D.update_chr_start_end_index(["2", "2", "1", "1"])
    # interleaved: the second chromosome's last index precedes the first's
    with pytest.raises(ValueError, match="ordered"):
        D.update_chr_start_end_index(["1", "2", "1"])


@pytest.fixture(scope="module")
def geno():
    N, M = 150, 500
    body, _, _ = random_bed(N, M, 0.01, seed=31, maf_lo=0.02, maf_hi=0.5)
    g = O.setgeno(body, N, M, min_maf=0.01)
    return g


@pytest.mark.parametrize("start,end", [(0, 99), (120, 260), (300, None)])
def test_loco_grm_pinned(geno, start, end):
    g = geno
    M = g["Mqc"]
    end = M - 1 if end is None else end
    L = O.loco_grm(g, start, end)
    cnt = end - start + 1
    # brute force: drop columns start..end (inclusive) of the dense Z
    Z = O.std_matrix(g).astype(np.float64)
    Zo = np.delete(Z, np.arange(start, end + 1), axis=1)
    assert Zo.shape[1] == M - cnt
    Kb = Zo @ Zo.T / (M - cnt)
    assert np.allclose(L.K, Kb, rtol=1e-12, atol=1e-12)
    # the reference's route, full minus range, agrees in FP64
    S = O.grm_f64(g)
    Sr = O.grm_f64(O.marker_subset(g, np.arange(start, end + 1)))
    assert np.allclose(L.K, (S - Sr) / (M - cnt), rtol=1e-10, atol=1e-10)
    # diag_loco (FP32, reference order) against the brute-force FP64 diagonal
    dl, c2 = O.loco_diag_std(g, start, end)
    assert c2 == cnt
    assert np.allclose(dl, np.sum(Zo * Zo, axis=1), rtol=2e-5)
    assert np.allclose(L.kdiag, dl / (M - cnt))
    # a one-off range is a different matrix
    L2 = O.loco_grm(g, start, end - 1) if end > start else None
    if L2 is not None:
        assert not np.allclose(L.K, L2.K, rtol=1e-6)


@pytest.mark.parametrize("trait", ["binary", "quantitative"])
def test_loco_driver_matches_oracle(geno, trait):
    g = geno
    N = g["Nnm"]
    rng = np.random.default_rng(8)
    X = np.column_stack([np.ones(N), rng.normal(0, 1, N)])
    Xt, _, _ = plink.covariate_transform(X)
    gz = O.std_matrix(g).astype(np.float64)[:, :80] @ rng.normal(0, np.sqrt(0.5 / 80), 80)
    if trait == "binary":
        y = (rng.random(N) < 1 / (1 + np.exp(-(-0.5 + gz)))).astype(float)
        fam = D.Binomial
    else:
        y = gz + rng.normal(0, 0.7, N)
        fam = D.Gaussian
    st = D.glm_start(y, Xt, fam, np.zeros(N))
    fit = {"theta": np.array([1.0, 0.3]) if trait == "binary" else np.array([0.5, 0.4]),
           "coefficients": st["coef"], "linear_predictors": st["eta"]}
    L = O.loco_grm(g, 100, 299)
    opt = D.Options(tolPCG=1e-10, maxiterPCG=2000)
    res = DenseExecutor(L, opt).run(D.loco_co(fit, y, Xt, trait, opt))
    ref = O.loco_chromosome(L, fit, y, Xt, trait, maxiterPCG=2000, tolPCG=1e-10)
    assert res["isLOCO"]
    for k in ("coefficients", "linear_predictors", "fitted_values", "Y", "residuals", "cov"):
        assert np.allclose(res[k], ref[k], rtol=1e-9, atol=1e-12), k
    for k in ("XV", "XVX", "XXVX_inv", "S_a"):
        assert np.allclose(res["obj_noK"][k], ref["obj_noK"][k], rtol=1e-9, atol=1e-12), k
    # the full GRM gives a different answer (the range matters)
    F = O.DenseGRM(O.grm_f64(g), g["Mqc"], O.diag_std(g))
    full = O.loco_chromosome(F, fit, y, Xt, trait, maxiterPCG=2000, tolPCG=1e-10)
    assert not np.allclose(full["linear_predictors"], ref["linear_predictors"], rtol=1e-6)
