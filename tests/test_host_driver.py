"""Host logic of the product's fit drivers (saige_gpu.driver, the FG.R
mirror) checked on the CPU: the coroutines are run by a test executor that
answers solve/matvec requests with the oracle's dense FP64 PCG, and must
reproduce the oracle's own restatement of FG.R step for step."""
import numpy as np
import pytest

from fixtures import random_bed
from oracle import oracle as O
from saige_gpu import driver as D
from saige_gpu import plink


class DenseExecutor:
    """Test-only executor: answers the coroutine's requests on the CPU."""

    def __init__(self, grm, opt):
        self.grm = grm
        self.opt = opt

    def run(self, gen):
        value = None
        while True:
            try:
                req = gen.send(value)
            except StopIteration as e:
                return e.value
            if req[0] == "solve":
                value = self._solve(*req[1:])
            elif req[0] == "matvec":
                value = self.grm.K @ req[1]
            else:
                value = D.trace_round_host(req, lambda B: self.grm.K @ B, self._solve)

    def _solve(self, w, tau, B):
        X, _ = self.grm.pcg(np.asarray(w), np.asarray(tau, np.float64), B, self.opt.maxiterPCG, self.opt.tolPCG)
        return X


@pytest.fixture(scope="module")
def cohort():
    N, M = 200, 600
    body, G, miss = random_bed(N, M, 0.005, seed=21, maf_lo=0.05, maf_hi=0.5)
    g = O.setgeno(body, N, M, min_maf=0.01)
    grm = O.DenseGRM(O.grm_f64(g), g["Mqc"], O.diag_std(g))
    rng = np.random.default_rng(3)
    X = np.column_stack([np.ones(N), rng.normal(50, 10, N), rng.random(N) < 0.5])
    Xt, _, _ = plink.covariate_transform(X)
    gz = O.std_matrix(g).astype(np.float64)[:, :100] @ rng.normal(0, np.sqrt(0.4 / 100), 100)
    yb = (rng.random(N) < 1 / (1 + np.exp(-(-1.0 + gz)))).astype(float)
    yq = gz + rng.normal(0, np.sqrt(0.6), N)
    return grm, Xt, yb, yq


@pytest.mark.parametrize("trait", ["binary", "quantitative"])
def test_driver_matches_oracle(cohort, trait):
    grm, X, yb, yq = cohort
    y = yb if trait == "binary" else yq
    opt = D.Options(nrun=10, traceCVcutoff=1.0, tolPCG=1e-8, maxiterPCG=1000)
    probes = D.RademacherProbes(len(y), 200)
    res = DenseExecutor(grm, opt).run(D.fit_co(trait, y, X, probes, opt))
    kw = dict(nrun=10, tolPCG=1e-8, maxiterPCG=1000, cv_cutoff=1.0)
    if trait == "binary":
        ref = O.glmmkin_binary(grm, y, X, probes, **kw)
    else:
        ref = O.glmmkin_quantitative(grm, y, X, probes, **kw)
    assert np.allclose(res["theta"], ref["theta"], rtol=1e-9, atol=1e-12)
    assert np.allclose(res["coefficients"], ref["coefficients"], rtol=1e-8)
    assert res["iterations"] == ref["iterations"]
    assert len(res["history"]) == len(ref["history"])


def test_covariate_transform():
    rng = np.random.default_rng(1)
    N = 300
    X = np.column_stack([np.ones(N), rng.normal(50, 10, N), rng.random(N) < 0.5])
    Xt, R, keep = plink.covariate_transform(X)
    assert np.allclose(Xt.T @ Xt, N * np.eye(3))
    assert np.allclose(Xt / np.sqrt(N) @ R, X)
    Xc = np.column_stack([X, X[:, 1] * 2])         # collinear column dropped
    Xt2, _, keep2 = plink.covariate_transform(Xc)
    assert Xt2.shape[1] == 3 and list(keep2) == [0, 1, 2]


def test_pack_bed_round_trip():
    from saige_gpu.synth import decode_bed_rows
    rng = np.random.default_rng(2)
    G = rng.integers(0, 3, (23, 9))
    miss = rng.random((23, 9)) < 0.2
    body = plink.pack_bed(G, miss)
    G2, m2 = decode_bed_rows(body, 23)
    assert np.array_equal(m2, miss)
    assert np.array_equal(np.where(miss, 0, G2), np.where(miss, 0, G))


def test_variance_ratio_matches_oracle(cohort):
    """variance_ratio_co (FG.R:2152-2423 core) against the oracle's per-marker
    formula on the same markers, with a dense FP64 test executor."""
    grm, X, yb, yq = cohort
    opt = D.Options(nrun=10, traceCVcutoff=1.0, tolPCG=1e-10, maxiterPCG=2000)
    probes = D.RademacherProbes(len(yb), 200)
    ex = DenseExecutor(grm, opt)
    fit = ex.run(D.fit_co("binary", yb, X, probes, opt))
    rng = np.random.default_rng(5)
    Gpool = rng.binomial(2, rng.uniform(0.05, 0.5, 80)[None, :], (len(yb), 80))
    cats = [list(range(0, 40)), list(range(40, 80))]
    out = ex.run(D.variance_ratio_co(fit, yb, X, "binary", lambda i: Gpool[:, i], cats, opt,
                                     num_markers=12, ratio_cv_cutoff=1.0))
    mu = fit["fitted_values"]
    w = mu * (1 - mu)
    Sigma_iX, _ = grm.pcg(w, fit["theta"], X, 2000, 1e-10)
    noK = O.score_test_null(mu, mu * (1 - mu), yb, X)
    for k, cat in enumerate(cats):
        assert len(out["ratios"][k]) == 12
        ref = [O.variance_ratio_marker(grm, Gpool[:, i], X, yb, mu, w, fit["theta"], Sigma_iX, noK["XV"],
                                       noK["XXVX_inv"], "binary", 2000, 1e-10) for i in cat[:12]]
        assert np.allclose(out["ratios"][k], ref, rtol=1e-6)
        assert np.isclose(out["table"][k][0], np.mean(ref), rtol=1e-6)
    # CV growth: a cutoff below the 12-marker CV adds markers in steps of 10
    out2 = ex.run(D.variance_ratio_co(fit, yb, X, "binary", lambda i: Gpool[:, i], cats[:1], opt,
                                      num_markers=12, ratio_cv_cutoff=0.0))
    assert len(out2["ratios"][0]) == 40          # runs out of markers: 12, 22, 32, 40
    import tempfile, os
    with tempfile.TemporaryDirectory() as d:
        pth = os.path.join(d, "vr.txt")
        D.write_variance_ratio(pth, out["table"])
        lines = open(pth).read().split("\n")
        assert lines[0].split()[1:] == ["null", "1"]
        assert abs(float(lines[0].split()[0]) - out["table"][0][0]) <= 1e-14 * abs(out["table"][0][0])
