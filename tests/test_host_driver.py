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
                _, w, tau, B = req
                value, _ = self.grm.pcg(np.asarray(w), np.asarray(tau, np.float64), B, self.opt.maxiterPCG,
                                        self.opt.tolPCG)
            else:
                value = self.grm.K @ req[1]


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
