"""The C++ null-fit driver (saige-gpu_amd/csrc/fit_core.cpp: FG.R:2-35,
79-304, 309-549, 579-591, 1612-1647 with FG.cpp:3113-3160, 3167-3199,
3279-3341, 3409-3662) against the Python driver (saige_gpu.driver) on the
CPU.  Both answer every solve with an exact dense FP64 solve of
Sigma = tau0 diag(1/w) + tau1 K and every product with K b (the C++ side
through the test harness tests/native/fit_dense.cpp, compiled here with g++),
so the fits must agree to FP64 rounding: same iteration counts, same tau
path, same coefficients.  The Python driver is itself pinned to the oracle's
restatement of FG.R (test_host_driver.py)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from fixtures import random_bed
from oracle import oracle as O
from saige_gpu import driver as D
from saige_gpu import plink

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "saige-gpu_amd", "csrc")
P = ctypes.POINTER


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("fitdense")
    out, vm = str(tmp / "libfitdense.so"), str(tmp / "vmath.o")
    # the array exp/log file is built as the library's Makefile builds it
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffast-math", "-fopenmp-simd", "-c",
                    os.path.join(CSRC, "vmath.cpp"), "-o", vm], check=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", CSRC,
                    os.path.join(HERE, "native", "fit_dense.cpp"), os.path.join(CSRC, "fit_core.cpp"), vm,
                    "-lmvec", "-lm", "-o", out], check=True)
    L = ctypes.CDLL(out)
    d, i64, u64 = P(ctypes.c_double), ctypes.c_int64, ctypes.c_uint64
    L.fitnull_dense.argtypes = [ctypes.c_int, i64, ctypes.c_int, d, d, d, ctypes.c_int, ctypes.c_double,
                                ctypes.c_int, ctypes.c_double, u64, d, d, d, d, d, d, P(ctypes.c_int),
                                P(ctypes.c_int), d, P(ctypes.c_int)]
    L.probes.argtypes = [u64, i64, i64, d]
    L.calcv.argtypes = [d, i64]
    L.calcv.restype = ctypes.c_double
    L.covtrans.argtypes = [d, i64, ctypes.c_int, d, d, P(ctypes.c_int), P(ctypes.c_int)]
    L.scorenull.argtypes = [d, d, d, d, i64, ctypes.c_int, d, d, d, d]
    return L


def ptr(a):
    return a.ctypes.data_as(P(ctypes.c_double))


class ExactExecutor:
    """Answers the Python coroutines' requests with exact dense FP64 algebra."""

    def __init__(self, K):
        self.K = K

    def solve(self, w, tau, B):
        Sig = tau[0] * np.diag(1.0 / np.asarray(w)) + tau[1] * self.K
        return np.linalg.solve(Sig, B)

    def run(self, gen):
        value = None
        while True:
            try:
                req = gen.send(value)
            except StopIteration as e:
                return e.value
            if req[0] == "solve":
                value = self.solve(*req[1:])
            elif req[0] == "matvec":
                value = self.K @ req[1]
            else:
                value = D.trace_round_host(req, lambda B: self.K @ B, self.solve)


@pytest.fixture(scope="module")
def cohort():
    N, M = 160, 500
    body, G, miss = random_bed(N, M, 0.005, seed=31, maf_lo=0.05, maf_hi=0.5)
    g = O.setgeno(body, N, M, min_maf=0.01)
    K = O.grm_f64(g) / g["Mqc"]
    rng = np.random.default_rng(7)
    X = np.column_stack([np.ones(N), rng.normal(50, 10, N), rng.random(N) < 0.5])
    Xt, _, _ = plink.covariate_transform(X)
    gz = O.std_matrix(g).astype(np.float64)[:, :100] @ rng.normal(0, np.sqrt(0.5 / 100), 100)
    yb = (rng.random(N) < 1 / (1 + np.exp(-(-0.8 + gz)))).astype(float)
    yq = gz + rng.normal(0, np.sqrt(0.5), N)
    return K, np.asfortranarray(Xt), yb, yq


def run_cpp(lib, trait, K, y, X, opt):
    n, p = X.shape
    theta, coef = np.zeros(2), np.zeros(p)
    eta, mu, Yw, cov = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros((p, p), order="F")
    iters, conv, nh = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    hist = np.zeros(2 * (opt.maxiter + 2))
    Kf = np.asfortranarray(K)
    rc = lib.fitnull_dense(0 if trait == "binary" else 1, n, p, ptr(Kf), ptr(np.ascontiguousarray(y)), ptr(X),
                           opt.maxiter, opt.tol, opt.nrun, opt.traceCVcutoff, 200, ptr(theta), ptr(coef), ptr(eta),
                           ptr(mu), ptr(Yw), ptr(cov), ctypes.byref(iters), ctypes.byref(conv), ptr(hist),
                           ctypes.byref(nh))
    assert rc == 0
    return {"theta": theta, "coefficients": coef, "linear_predictors": eta, "fitted_values": mu, "Y": Yw, "cov": cov,
            "iterations": iters.value, "converged": bool(conv.value), "history": hist[:2 * nh.value].reshape(-1, 2)}


@pytest.mark.parametrize("trait,cv_cutoff", [("binary", 1.0), ("quantitative", 1.0), ("binary", 0.0025),
                                              ("quantitative", 0.0025)])
def test_cpp_driver_matches_python_driver(lib, cohort, trait, cv_cutoff):
    K, X, yb, yq = cohort
    y = yb if trait == "binary" else yq
    opt = D.Options(nrun=10, traceCVcutoff=cv_cutoff)
    probes = D.RademacherProbes(len(y), 200)
    ref = ExactExecutor(K).run(D.fit_co(trait, y, X, probes, opt))
    got = run_cpp(lib, trait, K, y, X, opt)
    assert got["iterations"] == ref["iterations"] and got["converged"] == ref["converged"]
    assert np.allclose(got["history"], np.array(ref["history"]), rtol=1e-9, atol=1e-13)
    assert np.allclose(got["theta"], ref["theta"], rtol=1e-9, atol=1e-13)
    assert ref["theta"][1] > 0                      # a fit that moved off the start values
    for k in ("coefficients", "linear_predictors", "fitted_values", "Y", "cov"):
        assert np.allclose(got[k], ref[k], rtol=1e-8, atol=1e-11), k


def test_cpp_probes_and_cv(lib):
    for i in (0, 7, 123):
        a = np.zeros(257)
        lib.probes(200, i, 257, ptr(a))
        assert np.array_equal(a, D.probe_vector(200, i, 257))
    x = np.random.default_rng(2).normal(3.0, 0.5, 37)
    assert np.isclose(lib.calcv(ptr(x), len(x)), D.cal_cv(x), rtol=1e-13)


def test_cpp_covariate_transform_and_score_null(lib):
    rng = np.random.default_rng(4)
    N = 211
    X = np.column_stack([np.ones(N), rng.normal(50, 10, N), rng.random(N) < 0.5])
    for Xin in (X, np.column_stack([X, 2 * X[:, 1] - X[:, 0]])):       # the second has a collinear column
        Xin = np.asfortranarray(Xin)
        p = Xin.shape[1]
        Xo, R = np.zeros(N * p), np.zeros(p * p)          # flat: the C++ side writes N x k, k x k
        kept, pk = np.zeros(p, np.int32), ctypes.c_int(0)
        lib.covtrans(ptr(Xin), N, p, ptr(Xo), ptr(R), kept.ctypes.data_as(P(ctypes.c_int)), ctypes.byref(pk))
        Xt, Rt, keep = plink.covariate_transform(Xin)
        k = pk.value
        assert list(kept[:k]) == list(keep)
        Xo = Xo[:N * k].reshape((N, k), order="F")
        Ro = R[:k * k].reshape((k, k), order="F")
        assert np.allclose(Xo, Xt, rtol=1e-10, atol=1e-10)
        assert np.allclose(Ro, Rt, rtol=1e-10, atol=1e-10)
    mu = rng.uniform(0.1, 0.9, N)
    y = (rng.random(N) < mu).astype(float)
    Xt = np.asfortranarray(Xt)
    p = Xt.shape[1]
    XVX, XVXi, S_a, XVXiXV = (np.zeros((p, p), order="F"), np.zeros((p, p), order="F"), np.zeros(p),
                              np.zeros((N, p), order="F"))
    V = mu * (1 - mu)
    lib.scorenull(ptr(mu), ptr(V), ptr(y), ptr(Xt), N, p, ptr(XVX), ptr(XVXi), ptr(S_a), ptr(XVXiXV))
    ref = D.score_test_null(mu, V, y, Xt)
    for got, k in ((XVX, "XVX"), (XVXi, "XVX_inv"), (S_a, "S_a"), (XVXiXV, "XVX_inv_XV")):
        assert np.allclose(got, ref[k], rtol=1e-10, atol=1e-12), k


def test_library_covariate_transform_export():
    """The same code exported by libsaigehip (saige_covariate_transform),
    called without a device."""
    from saige_gpu.api import covariate_transform_native
    rng = np.random.default_rng(8)
    N = 97
    X = np.column_stack([np.ones(N), rng.normal(50, 10, N), rng.random(N) < 0.5, np.ones(N) * 3])
    Xt, R, keep = covariate_transform_native(X)
    Xr, Rr, keepr = plink.covariate_transform(X)
    assert list(keep) == list(keepr) == [0, 1, 2]
    assert np.allclose(Xt, Xr, rtol=1e-10, atol=1e-10) and np.allclose(R, Rr, rtol=1e-10, atol=1e-10)


def test_library_glm_start_export():
    """saige_glm_start (the C++ glm.fit restatement the batched path takes
    its start values from), called without a device, against the Python
    glm_start for both families, with and without an offset.  The start
    values also satisfy glm's score equations X^T (y - mu) = 0 (logit, no
    offset), which a wrong weight or working response would break."""
    from saige_gpu.api import glm_start_native
    rng = np.random.default_rng(9)
    N = 3000
    X, _, _ = plink.covariate_transform(np.column_stack([np.ones(N), rng.normal(50, 10, N), rng.random(N) < 0.5]))
    lin = X @ np.array([-0.03, 0.4, -0.2])
    yb = (rng.random(N) < 1 / (1 + np.exp(-lin))).astype(np.float64)
    yq = lin + rng.standard_normal(N)
    off = 0.1 * rng.standard_normal(N)
    for trait, fam, y in (("binary", D.Binomial, yb), ("quantitative", D.Gaussian, yq)):
        for o in (None, off):
            got = glm_start_native(trait, y, X, o)
            ref = D.glm_start(y, X, fam, o)
            for k in ("coef", "eta", "mu"):
                assert np.allclose(got[k], ref[k], rtol=1e-10, atol=1e-12), (trait, o is None, k)
    g = glm_start_native("binary", yb, X)
    assert np.max(np.abs(X.T @ (yb - g["mu"]))) < 1e-6 * N


def test_library_get_coef_helpers():
    """saige_glm_working and saige_coefficients (Get_Coef's working response
    and getCoefficients' algebra after the solve, which the GPU paths take
    from the library) against the numpy forms of driver.get_coef_co, for
    both families, with and without an offset.  The coefficients are also
    pinned by a closed form: with Sigma = diag(1/w) (tau = (1, 0)), sol is
    W [Y, X] and alpha is the weighted least-squares fit of Y on X."""
    from saige_gpu.api import coefficients_native, glm_working_native
    rng = np.random.default_rng(12)
    N = 2500
    X, _, _ = plink.covariate_transform(np.column_stack([np.ones(N), rng.normal(50, 10, N), rng.random(N) < 0.5]))
    Xf = np.asfortranarray(X)
    eta = X @ np.array([-0.05, 0.3, 0.1]) + 0.2 * rng.standard_normal(N)
    yb = (rng.random(N) < 1 / (1 + np.exp(-eta))).astype(np.float64)
    yq = eta + rng.standard_normal(N)
    off = 0.1 * rng.standard_normal(N)
    for trait, fam, y in (("binary", D.Binomial, yb), ("quantitative", D.Gaussian, yq)):
        for o in (None, off):
            oo = np.zeros(N) if o is None else o
            mu, Y, W = glm_working_native(trait, y, eta, o)
            mr = fam.linkinv(eta)
            me = fam.mu_eta(eta)
            assert np.allclose(mu, mr, rtol=1e-14, atol=0), (trait, o is None)
            assert np.allclose(Y, eta - oo + (y - mr) / me, rtol=1e-13, atol=1e-14), (trait, o is None)
            assert np.allclose(W, (me / np.sqrt(fam.variance(mr))) ** 2, rtol=1e-13, atol=0), (trait, o is None)
    w = 0.2 + rng.random(N)
    Y = rng.standard_normal(N)
    sol = np.asfortranarray(np.column_stack([Y * w, X * w[:, None]]))
    cov, alpha, e = coefficients_native(Y, Xf, w, 1.0, sol)
    XtWX = X.T @ (X * w[:, None])
    assert np.allclose(cov, np.linalg.inv(XtWX), rtol=1e-10, atol=1e-14)
    assert np.allclose(alpha, np.linalg.solve(XtWX, X.T @ (w * Y)), rtol=1e-10, atol=1e-12)
    assert np.allclose(e, X @ alpha, rtol=1e-9, atol=1e-11)      # Y - (WY - WX alpha) / w
    # a general solution block against the numpy algebra of get_coefficients_co
    sol = np.asfortranarray(rng.standard_normal((N, 4)))
    cov, alpha, e = coefficients_native(Y, Xf, w, 0.7, sol)
    siY, siX = sol[:, 0], sol[:, 1:]
    cr = D._inv_sympd(X.T @ siX)
    ar = cr @ (siX.T @ Y)
    assert np.allclose(cov, cr, rtol=1e-9, atol=1e-13)
    assert np.allclose(alpha, ar, rtol=1e-9, atol=1e-13)
    assert np.allclose(e, Y - 0.7 * (siY - siX @ ar) / w, rtol=1e-9, atol=1e-12)


def test_logit_clamps_follow_r_and_bad_inputs_are_refused():
    """R's binomial() beyond |eta| > 30 (logit_linkinv / logit_mu_eta take
    exp(eta) as DBL_EPSILON or 1/DBL_EPSILON, mu_eta = DBL_EPSILON): the C++
    family, the Python driver and the oracle give these values, so the
    working response of a quasi-separated phenotype stays finite; inputs R's
    glm() refuses (binomial y outside [0, 1], non-finite values) are refused
    with SAIGE_E_ARG."""
    from saige_gpu.api import glm_start_native, glm_working_native
    from saige_gpu._lib import SaigeError
    eps = np.finfo(np.float64).eps
    eta = np.array([-800.0, -31.0, -30.5, -30.0, -1.0, 0.0, 1.0, 30.0, 30.5, 31.0, 800.0])
    y = np.array([0, 0, 1, 0, 1, 0, 1, 1, 0, 1, 1], np.float64)
    big = np.abs(eta) > 30
    # R's values beyond the threshold, typed from its definition
    mu_r = np.where(eta > 30, (1 / eps) / (1 + 1 / eps), eps / (1 + eps))
    for F in (D.Binomial, O.Binomial):
        assert np.array_equal(F.mu_eta(eta)[big], np.full(big.sum(), eps))
        assert np.array_equal(F.linkinv(eta)[big], mu_r[big])
        # inside, the textbook logistic (to rounding)
        assert np.allclose(F.linkinv(eta[~big]), 1 / (1 + np.exp(-eta[~big])), rtol=1e-15)
    mu, Y, W = glm_working_native("binary", y, eta)
    assert np.array_equal(mu[big], mu_r[big])
    assert np.all(np.isfinite(Y)) and np.all(np.isfinite(W)) and np.all(W > 0)
    assert np.allclose(Y, eta + (y - mu) / D.Binomial.mu_eta(eta), rtol=1e-12)
    # a quasi-separated covariate: the start values stay finite
    rng = np.random.default_rng(3)
    n = 400
    x = rng.standard_normal(n)
    ys = (x > 0).astype(float)
    ys[:3] = 1 - ys[:3]
    X = np.column_stack([np.ones(n), x])
    g = glm_start_native("binary", ys, X)
    assert np.all(np.isfinite(g["coef"])) and np.all(np.isfinite(g["eta"]))
    assert np.all((g["mu"] > 0) & (g["mu"] < 1))
    ref = D.glm_start(ys, X, D.Binomial, np.zeros(n))
    assert np.allclose(g["coef"], ref["coef"], rtol=1e-9)
    for bad_y in (np.where(np.arange(n) == 5, 2.0, ys), np.where(np.arange(n) == 7, np.nan, ys)):
        with pytest.raises(SaigeError):
            glm_start_native("binary", bad_y, X)
    Xbad = X.copy()
    Xbad[9, 1] = np.inf
    with pytest.raises(SaigeError):
        glm_start_native("quantitative", ys, Xbad)
    with pytest.raises(SaigeError):
        glm_working_native("binary", y, np.where(eta == 0, np.nan, eta))
