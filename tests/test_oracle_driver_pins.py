"""Pins the oracle's driver layer -- Get_Coef (FG.R:2-35), the binary and
quantitative AI-REML drivers (FG.R:79-304, 309-549), the tau-update rules of
fitglmmaiRPCG[_q] (FG.cpp:3302-3341, 3610-3662), the variance-ratio core
(FG.R:2309-2368) and one LOCO chromosome (FG.R:42-73, 1240-1272) -- with
checks the mathematics fixes, not with a second restatement of FG.R:

* one Get_Coef step at tau = (1, 0) is the Newton step of the GLM score
  equations, and its fixed point is glm()'s maximum-likelihood estimate;
* at tau1 > 0, Get_Coef's fixed point solves the penalised quasi-likelihood
  equations X'(y - mu) = 0 and eta - X alpha - offset = (tau1/tau0) K (y - mu)
  (Henderson's mixed-model equations for the working response);
* with probes sqrt(N) e_i, i = 1..N, the Hutchinson trace is exact, so the
  quantitative AI-REML fixed point is the REML optimum (found here by scipy
  on the dense REML log-likelihood) and the binary one zeroes the dense
  score Y'PKPY - tr(PK);
* at tau1 = 0, G adjusted for X is W-orthogonal to X, so the variance ratio
  is exactly 1 (binary) or 1/tau0 (quantitative);
* the update rules are checked on hand-computed cases that reach every
  branch: acceptance, zeroing below tol, the binary rule's zeroing of a
  negative step (the reference zeroes before its halving loop, so a binary
  step is never halved), and the quantitative step halving with GMMAT's
  zero mask.
"""
import numpy as np
import pytest
from scipy.optimize import minimize

from fixtures import random_bed
from oracle import oracle as O


def _problem(N, M, seed):
    body, _, _ = random_bed(N, M, 0.0, seed=seed, maf_lo=0.05, maf_hi=0.5)
    g = O.setgeno(body, N, M, min_maf=0.01)
    return g, O.DenseGRM(O.grm_f64(g), g["Mqc"], O.diag_std(g))


def _polygenic(grm, rng, var):
    L = np.linalg.cholesky(grm.K + 1e-9 * np.eye(grm.N))
    return L @ rng.standard_normal(grm.N) * np.sqrt(var)


@pytest.fixture(scope="module")
def binary_problem():
    N = 160
    g, grm = _problem(N, 400, 51)
    rng = np.random.default_rng(51)
    X = np.column_stack([np.ones(N), rng.standard_normal(N)])
    eta = X @ np.array([-0.3, 0.5]) + _polygenic(grm, rng, 3.0)
    y = (rng.random(N) < 1 / (1 + np.exp(-eta))).astype(float)
    return grm, X, y


@pytest.fixture(scope="module")
def quant_problem():
    N = 120
    g, grm = _problem(N, 500, 41)
    rng = np.random.default_rng(42)
    X = np.column_stack([np.ones(N), rng.standard_normal(N)])
    y = X @ np.array([0.5, 0.3]) + _polygenic(grm, rng, 0.6) + rng.standard_normal(N) * np.sqrt(0.4)
    return grm, X, y


def _exact_probes(N):
    E = np.sqrt(N) * np.eye(N)
    return lambda i: E[:, i]


# ------------------------------------------------------------ Get_Coef

def test_get_coef_step_is_newton_at_tau1_zero(binary_problem):
    """Sigma = W^-1: one Get_Coef iteration is alpha0 + (X'WX)^-1 X'(y - mu0),
    the Newton step of the logistic score equations (canonical link); the
    offset enters eta, not the working response's regression."""
    grm, X, y = binary_problem
    N = len(y)
    off = np.linspace(-0.2, 0.2, N)
    alpha0 = np.array([0.4, -0.6])
    eta0 = X @ alpha0 + off
    rc = O.get_coef(grm, y, X, np.array([1.0, 0.0]), O.Binomial, alpha0, eta0, off,
                    maxiterPCG=50, tolPCG=1e-20, maxiter=1)
    mu0 = 1 / (1 + np.exp(-eta0))
    W0 = mu0 * (1 - mu0)
    newton = alpha0 + np.linalg.solve(X.T @ (W0[:, None] * X), X.T @ (y - mu0))
    assert np.allclose(rc["alpha"], newton, rtol=1e-10, atol=1e-12)
    # eta = X alpha + offset at tau1 = 0 (no random effect)
    assert np.allclose(rc["eta"], X @ rc["alpha"] + off, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("fam", ["binomial", "gaussian"])
def test_get_coef_fixed_point_is_glm_mle(binary_problem, quant_problem, fam):
    grm, X, y = binary_problem if fam == "binomial" else quant_problem
    F = O.Binomial if fam == "binomial" else O.Gaussian
    N = len(y)
    tau = np.array([1.0, 0.0])
    alpha, eta = np.zeros(X.shape[1]), np.zeros(N)
    for _ in range(30):
        rc = O.get_coef(grm, y, X, tau, F, alpha, eta, np.zeros(N), maxiterPCG=50, tolPCG=1e-20)
        done = np.max(np.abs(rc["alpha"] - alpha)) < 1e-13
        alpha, eta = rc["alpha"], rc["eta"]
        if done:
            break
    mle = O.glm_fit(y, X, F)["coef"]
    assert np.allclose(alpha, mle, rtol=1e-8, atol=1e-10)
    # the MLE itself: X'(y - mu) = 0
    assert np.max(np.abs(X.T @ (y - F.linkinv(X @ alpha)))) < 1e-8 * N


@pytest.mark.parametrize("fam", ["binomial", "gaussian"])
def test_get_coef_fixed_point_solves_pql(binary_problem, quant_problem, fam):
    grm, X, y = binary_problem if fam == "binomial" else quant_problem
    F = O.Binomial if fam == "binomial" else O.Gaussian
    tau = np.array([1.0, 0.7]) if fam == "binomial" else np.array([0.6, 0.4])
    N = len(y)
    off = np.full(N, 0.1)
    st = O.glm_fit(y, X, F, off)
    alpha, eta = st["coef"], st["eta"]
    for _ in range(60):
        rc = O.get_coef(grm, y, X, tau, F, alpha, eta, off, maxiterPCG=5000, tolPCG=1e-22)
        done = np.max(np.abs(rc["alpha"] - alpha)) < 1e-12
        alpha, eta = rc["alpha"], rc["eta"]
        if done:
            break
    mu = F.linkinv(eta)
    assert np.max(np.abs(X.T @ (y - mu))) < 1e-8
    b = eta - X @ alpha - off
    blup = tau[1] / tau[0] * grm.K @ (y - mu)
    assert np.max(np.abs(b - blup)) < 1e-8 * np.max(np.abs(b))
    # a random effect of the right size: not the tau1 = 0 answer
    assert np.max(np.abs(b)) > 0.05


# ------------------------------------------------------------ AI-REML with exact traces

def test_quantitative_aireml_reaches_reml_optimum(quant_problem):
    grm, X, y = quant_problem
    N = len(y)
    fit = O.glmmkin_quantitative(grm, y, X, _exact_probes(N), nrun=N, tol=1e-9, maxiter=100,
                                 tolPCG=1e-16, maxiterPCG=5000, cv_cutoff=1e9)
    assert fit["converged"]

    def neg_reml(lt):
        t0, t1 = np.exp(lt)
        S = t0 * np.eye(N) + t1 * grm.K
        Si = np.linalg.inv(S)
        XSX = X.T @ Si @ X
        P = Si - Si @ X @ np.linalg.solve(XSX, X.T @ Si)
        return 0.5 * (np.linalg.slogdet(S)[1] + np.linalg.slogdet(XSX)[1] + y @ P @ y)

    opt = minimize(neg_reml, np.log([0.5, 0.5]), method="Nelder-Mead",
                   options=dict(xatol=1e-12, fatol=1e-15, maxiter=20000))
    assert np.allclose(fit["theta"], np.exp(opt.x), rtol=1e-5)
    # the REML score components vanish at the driver's answer
    t = fit["theta"]
    Si = np.linalg.inv(t[0] * np.eye(N) + t[1] * grm.K)
    P = Si - Si @ X @ np.linalg.solve(X.T @ Si @ X, X.T @ Si)
    Py = P @ y
    s0 = Py @ Py - np.trace(P)
    s1 = Py @ grm.K @ Py - np.trace(P @ grm.K)
    assert abs(s0) < 1e-6 * np.trace(P) and abs(s1) < 1e-6 * np.trace(P @ grm.K)
    # both components interior
    assert t[0] > 0.02 and t[1] > 0.02


def test_binary_aireml_zeroes_dense_score(binary_problem):
    grm, X, y = binary_problem
    N = len(y)
    fit = O.glmmkin_binary(grm, y, X, _exact_probes(N), nrun=N, tol=1e-9, maxiter=100,
                           tolPCG=1e-16, maxiterPCG=5000, cv_cutoff=1e9)
    assert fit["converged"]
    tau = fit["theta"]
    assert tau[0] == 1.0 and tau[1] > 0.05
    mu = fit["fitted_values"]
    W = mu * (1 - mu)
    Si = np.linalg.inv(tau[0] * np.diag(1 / W) + tau[1] * grm.K)
    P = Si - Si @ X @ np.linalg.solve(X.T @ Si @ X, X.T @ Si)
    # the working response at the fit (FG.R:4-8)
    eta = fit["linear_predictors"]
    Y = eta + (y - mu) / W
    assert np.allclose(fit["Y"], Y, rtol=1e-8, atol=1e-10)
    PY = P @ Y
    score = PY @ grm.K @ PY - np.trace(P @ grm.K)
    AI = PY @ grm.K @ P @ grm.K @ PY
    assert abs(score / AI) < 1e-8 * tau[1]
    # and the fixed effects / random effect solve the PQL equations there
    assert np.max(np.abs(X.T @ (y - mu))) < 1e-7
    b = eta - X @ fit["coefficients"]
    assert np.max(np.abs(b - tau[1] * grm.K @ (y - mu))) < 1e-7 * np.max(np.abs(b))


# ------------------------------------------------------------ tau update rules

def _fake_scores(monkeypatch, name, values):
    monkeypatch.setattr(O, name, lambda *a, **k: values)


@pytest.mark.parametrize("tau1,dtau,expect", [
    (0.3, 0.2, 0.5),     # accepted
    (0.3, -0.27, 0.03),  # accepted, above tol
    (0.3, -0.29, 0.0),   # 0.01 < tol: zeroed
    (0.3, -0.5, 0.0),    # negative: zeroed before the halving loop, never halved to 0.05
    (0.0, 0.01, 0.0),    # from zero, a step below tol stays at zero
])
def test_binary_tau_update_rule(monkeypatch, tau1, dtau, expect):
    # score1 = YPAPY - Trace = dtau * AI with AI = 2
    _fake_scores(monkeypatch, "get_ai_score", {"YPAPY": 5.0 + 2.0 * dtau, "Trace": 5.0, "AI": 2.0})
    tau, _ = O.fitglmmai_rpcg(None, None, None, None, np.array([1.0, tau1]), None, None, None, 30, None,
                              tol=0.02)
    assert tau[0] == 1.0
    assert tau[1] == pytest.approx(expect, abs=1e-15)


@pytest.mark.parametrize("tau,dtau,expect", [
    ((1.0, 0.5), (-0.2, 0.1), (0.8, 0.6)),        # accepted
    ((1.0, 0.5), (-0.2, -0.8), (0.9, 0.1)),       # tau1 < 0: halved once
    ((0.5, 0.3), (-1.5, 0.2), (0.125, 0.35)),     # tau0 < 0: halved twice
    ((1.0, 0.0), (0.1, -0.5), (1.1, 0.0)),        # tau1 started below tol: masked to 0, no halving
    ((1.0, 0.05), (0.0, -0.04), (1.0, 0.0)),      # lands in (0, tol): zeroed at the end
    ((1.0, 0.0), (0.0, 0.3), (1.0, 0.3)),         # leaves zero when the step is large enough
])
def test_quantitative_tau_update_rule(monkeypatch, tau, dtau, expect):
    AI = np.array([[2.0, 0.5], [0.5, 1.0]])
    score = AI @ np.array(dtau)
    _fake_scores(monkeypatch, "get_ai_score_q", {"YPAPY": 7.0 + score[1], "YPA0PY": 3.0 + score[0],
                                                 "Trace": np.array([3.0, 7.0]), "AI": AI})
    new, _ = O.fitglmmai_rpcg_q(None, None, None, None, np.array(tau), None, None, None, 30, None,
                                tol=0.02)
    assert np.allclose(new, expect, rtol=0, atol=1e-14)


# ------------------------------------------------------------ variance ratio

@pytest.mark.parametrize("trait", ["binary", "quantitative"])
def test_variance_ratio_at_tau1_zero(binary_problem, trait):
    grm, X, yb = binary_problem
    N = len(yb)
    rng = np.random.default_rng(7)
    if trait == "binary":
        y = yb
        mu = O.glm_fit(y, X, O.Binomial)["mu"]
        w = mu * (1 - mu)
        tau = np.array([1.0, 0.0])
        mu2 = w
        expect = 1.0
    else:
        y = X @ np.array([1.0, 0.5]) + rng.standard_normal(N)
        mu = X @ np.linalg.lstsq(X, y, rcond=None)[0]
        w = np.ones(N)
        tau = np.array([0.7, 0.0])
        mu2 = np.full(N, 1 / tau[0])
        expect = 1 / tau[0]
    Sigma_iX, _ = grm.pcg(w, tau, X, 100, 1e-24)
    noK = O.score_test_null(mu, mu2, y, X)
    for j in range(6):
        G0 = rng.binomial(2, 0.1 + 0.07 * j, N).astype(float)
        r = O.variance_ratio_marker(grm, G0, X, y, mu, w, tau, Sigma_iX, noK["XV"], noK["XXVX_inv"],
                                    trait, maxiterPCG=100, tolPCG=1e-24)
        assert r == pytest.approx(expect, rel=1e-10)


def test_variance_ratio_below_one_with_polygenic_component(binary_problem):
    """At tau1 > 0 the ratio is G'PG / G'WG with P at (W, tau): below 1,
    because Sigma = W^-1 + tau1 K >= W^-1 makes P <= Sigma^-1 <= W (a dropped
    tau1 term would give exactly 1)."""
    grm, X, y = binary_problem
    N = len(y)
    mu = O.glm_fit(y, X, O.Binomial)["mu"]
    w = mu * (1 - mu)
    tau = np.array([1.0, 0.5])
    Sigma_iX, _ = grm.pcg(w, tau, X, 5000, 1e-24)
    noK = O.score_test_null(mu, w, y, X)
    rng = np.random.default_rng(8)
    G0 = rng.binomial(2, 0.3, N).astype(float)
    r = O.variance_ratio_marker(grm, G0, X, y, mu, w, tau, Sigma_iX, noK["XV"], noK["XXVX_inv"],
                                "binary", maxiterPCG=5000, tolPCG=1e-24)
    # dense: var1 = G'PG / AC with P at (W, tau); var2 = G'WG / AC
    if G0.sum() / (2 * N) > 0.5:
        G0 = 2 - G0
    G = G0 - noK["XXVX_inv"] @ (noK["XV"] @ G0)
    Si = np.linalg.inv(grm.sigma(w, tau))
    P = Si - Si @ X @ np.linalg.solve(X.T @ Si @ X, X.T @ Si)
    assert r == pytest.approx((G @ P @ G) / (G @ (w * G)), rel=1e-9)
    assert r < 1.0   # P <= W: the GRM term can only reduce G's precision-weighted variance


# ------------------------------------------------------------ LOCO

@pytest.mark.parametrize("trait", ["binary", "quantitative"])
def test_loco_chromosome_is_get_coef_on_bruteforce_grm(trait):
    N, M = 140, 420
    body, G, _ = random_bed(N, M, 0.0, seed=61, maf_lo=0.05, maf_hi=0.5)
    g = O.setgeno(body, N, M, min_maf=0.01)
    start, end = 100, 249
    # brute force: standardise the retained markers' integer genotypes in FP64
    keep = np.r_[np.arange(0, start), np.arange(end + 1, g["Mqc"])]
    qc_idx = np.nonzero(g["qc_mask"])[0][keep]
    Gk = G[:, qc_idx].astype(np.float64)
    p = Gk.mean(0) / 2
    Zk = (Gk - 2 * p) / np.sqrt(2 * p * (1 - p))
    Kb = O.DenseGRM(Zk @ Zk.T, Zk.shape[1])
    rng = np.random.default_rng(62)
    X = np.column_stack([np.ones(N), rng.standard_normal(N)])
    if trait == "binary":
        y = (rng.random(N) < 0.35).astype(float)
        F, theta, mu2 = O.Binomial, np.array([1.0, 0.4]), None
    else:
        y = rng.standard_normal(N) + 0.5 * X[:, 1]
        F, theta, mu2 = O.Gaussian, np.array([0.6, 0.3]), None
    st = O.glm_fit(y, X, F)
    fit = {"theta": theta, "coefficients": st["coef"], "linear_predictors": st["eta"]}
    res = O.loco_chromosome(O.loco_grm(g, start, end), fit, y, X, trait, maxiterPCG=5000, tolPCG=1e-20)
    ref = O.get_coef(Kb, y, X, theta, F, st["coef"], st["eta"], np.zeros(N), maxiterPCG=5000,
                     tolPCG=1e-20)
    # FP32 standardisation in the oracle's GRM against FP64 here
    for a, b in (("coefficients", "alpha"), ("linear_predictors", "eta"), ("Y", "Y"), ("cov", "cov")):
        assert np.allclose(res[a], ref[b], rtol=2e-5, atol=2e-6), a
    mu = ref["mu"]
    mu2 = mu * (1 - mu) if trait == "binary" else np.full(N, 1 / theta[0])
    XVX = X.T @ (mu2[:, None] * X)
    assert np.allclose(res["obj_noK"]["XVX"], XVX, rtol=2e-5)
    assert np.allclose(res["obj_noK"]["XXVX_inv"], X @ np.linalg.inv(XVX), rtol=2e-5, atol=1e-7)
    assert np.allclose(res["obj_noK"]["S_a"], X.T @ (y - mu), atol=2e-5)
