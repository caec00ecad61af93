"""Null-model fitting drivers: the host counterpart of the reference's R
drivers (src/SAIGE/R/SAIGE_fitGLMM_fast.R, FG.R below) and of the
solver-level Rcpp functions (src/SAIGE/src/SAIGE_fitGLMM_fast.cpp, FG.cpp).

Each fit is written once, as a coroutine that yields work requests:
    ("solve",  w, tau, B)  ->  X = Sigma^-1 B     (B: n x k, float64)
    ("matvec", B)          ->  K B
and receives the results.  Two executors run the coroutines:

* PlainExecutor: every solve column goes through saige_pcg1 (single RHS,
  GEMV on the resident GRM), one after another -- the reference's call
  pattern (getPCG1ofSigmaAndVector per vector), i.e. the drop-in path.
* BatchScheduler: many fits share the block-PCG slot engine (64 columns,
  FP32 MFMA GEMM per iteration); columns join and leave between iterations
  (continuous batching), so requests of different fits, sample subsets and
  phases advance together.

Host arithmetic (p x p inverses, projections, dot products of trace probes,
the tau update) is FP64.  All O(N^2) work -- every GRM product inside the
solves and every K u -- runs on the device.
"""
import collections
import concurrent.futures
import ctypes
import heapq
import time

import numpy as np

# ----------------------------------------------------------------- families


_EPS = np.finfo(np.float64).eps


class Binomial:
    """R's binomial(link = "logit") as used by FG.R, with its clamps beyond
    |eta| > 30 (R's logit_linkinv / logit_mu_eta take exp(eta) as
    1/DBL_EPSILON or DBL_EPSILON there; mu_eta is DBL_EPSILON)."""
    name = "binomial"

    @staticmethod
    def linkinv(eta):
        eta = np.asarray(eta, np.float64)
        with np.errstate(over="ignore"):
            mu = 1.0 / (1.0 + np.exp(-eta))
        return np.where(eta > 30.0, 1.0 / (1.0 + _EPS), np.where(eta < -30.0, _EPS / (1.0 + _EPS), mu))

    @staticmethod
    def mu_eta(eta):
        eta = np.asarray(eta, np.float64)
        e = np.exp(-np.abs(eta))
        return np.where(np.abs(eta) > 30.0, _EPS, e / (1.0 + e) ** 2)

    @staticmethod
    def variance(mu):
        return mu * (1.0 - mu)


class Gaussian:
    """R's gaussian(link = "identity")."""
    name = "gaussian"

    @staticmethod
    def linkinv(eta):
        return eta

    @staticmethod
    def mu_eta(eta):
        return np.ones_like(eta)

    @staticmethod
    def variance(mu):
        return np.ones_like(mu)


def glm_start(y, X, family, offset=None, maxit=25, epsilon=1e-8):
    """Start values of fitNULLGLMM's fit0 = glm(...) (FG.R:1119, 1370): R's
    glm.fit IRLS with its deviance stopping rule (glm.control defaults)."""
    n = len(y)
    off = np.zeros(n) if offset is None else np.asarray(offset, np.float64)
    y = np.asarray(y, np.float64)
    if family.name == "binomial":
        mu = (y + 0.5) / 2.0
        eta = np.log(mu / (1.0 - mu))
    else:
        mu = y.copy()
        eta = mu.copy()

    binary01 = family.name == "binomial" and bool(np.all((y == 0) | (y == 1)))

    def deviance(mu):
        if binary01:
            # y in {0, 1}: the unit deviance is -2 log(mu) or -2 log(1 - mu)
            return -2.0 * float(np.sum(np.log(np.where(y > 0, mu, 1.0 - mu))))
        if family.name == "binomial":
            with np.errstate(divide="ignore", invalid="ignore"):
                a = np.where(y > 0, y * np.log(y / mu), 0.0)
                b = np.where(y < 1, (1 - y) * np.log((1 - y) / (1 - mu)), 0.0)
            return 2.0 * np.sum(a + b)
        return float(np.sum((y - mu) ** 2))

    dev_old = deviance(mu)
    beta = np.zeros(X.shape[1])
    for _ in range(maxit):
        me = family.mu_eta(eta)
        z = (eta - off) + (y - mu) / me
        wt = me ** 2 / family.variance(mu)
        Xw = X * wt[:, None]
        beta = np.linalg.solve(X.T @ Xw, Xw.T @ z)
        eta = X @ beta + off
        mu = family.linkinv(eta)
        dev = deviance(mu)
        if abs(dev - dev_old) / (abs(dev) + 0.1) < epsilon:
            break
        dev_old = dev
    return {"coef": beta, "eta": eta, "mu": mu, "offset": off}


def glm_start_library(y, X, family, offset=None):
    """glm_start by the library's C++ IRLS (saige_glm_start, the same
    restatement in FP64 with array exp/log, about 4 ms at n = 100k; the call
    holds no Python lock).  The drivers take it with lib=True, as the GPU
    paths (fit_batched, fit_plain, the serving loop) run them; the dense test
    executors keep glm_start."""
    from .api import glm_start_native
    return glm_start_native("binary" if family.name == "binomial" else "quantitative", y, X, offset)


def _free_trace_buffers(sess, bufs):
    for per_sys in bufs.values():
        for b in per_sys.values():
            sess.free_device(b)


def _take_trace_buffers(sess, N):
    """The trace rounds' device buffers (system -> buffers) parked on sess by
    the last scheduler, if they were made for N samples; the session holds
    none while a scheduler has them."""
    park = getattr(sess, "_trace_buffers", None)
    sess._trace_buffers = None
    if park is None:
        return {}
    n, bufs = park
    if n == N:
        return bufs
    _free_trace_buffers(sess, bufs)
    return {}


def release_trace_buffers(sess):
    """Free the trace rounds' device buffers parked on sess."""
    park = getattr(sess, "_trace_buffers", None)
    sess._trace_buffers = None
    if park is not None:
        _free_trace_buffers(sess, park[1])


_POOLS = {}


def _shared_pool(name, workers):
    """A process-wide thread pool, made on first use and kept: starting a
    pool's threads anew costs milliseconds in every round of fits.  The
    "step" pool has one worker, so engine iterations handed to it run one at
    a time."""
    pool = _POOLS.get(name)
    if pool is None:
        pool = _POOLS[name] = concurrent.futures.ThreadPoolExecutor(max_workers=workers,
                                                                    thread_name_prefix=f"saige-{name}")
    return pool


# ------------------------------------------------------------------ probes


_M1, _M2 = np.uint64(0xBF58476D1CE4E5B9), np.uint64(0x94D049BB133111EB)
_C1, _C2, _C3 = np.uint64(0x9E3779B97F4A7C15), np.uint64(0xC2B2AE3D27D4EB4F), np.uint64(0x5851F42D)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def probe_vector(seed, i, n):
    """Probe i: entry r is -1 if the top bit of
    mix(seed ^ mix(i*C1 + mix(r*C2 + C3))) is set, else +1 (64-bit wrapping
    arithmetic; the same hash the device's saige_probes_device evaluates)."""
    r = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        inner = _mix64(r * _C2 + _C3)
        h = _mix64(np.uint64(seed) ^ _mix64(np.uint64(i) * _C1 + inner))
    return np.where((h >> np.uint64(63)) != 0, -1.0, 1.0)


class RademacherProbes:
    """Deterministic Rademacher probe vectors u_i in {-1, +1}^n.

    The reference draws u = 2*rbinom(n, 1, 0.5) - 1 from R's RNG after
    comm.set.seed(200) at every GetTrace call (FG.cpp:3114, 3133-3137), so
    every AI iteration reuses the same probes; here probe i is a fixed
    function of (seed, i) with the same property (probe_vector), which the
    device can also generate in place."""

    _shared = {}

    def __init__(self, n, seed=200):
        self.n = n
        self.seed = seed
        self._cache = {}

    @classmethod
    def shared(cls, n, seed=200):
        """One instance per (n, seed) for the process: the probes are a fixed
        function of (seed, i), so fits of the same size share the vectors
        (and their cached blocks) instead of regenerating them."""
        key = (int(n), int(seed))
        p = cls._shared.get(key)
        if p is None:
            p = cls._shared[key] = cls(n, seed)
        return p

    def __call__(self, i):
        v = self._cache.get(i)
        if v is None:
            v = self._cache[i] = probe_vector(self.seed, i, self.n)
        return v

    def block(self, start, end):
        """Probes start..end-1 as an n x (end-start) array (cached)."""
        key = ("block", start, end)
        U = self._cache.get(key)
        if U is None:
            U = np.empty((self.n, end - start), order="F")
            for c, i in enumerate(range(start, end)):
                U[:, c] = self(i)
            self._cache[key] = U
        return U


# -------------------------------------------------- solver-level coroutines

class Options:
    def __init__(self, maxiter=20, tol=0.02, nrun=30, tolPCG=1e-5, maxiterPCG=500,
                 traceCVcutoff=0.0025):
        self.maxiter = maxiter
        self.tol = tol
        self.nrun = nrun
        self.tolPCG = tolPCG
        self.maxiterPCG = maxiterPCG
        self.traceCVcutoff = traceCVcutoff


def _inv_sympd(A):
    """arma::inv_sympd(symmatu(A)) with the pinv fallback (FG.cpp:3183-3191)."""
    A = np.triu(A) + np.triu(A, 1).T
    try:
        np.linalg.cholesky(A)
        return np.linalg.inv(A)
    except np.linalg.LinAlgError:
        return np.linalg.pinv(A)


def cols(*parts):
    """Column-major block from vectors and column blocks (the layout the
    C ABI takes, so no transposes on the way to the device)."""
    n = parts[0].shape[0]
    k = sum(1 if p.ndim == 1 else p.shape[1] for p in parts)
    out = np.empty((n, k), order="F")
    c = 0
    for p in parts:
        if p.ndim == 1:
            out[:, c] = p
            c += 1
        else:
            out[:, c:c + p.shape[1]] = p
            c += p.shape[1]
    return out


def cal_cv(x):
    """calCV (FG.cpp:3104-3110)."""
    x = np.asarray(x, np.float64)
    return (np.std(x, ddof=1) / np.mean(x)) / len(x)


def get_coefficients_co(Y, X, w, tau, opt, lib=False):
    """getCoefficients (FG.cpp:3167-3199): Sigma^-1 [Y, X] in one request.
    lib: the algebra after the solve by the library's C++
    (saige_coefficients; X column-major FP64), else numpy."""
    sol = yield ("solve", w, tau, cols(Y, X))
    if lib:
        from .api import coefficients_native
        sol = np.asfortranarray(sol, np.float64)
        cov, alpha, eta = coefficients_native(Y, X, w, tau[0], sol)
        return {"Sigma_iY": sol[:, 0], "Sigma_iX": sol[:, 1:], "cov": cov, "alpha": alpha, "eta": eta}
    Sigma_iY = np.ascontiguousarray(sol[:, 0])
    Sigma_iX = np.ascontiguousarray(sol[:, 1:])
    cov = _inv_sympd(X.T @ Sigma_iX)
    alpha = cov @ (Sigma_iX.T @ Y)
    eta = Y - tau[0] * (Sigma_iY - Sigma_iX @ alpha) / w
    return {"Sigma_iY": Sigma_iY, "Sigma_iX": Sigma_iX, "cov": cov, "alpha": alpha, "eta": eta}


def get_coef_co(y, X, tau, fam, alpha0, eta0, offset, opt, lib=False):
    """Get_Coef (FG.R:2-35).  lib: the working response and the algebra
    after each solve by the library's C++ (saige_glm_working,
    saige_coefficients), else numpy."""
    tol_coef = 0.1
    if lib:
        from .api import glm_working_native
        trait = "binary" if fam.name == "binomial" else "quantitative"
        X = np.asfortranarray(X, np.float64)

        def working(e):
            return glm_working_native(trait, y, e, offset)
    else:
        def working(e):
            mu = fam.linkinv(e)
            me = fam.mu_eta(e)
            return mu, e - offset + (y - mu) / me, (me / np.sqrt(fam.variance(mu))) ** 2
    mu, Y, W = working(eta0)
    for _ in range(opt.maxiter):
        rc = yield from get_coefficients_co(Y, X, W, tau, opt, lib)
        alpha = rc["alpha"]
        eta = rc["eta"] + offset
        mu, Y, W = working(eta)
        if np.max(np.abs(alpha - alpha0) / (np.abs(alpha) + np.abs(alpha0) + tol_coef)) < tol_coef:
            break
        alpha0 = alpha
    return {"Y": Y, "alpha": alpha, "eta": eta, "W": W, "cov": rc["cov"],
            "Sigma_iY": rc["Sigma_iY"], "Sigma_iX": rc["Sigma_iX"], "mu": mu}


def _proj(V, Sigma_iV, Sigma_iX, cov):
    """P v = Sigma^-1 v - Sigma^-1 X cov (Sigma^-1 X)^T v, column-wise."""
    return Sigma_iV - Sigma_iX @ (cov @ (Sigma_iX.T @ V))


def ai_score_co(Y, X, w, tau, Sigma_iY, Sigma_iX, cov, probes, opt, quantitative=False):
    """getAIScore (FG.cpp:3279-3295) / getAIScore_q (FG.cpp:3479-3531) with
    GetTrace / GetTrace_q (FG.cpp:3113-3160, 3409-3472).

    Each round of probes is one "trace" request (see trace_round_host): the
    executor returns the FP64 scalars the estimator needs, and with PY also
    K PY and the PAPY (and PA0PY) solves.  With P u = Sigma^-1 u -
    Sigma^-1 X cov (Sigma^-1 X)^T u, the probe terms are
      (K u)^T P u = d - h^T cov g,   u^T P u = e - g^T cov g,
    d = (K u)^T Sigma^-1 u, e = u^T Sigma^-1 u, h = (Sigma^-1 X)^T K u,
    g = (Sigma^-1 X)^T u."""
    if quantitative:
        cov = _inv_sympd(X.T @ Sigma_iX)   # getAIScore_q recomputes cov (FG.cpp:3487-3494)
    PY = Sigma_iY - Sigma_iX @ (cov @ (Sigma_iX.T @ Y))
    t1, t0 = [], []
    start, end = 0, opt.nrun
    APY = None
    extra = None
    while True:
        r = yield ("trace", w, tau, probes, start, end, Sigma_iX, quantitative, PY if APY is None else None)
        if APY is None:
            APY, extra = r["APY"], r["extra"]
        hcg = np.einsum("jp,pq,jq->j", r["h"], cov, r["g"])
        t1.extend((r["d"] - hcg).tolist())
        if quantitative:
            t0.extend((r["e"] - np.einsum("jp,pq,jq->j", r["g"], cov, r["g"])).tolist())
        cv = cal_cv(t1)
        cv0 = cal_cv(t0) if quantitative else 0.0
        if cv > opt.traceCVcutoff or (quantitative and cv0 > opt.traceCVcutoff):
            start, end = end, end + 10    # FG.cpp:3147-3154
            continue
        break
    YPAPY = PY @ APY
    PAPY = _proj(APY, extra[:, 0], Sigma_iX, cov)
    out = {"YPAPY": YPAPY, "PY": PY, "nprobe": len(t1)}
    if not quantitative:
        out["Trace"] = float(np.mean(t1))
        out["AI"] = APY @ PAPY
    else:
        A0PY = PY
        PA0PY = _proj(A0PY, extra[:, 1], Sigma_iX, cov)
        out["YPA0PY"] = PY @ A0PY
        out["Trace"] = np.array([np.mean(t0), np.mean(t1)])
        out["AI"] = np.array([[A0PY @ PA0PY, A0PY @ PAPY], [A0PY @ PAPY, APY @ PAPY]])
    return out


def probe_block(probes, start, end):
    if hasattr(probes, "block"):
        return probes.block(start, end)
    return np.column_stack([probes(i) for i in range(start, end)])


def trace_round_co(req):
    """A "trace" request as ordinary products and solves: K [PY, U], then
    Sigma^-1 [U, APY, (PY)], and the scalars d, e, h, g in FP64 on the host.
    Executors without device-resident trace rounds run it in place of the
    request."""
    _, w, tau, probes, start, end, SX, want_t0, PY = req
    U = probe_block(probes, start, end)
    k = U.shape[1]
    out = {}
    if PY is not None:
        KB = yield ("matvec", cols(PY, U))
        APY, AU = KB[:, 0], KB[:, 1:]
        sol = yield ("solve", w, tau, cols(*([U, APY] + ([PY] if want_t0 else []))))
        SU = sol[:, :k]
        out["APY"], out["extra"] = APY, sol[:, k:]
    else:
        AU = yield ("matvec", U)
        SU = yield ("solve", w, tau, U)
    out["d"] = np.array([np.dot(AU[:, j], SU[:, j]) for j in range(k)])
    out["h"] = AU.T @ SX
    out["g"] = U.T @ SX
    if want_t0:
        out["e"] = np.array([np.dot(U[:, j], SU[:, j]) for j in range(k)])
    return out


def trace_round_host(req, matvec, solve):
    """Run trace_round_co with a product and a solve function."""
    gen = trace_round_co(req)
    value = None
    while True:
        try:
            r = gen.send(value)
        except StopIteration as e:
            return e.value
        value = matvec(r[1]) if r[0] == "matvec" else solve(*r[1:])


def fitglmmai_rpcg_co(Y, X, w, tau, Sigma_iY, Sigma_iX, cov, probes, opt):
    """fitglmmaiRPCG (FG.cpp:3302-3341)."""
    re = yield from ai_score_co(Y, X, w, tau, Sigma_iY, Sigma_iX, cov, probes, opt)
    Dtau = (re["YPAPY"] - re["Trace"]) / re["AI"]
    tau0 = np.array(tau, np.float64)
    tau = tau0.copy()
    tau[1] = tau0[1] + Dtau
    tau[tau < opt.tol] = 0.0
    step = 1.0
    while tau[1] < 0.0:
        step *= 0.5
        tau[1] = tau0[1] + step * Dtau
    tau[tau < opt.tol] = 0.0
    return tau, re


def fitglmmai_rpcg_q_co(Y, X, w, tau, Sigma_iY, Sigma_iX, cov, probes, opt):
    """fitglmmaiRPCG_q (FG.cpp:3610-3662): 2x2 AI system, GMMAT zeroing."""
    tau = np.array(tau, np.float64)
    zero = tau < opt.tol
    re = yield from ai_score_co(Y, X, w, tau, Sigma_iY, Sigma_iX, cov, probes, opt, quantitative=True)
    score = np.array([re["YPA0PY"] - re["Trace"][0], re["YPAPY"] - re["Trace"][1]])
    Dtau = np.linalg.solve(re["AI"], score)
    tau0 = tau.copy()
    tau = tau0 + Dtau
    tau[zero & (tau < opt.tol)] = 0.0
    step = 1.0
    while tau[0] < 0.0 or tau[1] < 0.0:
        step *= 0.5
        tau = tau0 + step * Dtau
        tau[zero & (tau < opt.tol)] = 0.0
    tau[tau < opt.tol] = 0.0
    return tau, re


# ------------------------------------------------------------ R drivers

def fit_binary_co(y, X, probes, opt, offset=None, fit0=None, tau_init=(0.0, 0.0), lib=False):
    """glmmkin.ai_PCG_Rcpp_Binary (FG.R:79-304) without LOCO.  fit0: the
    glm() start values, else computed here.  lib: the start values, Get_Coef's
    working response and getCoefficients' algebra by the library's C++
    (glm_start_library, get_coef_co), else by numpy."""
    y = np.asarray(y, np.float64)
    X = np.asfortranarray(X, np.float64) if lib else np.asarray(X, np.float64)
    n = len(y)
    offset = np.zeros(n) if offset is None else np.asarray(offset, np.float64)
    fam = Binomial
    if fit0 is None:
        fit0 = (glm_start_library if lib else glm_start)(y, X, fam, offset)
    eta = fit0["eta"]
    alpha0 = fit0["coef"]
    eta0 = eta
    tau = np.array([1.0, 0.1 if tau_init[1] == 0 else float(tau_init[1])])   # FG.R:148-157
    tau0 = tau.copy()
    rc = yield from get_coef_co(y, X, tau, fam, alpha0, eta0, offset, opt, lib)
    re = yield from ai_score_co(rc["Y"], X, rc["W"], tau, rc["Sigma_iY"], rc["Sigma_iX"], rc["cov"],
                                probes, opt)
    tau[1] = max(0.0, tau0[1] + tau0[1] ** 2 * (re["YPAPY"] - re["Trace"]) / n)   # FG.R:165
    history = [tau.copy()]
    it = 0
    for it in range(1, opt.maxiter + 1):
        alpha0 = rc["alpha"]
        tau0 = tau.copy()
        eta0 = eta
        rc = yield from get_coef_co(y, X, tau, fam, alpha0, eta0, offset, opt, lib)
        tau, _ = yield from fitglmmai_rpcg_co(rc["Y"], X, rc["W"], tau, rc["Sigma_iY"], rc["Sigma_iX"],
                                              rc["cov"], probes, opt)
        eta = rc["eta"]
        history.append(tau.copy())
        if tau[1] == 0:
            break
        if np.max(np.abs(tau - tau0) / (np.abs(tau) + np.abs(tau0) + opt.tol)) < opt.tol:   # FG.R:206
            break
        if np.max(tau) > opt.tol ** (-2):
            it = opt.maxiter
            break
    rc = yield from get_coef_co(y, X, tau, fam, rc["alpha"], eta, offset, opt, lib)  # FG.R:218
    return {"theta": tau, "coefficients": rc["alpha"], "linear_predictors": rc["eta"],
            "fitted_values": rc["mu"], "Y": rc["Y"], "cov": rc["cov"], "converged": it < opt.maxiter,
            "history": history, "iterations": it, "traitType": "binary"}


def fit_quantitative_co(y, X, probes, opt, offset=None, fit0=None, lib=False):
    """glmmkin.ai_PCG_Rcpp_Quantitative (FG.R:309-549) without LOCO; fit0
    and lib as in fit_binary_co."""
    y = np.asarray(y, np.float64)
    X = np.asfortranarray(X, np.float64) if lib else np.asarray(X, np.float64)
    n = len(y)
    offset = np.zeros(n) if offset is None else np.asarray(offset, np.float64)
    fam = Gaussian
    if fit0 is None:
        fit0 = (glm_start_library if lib else glm_start)(y, X, fam, offset)
    eta = fit0["eta"]
    alpha = fit0["coef"]
    alpha0 = alpha
    eta0 = eta
    tau = np.array([1.0, 0.0])                                   # FG.R:389-398
    tau0 = tau.copy()
    rc = yield from get_coef_co(y, X, tau, fam, alpha0, eta0, offset, opt, lib)
    re = yield from ai_score_co(rc["Y"], X, rc["W"], tau, rc["Sigma_iY"], rc["Sigma_iX"], rc["cov"],
                                probes, opt, quantitative=True)
    tau[1] = max(0.0, tau0[1] + tau0[1] ** 2 * (re["YPAPY"] - re["Trace"][1]) / n)   # FG.R:405-406
    tau[0] = max(0.0, tau0[0] + tau0[0] ** 2 * (re["YPA0PY"] - re["Trace"][0]) / n)
    history = [tau.copy()]
    it = 0
    for it in range(1, opt.maxiter + 1):
        alpha0 = alpha
        tau0 = tau.copy()
        eta0 = eta
        rc = yield from get_coef_co(y, X, tau, fam, alpha0, eta0, offset, opt, lib)
        tau, _ = yield from fitglmmai_rpcg_q_co(rc["Y"], X, rc["W"], tau, rc["Sigma_iY"], rc["Sigma_iX"],
                                                rc["cov"], probes, opt)
        alpha = rc["alpha"]
        eta = rc["eta"]
        history.append(tau.copy())
        if tau[0] <= 0:
            raise RuntimeError("ERROR! The first variance component parameter estimate is 0")   # FG.R:449-451
        if tau[1] <= 0:
            break
        if np.max(np.abs(tau - tau0) / (np.abs(tau) + np.abs(tau0) + opt.tol)) < opt.tol:     # FG.R:456
            break
        if np.max(tau) > opt.tol ** (-2):
            it = opt.maxiter
            break
    rc = yield from get_coef_co(y, X, tau, fam, alpha, eta, offset, opt, lib)       # FG.R:475
    return {"theta": tau, "coefficients": rc["alpha"], "linear_predictors": rc["eta"],
            "fitted_values": rc["mu"], "Y": rc["Y"], "cov": rc["cov"], "converged": it < opt.maxiter,
            "history": history, "iterations": it, "traitType": "quantitative"}


def fit_co(trait, y, X, probes, opt, **kw):
    if trait == "binary":
        return fit_binary_co(y, X, probes, opt, **kw)
    if trait == "quantitative":
        return fit_quantitative_co(y, X, probes, opt, **kw)
    raise ValueError(f"unknown trait type {trait!r}")


# ------------------------------------------------ leave one chromosome out

def update_chr_start_end_index(chr_vec):
    """updateChrStartEndIndexVec (Util.R:29-68): for chromosomes 1..22 the
    first and last 0-based position of that chromosome in chr_vec (None where
    absent).  Raises if the chromosomes are not in order; LOCO is off when
    fewer than two autosomes are present.  Returns (loco, starts, ends)."""
    chr_vec = [str(c).strip() for c in chr_vec]
    starts, ends = [], []
    for i in range(1, 23):
        pos = [k for k, c in enumerate(chr_vec) if c == str(i)]
        if pos:
            starts.append(min(pos))
            ends.append(max(pos))
            prev_s, prev_e = starts[i - 2] if i > 1 else None, ends[i - 2] if i > 1 else None
            if prev_s is not None and (starts[-1] <= prev_s or ends[-1] <= prev_e):
                raise ValueError("ERROR! chromosomes need to be ordered from 1 to 22 in the bim file.")
        else:
            starts.append(None)
            ends.append(None)
    present = sum(s is not None for s in starts)
    return present > 1, starts, ends


def loco_co(fit, y, X, trait, opt, offset=None):
    """One chromosome of fitNULLGLMM's LOCO loop (FG.R:1240-1272 binary,
    FG.R:1486-1512 quantitative): Get_Coef_LOCO (FG.R:42-73) from the full
    fit's theta, coefficients and linear predictors, whose solves the
    executor runs on Sigma_loco, then ScoreTest_NULL_Model with mu2 =
    mu(1-mu) (binary) or 1/tau[0] (quantitative)."""
    y = np.asarray(y, np.float64)
    X = np.asarray(X, np.float64)
    offset = np.zeros(len(y)) if offset is None else np.asarray(offset, np.float64)
    fam = Binomial if trait == "binary" else Gaussian
    tau = np.asarray(fit["theta"], np.float64)
    rc = yield from get_coef_co(y, X, tau, fam, fit["coefficients"], fit["linear_predictors"], offset, opt)
    mu = rc["mu"]
    mu2 = mu * (1 - mu) if trait == "binary" else np.full(len(y), 1.0 / tau[0])
    return {"isLOCO": True, "coefficients": rc["alpha"], "linear_predictors": rc["eta"],
            "fitted_values": mu, "Y": rc["Y"], "residuals": y - mu, "cov": rc["cov"],
            "obj_noK": score_test_null(mu, mu2, y, X)}


# ---------------------------------------------------------- variance ratio

def score_test_null(mu, mu2, y, X):
    """ScoreTest_NULL_Model (FG.R:579-591)."""
    V = np.asarray(mu2, np.float64)
    XV = (X * V[:, None]).T
    XVX = X.T @ XV.T
    XVX_inv = np.linalg.inv(XVX)
    XXVX_inv = X @ XVX_inv
    return {"XV": XV, "XVX": XVX, "XXVX_inv": XXVX_inv, "XVX_inv": XVX_inv,
            "S_a": (X * (y - mu)[:, None]).sum(0), "XVX_inv_XV": XXVX_inv * V[:, None], "V": V}


def variance_ratio_co(fit, y, X, trait, genos, categories, opt, num_markers=30, ratio_cv_cutoff=0.001):
    """The variance-ratio core of extractVarianceRatio (FG.R:2152-2423) for a
    fitted null model.

    genos(i): integer genotypes (length n) of candidate marker i, imputed as
    repacked (Get_OneSNP_Geno[_forVarRatio], FG.cpp:2281-2300).
    categories: one list of candidate ids per MAC/MAF category, already in
    the random order the reference draws with sample() (FG.R:2239-2281).
    Each round's Sigma^-1 G solves go out as one block request.  Returns
    the table rows (ratio, "null", k) and the per-marker ratios."""
    y = np.asarray(y, np.float64)
    X = np.asarray(X, np.float64)
    fam = Binomial if trait == "binary" else Gaussian
    eta = np.asarray(fit["linear_predictors"], np.float64)
    mu = np.asarray(fit["fitted_values"], np.float64)
    W = (fam.mu_eta(eta) / np.sqrt(fam.variance(mu))) ** 2            # FG.R:2194-2196
    tau = np.asarray(fit["theta"], np.float64)
    Sigma_iX = yield ("solve", W, tau, cols(X))                        # getSigma_X, FG.R:2200
    noK = score_test_null(mu, mu * (1.0 - mu), y, X)
    XtSiX_inv = np.linalg.inv(X.T @ Sigma_iX)
    n = len(y)
    table, per_cat = [], []
    for k, cand in enumerate(categories):
        ratios = []
        idx = 0
        target = num_markers
        cv = ratio_cv_cutoff + 0.1
        while cv > ratio_cv_cutoff:
            batch = []
            while len(ratios) + len(batch) < target and idx < len(cand):
                G0 = np.asarray(genos(cand[idx]), np.float64)
                idx += 1
                if G0.sum() / (2 * n) > 0.5:                           # FG.R:2306-2308
                    G0 = 2.0 - G0
                AC = G0.sum()
                G = G0 - noK["XXVX_inv"] @ (noK["XV"] @ G0)            # FG.R:2330
                batch.append((G, AC))
            if batch:
                SiG = yield ("solve", W, tau, cols(*[b[0] for b in batch]))   # getSigma_G, FG.R:2338
                for c, (G, AC) in enumerate(batch):
                    sg = SiG[:, c]
                    var1a = G @ sg - (G @ Sigma_iX) @ XtSiX_inv @ (X.T @ sg)    # FG.R:2341
                    var1 = var1a / AC
                    g = G / np.sqrt(AC)
                    var2 = np.sum(mu * (1.0 - mu) * g * g) if trait == "binary" else g @ g   # FG.R:2360-2364
                    ratios.append(var1 / var2)
            cv = cal_cv(ratios)
            if cv > ratio_cv_cutoff:
                target += 10                                           # FG.R:2381-2384
            if idx == len(cand):
                cv = ratio_cv_cutoff                                   # FG.R:2389-2393
        table.append((float(np.mean(ratios)), "null", k + 1))
        per_cat.append(np.array(ratios))
    return {"table": table, "ratios": per_cat}


def write_variance_ratio(path, table):
    """<prefix>.varianceRatio.txt as write.table(quote=F, col.names=F,
    row.names=F) prints it (FG.R:2416-2418): "value type category"."""
    with open(path, "w") as f:
        for v, kind, k in table:
            f.write(f"{v:.15g} {kind} {k}\n")


# --------------------------------------------------------------- executors

_ALLOCATOR_TUNED = False


def tune_host_allocator():
    """Serve the per-request N x k host blocks from the heap instead of fresh
    mmaps: glibc maps every allocation above its mmap threshold anew and
    returns it on free, so each block pays its page faults again (tens of ms
    per 100k x 16 block).  Raising the mmap and trim thresholds keeps freed
    blocks for reuse.  Process-wide and idempotent; returns whether glibc
    accepted both settings."""
    global _ALLOCATOR_TUNED
    if _ALLOCATOR_TUNED:
        return True
    try:
        libc = ctypes.CDLL("libc.so.6")
    except OSError:
        return False
    M_TRIM_THRESHOLD, M_MMAP_THRESHOLD = -1, -3
    ok = bool(libc.mallopt(M_MMAP_THRESHOLD, 1 << 30)) or bool(libc.mallopt(M_MMAP_THRESHOLD, 32 << 20))
    ok = bool(libc.mallopt(M_TRIM_THRESHOLD, (1 << 31) - 1)) and ok
    _ALLOCATOR_TUNED = ok
    return ok

class PlainExecutor:
    """Runs a fit coroutine the way the reference does: every right-hand side
    is its own single-RHS PCG (saige_pcg1), every K b its own GEMV.  With
    loco=True they are the _LOCO variants (getPCG1ofSigmaAndVector_LOCO,
    getCrossprodMatAndKin_LOCO) on the range set by setStartEndIndex."""

    def __init__(self, sess, opt, loco=False):
        self.sess = sess
        self.opt = opt
        self.n_solves = 0
        self.pcg_iters = 0
        self.n_products = 0
        self._pcg = sess.getPCG1ofSigmaAndVector_LOCO if loco else sess.getPCG1ofSigmaAndVector
        self._kin = sess.getCrossprodMatAndKin_LOCO if loco else sess.getCrossprodMatAndKin

    def _solve(self, w, tau, B):
        tau32 = np.asarray(tau, np.float32)
        out = np.zeros(B.shape, np.float64, order="F")   # the scheduler's result layout
        for j in range(B.shape[1]):
            x, it = self._pcg(w, tau32, B[:, j], self.opt.maxiterPCG, self.opt.tolPCG, return_iters=True)
            out[:, j] = x
            self.n_solves += 1
            self.pcg_iters += it
        return out

    def _matvec(self, B):
        self.n_products += B.shape[1]
        return np.asfortranarray(np.column_stack([self._kin(B[:, j]) for j in range(B.shape[1])]),
                                 dtype=np.float64)

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
                value = self._matvec(req[1])
            else:
                value = trace_round_host(req, self._matvec, self._solve)


class _Fit:
    __slots__ = ("fid", "gen", "subset", "n", "sys", "pending", "result_buf", "remaining", "result",
                 "error", "t_submit", "t_admit", "t_done", "tag", "iters", "sys_ready", "sys_pending",
                 "trace", "sub")

    def __init__(self, fid, gen, subset, tag):
        self.fid = fid
        self.gen = gen
        self.subset = subset
        self.sys = None
        self.pending = collections.deque()
        self.result_buf = None
        self.remaining = 0
        self.result = None
        self.error = None
        self.t_submit = None
        self.t_admit = None
        self.t_done = None
        self.tag = tag
        self.iters = 0
        self.sys_ready = False
        self.sys_pending = None
        self.trace = None
        self.sub = None          # a trace round run as products and solves


class BatchScheduler:
    """Continuous batching of PCG iterations over the shared resident GRM.

    Fits are admitted while a system slot (per-fit w, tau, sample mask on the
    device) is free; their solve columns fill the engine's slots in FIFO
    order, join at the next iteration, and leave when they converge; a fit
    resumes as soon as its request's last column is done.  GRM products
    (K u for the trace probes, K PY) are product slots that ride along in
    the next iteration's GEMM.  A column's arithmetic does not depend on its
    neighbours, so each fit gets the same result alone and in the mix.

    Host work overlaps the device: when columns finish, the next iteration of
    the still-running columns is launched from a worker thread (the library
    call releases the GIL) while the main thread resumes the finished fits'
    coroutines.  Nothing touches the library during that time; a resumed
    fit's new system and columns are applied once the iteration is back.

    Trace rounds stay on the device: the probes are generated (or uploaded)
    into a per-system device buffer, K u goes from its product slot to a
    second buffer with h = (Sigma^-1 X)^T K u, and the Sigma^-1 u solves are
    never downloaded: their slots return d = (K u)^T Sigma^-1 u (and
    e = u^T Sigma^-1 u) as FP64 dots.  Only PY, APY and the PAPY/PA0PY
    solves travel as vectors."""

    TRACE_MAX_PROBES = 64      # probe columns per round held on the device
    TRACE_MAX_COVARIATES = 14  # shared dot partners per take (16 - 2 pairwise)

    def __init__(self, sess, opt, width=64, n_systems=16, hbm_reserve_bytes=8 << 30, tune_malloc=True,
                 host_threads=1):
        if tune_malloc:
            tune_host_allocator()
        self.sess = sess
        self.host_threads = max(1, int(host_threads))
        self._host_pool = _shared_pool(f"host{self.host_threads}", self.host_threads) \
            if self.host_threads > 1 else None
        self.opt = opt
        self.width = width
        self.N = sess.getNnomissingOut()
        # admission by free HBM: never hold more systems than half of what
        # is free beyond a reserve allows
        free, _ = sess.device_memory()
        per = sess.engine_system_bytes()
        by_mem = max(1, int(max(0, free - hbm_reserve_bytes) // 2 // max(1, per)))
        self.n_systems = n_systems = max(1, min(n_systems, by_mem))
        self.hbm_reserve_bytes = hbm_reserve_bytes
        sess.engine_reserve(width, n_systems)
        # lowest free slot first: live columns pack into the low 32-column
        # half, and the GEMM skips a half with no occupied slot
        self.free_slots = list(range(width))
        self.free_sys = collections.deque(range(n_systems))
        self.slot_owner = {}
        self.waiting = collections.deque()      # submitted, not admitted
        self.active = []                        # admitted
        self.finished = []
        self._next_id = 0
        self.n_iterations = 0
        self.n_solves = 0
        self.busy_slot_iters = 0
        # wall-time split of run(): device iterations, slot traffic, fit logic
        self.timing = collections.Counter()
        self._pool = _shared_pool("step", 1)
        self._inflight = None                   # engine_step running in the worker
        self.n_overlapped = 0                   # iterations that ran during host work
        # system -> device buffers of its trace rounds, kept on the session
        # between schedulers (allocating and freeing them costs each round ms);
        # taken from the session at the first trace round, so a scheduler that
        # runs none (variance ratios, LOCO) leaves them parked there
        self._devbufs = None
        self._rows_owner = {}                   # system -> fit whose subset rows its buffer holds
        self.device_trace = hasattr(sess, "engine_take_batch_ex")

    # ---------------------------------------------------------------- API
    def submit(self, gen, subset=None, tag=None, t_submit=None):
        f = _Fit(self._next_id, gen, None if subset is None else np.asarray(subset, np.int64), tag)
        self._next_id += 1
        f.t_submit = time.perf_counter() if t_submit is None else t_submit
        self.waiting.append(f)
        return f

    def idle(self):
        return not self.waiting and not self.active

    def close(self):
        """Park the trace rounds' device buffers on the session for the next
        scheduler (release_trace_buffers frees them), or free them when
        another scheduler parked its own meanwhile.  Idempotent."""
        if self._devbufs is None:
            return
        if getattr(self.sess, "_trace_buffers", None) is None:
            self.sess._trace_buffers = (self.N, self._devbufs)
        else:
            _free_trace_buffers(self.sess, self._devbufs)
        self._devbufs = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def run(self, arrivals=None):
        """Run until every submitted fit finished.  arrivals: optional list of
        (t_offset_seconds, gen, subset, tag) released as wall time passes."""
        t0 = time.perf_counter()
        arrivals = collections.deque(sorted(arrivals or [], key=lambda a: a[0]))
        while arrivals or not self.idle():
            now = time.perf_counter()
            while arrivals and t0 + arrivals[0][0] <= now:
                toff, gen, subset, tag = arrivals.popleft()
                self.submit(gen, subset, tag, t_submit=t0 + toff)
            if self._inflight is None:           # the engine is only changed between iterations
                self._admit()
                self._fill()
            if self._inflight is not None or self.slot_owner:
                self._step()
            elif arrivals and self.idle():
                time.sleep(max(0.0, min(0.05, t0 + arrivals[0][0] - time.perf_counter())))
        return self.finished

    # ----------------------------------------------------------- internals
    def _admit(self):
        if self.waiting and self.free_sys:
            free, _ = self.sess.device_memory()
            if free < self.hbm_reserve_bytes:
                return
        admitted = []
        while self.waiting and self.free_sys:
            f = self.waiting.popleft()
            f.sys = self.free_sys.popleft()
            f.t_admit = time.perf_counter()
            self.active.append(f)
            admitted.append((f, None))
        if admitted:
            t = time.perf_counter()
            self._advance(admitted)
            self.timing["host"] += time.perf_counter() - t

    def _expand(self, f, v):
        if f.subset is None:
            return v
        out = np.zeros(self.N, v.dtype)
        out[f.subset] = v
        return out

    def _compact(self, f, v):
        return v if f.subset is None else v[f.subset]

    def _send(self, f, value):
        """Resume a fit's coroutine: ("req", request) | ("done", result) |
        ("error", exception).  Touches nothing but the fit's own generators,
        so several fits may be resumed on worker threads at once.  Without
        device-resident trace rounds a "trace" request runs as
        trace_round_co's products and solves."""
        try:
            while True:
                if f.sub is not None:
                    try:
                        req = f.sub.send(value)
                    except StopIteration as e:
                        f.sub, value = None, e.value
                        continue
                else:
                    req = f.gen.send(value)
                if req[0] == "trace" and not self._device_round(req):
                    f.sub, value = trace_round_co(req), None
                    continue
                return ("req", req)
        except StopIteration as e:
            return ("done", e.value)
        except Exception as e:   # a failing fit does not stop the others
            return ("error", e)

    def _apply(self, f, out):
        kind, v = out
        if kind == "done":
            f.result = v
            self._finish(f)
            return
        if kind == "error":
            f.error = v
            self._finish(f)
            return
        req = v
        if req[0] == "trace":
            self._trace_start(f, req)
            return
        if req[0] == "matvec":
            B = req[1]
            if not f.sys_ready:          # products only need the system's sample mask
                f.sys_pending = (np.ones(B.shape[0]), (1.0, 0.0))
                f.sys_ready = True
            kind = "product"
        else:
            _, w, tau, B = req
            f.sys_pending = (w, tau)
            f.sys_ready = True
            kind = "solve"
        f.result_buf = np.empty(B.shape, np.float64, order="F")    # every column is taken before resuming
        f.remaining = B.shape[1]
        for j in range(B.shape[1]):
            f.pending.append((j, B[:, j], kind, ("host", j)))

    def _advance(self, items):
        """Resume fits [(fit, value), ...] and queue their requests in order.
        With host_threads > 1 the coroutines run on a thread pool (numpy
        releases the GIL); on the benchmark the per-fit numpy work is too
        small for that to pay (thread hand-off costs more), so the default
        is 1."""
        if len(items) > 1 and self.host_threads > 1:
            outs = list(self._host_pool.map(lambda fv: self._send(*fv), items))
        else:
            outs = [self._send(f, v) for f, v in items]
        for (f, _), out in zip(items, outs):
            self._apply(f, out)

    # ------------------------------------------------------- trace rounds
    def _device_round(self, req):
        """Whether a trace round fits the device-resident path: at most
        TRACE_MAX_PROBES probes and TRACE_MAX_COVARIATES covariates; larger
        rounds run as trace_round_co's products and solves."""
        _, _, _, _, start, end, SX = req[:7]
        return self.device_trace and end - start <= self.TRACE_MAX_PROBES \
            and SX.shape[1] <= self.TRACE_MAX_COVARIATES

    def _buffers(self, sys):
        if self._devbufs is None:
            self._devbufs = _take_trace_buffers(self.sess, self.N)
        bufs = self._devbufs.get(sys)
        if bufs is None:
            n = self.N * 4
            bufs = self._devbufs[sys] = {
                "U": self.sess.alloc_device(n * self.TRACE_MAX_PROBES),
                "AU": self.sess.alloc_device(n * self.TRACE_MAX_PROBES),
                "SX": self.sess.alloc_device(n * self.TRACE_MAX_COVARIATES),
            }
        return bufs

    def _trace_start(self, f, req):
        """Phase 1 of a trace round: probes into the device buffer, products
        K [PY, U]; K u is kept on the device."""
        _, w, tau, probes, start, end, SX, want_t0, PY = req
        k, p = end - start, SX.shape[1]
        assert k <= self.TRACE_MAX_PROBES and p <= self.TRACE_MAX_COVARIATES   # _device_round
        # device work (buffers, probes, uploads) waits for _fill: this may run
        # while an iteration is in flight, when nothing may touch the library
        f.trace = {"phase": 1, "w": w, "tau": tau, "k": k, "p": p, "want_t0": want_t0, "PY": PY,
                   "probes": probes, "start": start, "end": end, "SX": SX, "setup": True,
                   "h": np.zeros((k, p)), "d": np.zeros(k), "e": np.zeros(k), "bufs": None}
        if not f.sys_ready:              # products only need the system's sample mask
            f.sys_pending = (np.ones(len(PY) if PY is not None else self._n_of(f)), (1.0, 0.0))
            f.sys_ready = True
        n_host = 1 if PY is not None else 0
        f.result_buf = np.empty((self._n_of(f), n_host), np.float64, order="F")
        f.remaining = n_host + k
        if PY is not None:
            f.pending.append((0, PY, "product", ("host", 0)))
        for j in range(k):
            f.pending.append((j, ("dev", "U", j), "product", ("stash", j)))

    def _trace_setup(self, f):
        """Device side of phase 1 (called from _fill, between iterations):
        probes into the system's U buffer (generated in place, or uploaded for
        a sample subset or other probe sources), Sigma^-1 X into SX, and
        g = (Sigma^-1 X)^T u."""
        tr = f.trace
        tr["setup"] = False
        bufs = tr["bufs"] = self._buffers(f.sys)
        k, p, probes = tr["k"], tr["p"], tr["probes"]
        sub = f.subset if f.subset is not None else slice(None)
        if f.subset is None and isinstance(probes, RademacherProbes) and probes.n == self.N:
            self.sess.probes_device(bufs["U"], 0, probes.seed, tr["start"], k)
        elif isinstance(probes, RademacherProbes) and probes.n == len(f.subset):
            # the subset's rows go to the device once per fit on this system
            if "rows" not in bufs:
                bufs["rows"] = self.sess.alloc_device(8 * self.N)
            if self._rows_owner.get(f.sys) != f.fid:
                self.sess.copy_bytes_to_device(bufs["rows"], f.subset)
                self._rows_owner[f.sys] = f.fid
            self.sess.probes_device_rows(bufs["U"], 0, probes.seed, tr["start"], k, bufs["rows"], len(f.subset))
        else:
            Uf = np.zeros((self.N, k), np.float32, order="F")
            Uf[sub] = probe_block(probes, tr["start"], tr["end"])
            self.sess.copy_host_to_device(bufs["U"], Uf)
        SXf = np.zeros((self.N, p), np.float32, order="F")
        SXf[sub] = tr["SX"]
        self.sess.copy_host_to_device(bufs["SX"], SXf)
        tr["g"] = self.sess.dots_device(bufs["U"], k, bufs["SX"], p)

    def _n_of(self, f):
        return self.N if f.subset is None else len(f.subset)

    def _trace_phase2(self, f):
        """Phase 2: solves Sigma^-1 [U, APY, (PY)]; the U solves return only dots."""
        tr = f.trace
        tr["phase"] = 2
        if tr["PY"] is not None:
            tr["APY"] = f.result_buf[:, 0].copy()
        f.sys_pending = (tr["w"], tr["tau"])
        f.sys_ready = True
        extra = [] if tr["PY"] is None else [tr["APY"]] + ([tr["PY"]] if tr["want_t0"] else [])
        f.result_buf = np.empty((self._n_of(f), len(extra)), np.float64, order="F")
        f.remaining = tr["k"] + len(extra)
        for j in range(tr["k"]):
            f.pending.append((j, ("dev", "U", j), "solve", ("pair", j)))
        for j, v in enumerate(extra):
            f.pending.append((j, v, "solve", ("host", j)))

    def _trace_result(self, f):
        tr = f.trace
        f.trace = None
        out = {"d": tr["d"], "h": tr["h"], "g": tr["g"]}
        if tr["want_t0"]:
            out["e"] = tr["e"]
        if tr["PY"] is not None:
            out["APY"], out["extra"] = tr["APY"], f.result_buf
        return out

    def _set_system(self, f, w, tau):
        mask = None
        wf = w
        if f.subset is not None:
            wf = np.ones(self.N)
            wf[f.subset] = w
            mask = np.zeros(self.N, np.uint8)
            mask[f.subset] = 1
        t = time.perf_counter()
        self.sess.engine_set_system(f.sys, wf, np.asarray(tau, np.float32), mask)
        self.timing["set_system"] += time.perf_counter() - t

    def _finish(self, f):
        f.t_done = time.perf_counter()
        self.active.remove(f)
        self.free_sys.append(f.sys)
        self.finished.append(f)

    def _fill(self):
        host = {"solve": ([], [], [], []), "product": ([], [], [], [])}
        dev = []                          # runs of consecutive device columns: [kind, buf, col0, slots, systems, owners]
        for f in self.active:
            if f.sys_pending is not None and f.pending and self.free_slots:
                w, tau = f.sys_pending
                f.sys_pending = None
                self._set_system(f, w, tau)
            if f.trace is not None and f.trace["setup"] and f.pending and self.free_slots:
                t = time.perf_counter()
                self._trace_setup(f)
                self.timing["trace_setup"] += time.perf_counter() - t
            while f.pending and self.free_slots:
                j, src, kind, action = f.pending.popleft()
                slot = heapq.heappop(self.free_slots)
                if isinstance(src, tuple):               # ("dev", buffer name, column)
                    _, name, col = src
                    buf = f.trace["bufs"][name]
                    run = dev[-1] if dev else None
                    if (run is None or run[0] != kind or run[1] is not buf or
                            run[2] + len(run[3]) != col or run[4][0] != f.sys):
                        run = [kind, buf, col, [], [], []]
                        dev.append(run)
                    run[3].append(slot)
                    run[4].append(f.sys)
                    run[5].append((f, j, action))
                else:
                    g = host[kind]
                    g[0].append(slot)
                    g[1].append(f.sys)
                    g[2].append((f, src))
                    g[3].append((f, j, action))
            if not self.free_slots:
                break
        t = time.perf_counter()
        for kind, (slots, systems, cols, owners) in host.items():
            if not slots:
                continue
            # column-major, as the ABI takes it; rows outside a subset stay 0
            alloc = np.zeros if any(f.subset is not None for f, _ in cols) else np.empty
            B = alloc((self.N, len(cols)), np.float32, order="F")
            for c, (f, b) in enumerate(cols):
                if f.subset is None:
                    B[:, c] = b
                else:
                    B[f.subset, c] = b
            if kind == "solve":
                self.sess.engine_add_batch(slots, systems, B, self.opt.maxiterPCG, self.opt.tolPCG)
            else:
                self.sess.engine_add_products(slots, systems, B)
            for slot, own in zip(slots, owners):
                self.slot_owner[slot] = own
        for kind, buf, col0, slots, systems, owners in dev:
            if kind == "solve":
                self.sess.engine_add_batch_device(slots, systems, buf, col0, self.opt.maxiterPCG, self.opt.tolPCG)
            else:
                self.sess.engine_add_products_device(slots, systems, buf, col0)
            for slot, own in zip(slots, owners):
                self.slot_owner[slot] = own
        self.timing["add"] += time.perf_counter() - t

    def _step(self):
        t = time.perf_counter()
        if self._inflight is not None:
            done, iters = self._inflight.result()
            self._inflight = None
            self.timing["step_wait"] += time.perf_counter() - t
        else:
            done, iters = self.sess.engine_step(1, self.width)
            self.timing["step"] += time.perf_counter() - t
        self.n_iterations += 1
        self.busy_slot_iters += len(self.slot_owner)
        finished = [slot for slot in self.slot_owner if done[slot]]
        if not finished:
            return
        host, dev = [], {}
        for slot in finished:
            f, j, action = self.slot_owner.pop(slot)
            heapq.heappush(self.free_slots, slot)
            if action[0] == "host":
                host.append((slot, f, j))
            else:
                dev.setdefault((f.fid, action[0]), (f, action[0], []))[2].append((slot, action[1]))
            f.iters += int(iters[slot])
            f.remaining -= 1
            self.n_solves += 1
        t = time.perf_counter()
        X = self.sess.engine_take_batch([s for s, _, _ in host]) if host else None
        for f, what, items in dev.values():          # trace columns: results stay on the device
            tr = f.trace
            slots = [s for s, _ in items]
            colsj = [c for _, c in items]
            if what == "stash":                       # K u -> AU buffer, h = (Sigma^-1 X)^T K u
                _, dots = self.sess.engine_take_batch_ex(slots, X_dev=tr["bufs"]["AU"], xcols=colsj,
                                                         Q=tr["bufs"]["SX"], m=tr["p"])
                tr["h"][colsj] = dots
            else:                                     # Sigma^-1 u: d (and e) only
                _, dots = self.sess.engine_take_batch_ex(
                    slots, P1=tr["bufs"]["AU"], P2=tr["bufs"]["U"] if tr["want_t0"] else None, pcols=colsj)
                tr["d"][colsj] = dots[:, 0]
                if tr["want_t0"]:
                    tr["e"][colsj] = dots[:, 1]
        self.timing["take"] += time.perf_counter() - t
        touched = {f.fid: f for _, f, _ in host}
        touched.update({f.fid: f for f, _, _ in dev.values()})
        resume = []
        for f in touched.values():
            if f.remaining == 0 and not f.pending:
                if f.trace is not None and f.trace["phase"] == 1:
                    continue                          # phase 2 is queued below, not a resume
                resume.append(f)
        phase2 = [f for f in touched.values()
                  if f.remaining == 0 and not f.pending and f.trace is not None and f.trace["phase"] == 1]
        if (resume or phase2) and self.slot_owner:
            # the columns still running take their next iteration meanwhile
            self._inflight = self._pool.submit(self.sess.engine_step, 1, self.width)
            self.n_overlapped += 1
        t = time.perf_counter()
        for c, (slot, f, j) in enumerate(host):       # contiguous column copies (F-order blocks)
            f.result_buf[:, j] = X[:, c] if f.subset is None else X[f.subset, c]
        self.timing["copy"] += time.perf_counter() - t
        for f in phase2:
            self._trace_phase2(f)
        t = time.perf_counter()
        self._advance([(f, self._trace_result(f) if f.trace is not None else f.result_buf) for f in resume])
        self.timing["host"] += time.perf_counter() - t


def fit_plain(sess, trait, y, X, opt=None, probe_seed=200, **kw):
    """One null fit through the plain single-RHS path."""
    opt = opt or Options()
    probes = RademacherProbes.shared(len(y), probe_seed)
    ex = PlainExecutor(sess, opt)
    kw.setdefault("lib", True)
    res = ex.run(fit_co(trait, y, X, probes, opt, **kw))
    res["n_solves"] = ex.n_solves
    res["pcg_iters"] = ex.pcg_iters
    res["n_products"] = ex.n_products
    return res


def variance_ratio(sess, fit, y, X, trait, categories, opt=None, vr_pool=True, num_markers=30,
                   ratio_cv_cutoff=0.001, width=64):
    """Run variance_ratio_co through the batch scheduler (all markers of a
    round in one block PCG).  vr_pool: candidates index the markers held
    out of the GRM (Get_OneSNP_Geno_forVarRatio) rather than GRM markers."""
    opt = opt or Options()
    get = sess.Get_OneSNP_Geno_forVarRatio if vr_pool else sess.Get_OneSNP_Geno
    with BatchScheduler(sess, opt, width=width, n_systems=1) as sch:
        h = sch.submit(variance_ratio_co(fit, y, X, trait, get, categories, opt, num_markers,
                                         ratio_cv_cutoff))
        sch.run()
    if h.error is not None:
        raise h.error
    return h.result


def loco(sess, jobs, starts, ends, opt=None, batched=True, width=64, n_systems=16, timing=None):
    """fitNULLGLMM's LOCO loop over chromosomes 1..22 for many fitted nulls.
    jobs: dicts with fit (a fit result), trait, y, X and optionally subset.
    starts/ends: update_chr_start_end_index's ranges, as indices into the
    GRM's QC'd marker list.  Per chromosome: setStartEndIndex builds the
    leave-out GRM on the device once, then every job's Get_Coef_LOCO runs,
    batched through the scheduler on the LOCO GRM or one solve at a time.
    Returns, per job, a list of 22 results ({"isLOCO": False} where the
    chromosome is absent).  timing: optional Counter, gets the wall seconds
    of the range builds ("set_range") and of the fits ("coef")."""
    opt = opt or Options()
    timing = collections.Counter() if timing is None else timing
    out = [[] for _ in jobs]
    for j in range(22):
        if starts[j] is None or ends[j] is None:
            for r in out:
                r.append({"isLOCO": False})
            continue
        t = time.perf_counter()
        sess.setStartEndIndex(starts[j], ends[j], j)
        timing["set_range"] += time.perf_counter() - t
        t = time.perf_counter()
        if batched:
            sess.engine_use_loco(True)
            try:
                with BatchScheduler(sess, opt, width=width, n_systems=n_systems) as sch:
                    hs = [sch.submit(loco_co(jb["fit"], jb["y"], jb["X"], jb["trait"], opt),
                                     jb.get("subset")) for jb in jobs]
                    sch.run()
            finally:
                sess.engine_use_loco(False)
            for r, h in zip(out, hs):
                if h.error is not None:
                    raise h.error
                r.append(h.result)
        else:
            if any(jb.get("subset") is not None for jb in jobs):
                raise ValueError("sample subsets need the batched path")
            ex = PlainExecutor(sess, opt, loco=True)
            for r, jb in zip(out, jobs):
                r.append(ex.run(loco_co(jb["fit"], jb["y"], jb["X"], jb["trait"], opt)))
        timing["coef"] += time.perf_counter() - t
    return out


def fit_batched(sess, jobs, opt=None, width=64, n_systems=16, probe_seed=200, lib=True):
    """Many null fits through the batching scheduler.  jobs: list of dicts
    with keys trait, y, X and optionally subset (sample indices), offset
    (the covariate offset of isCovariateOffset = TRUE, FG.R:1055-1065) and
    tag.  lib as in fit_binary_co."""
    opt = opt or Options()
    sch = BatchScheduler(sess, opt, width=width, n_systems=n_systems)
    # the glm() start values (host IRLS, no GRM) of all fits
    fams = {"binary": Binomial, "quantitative": Gaussian}

    start = glm_start_library if lib else glm_start

    # the inputs converted once here (one column-major copy per distinct
    # covariate matrix in lib mode), so the drivers' own conversions are
    # no-ops
    conv = {}

    def covariates(X):
        if id(X) not in conv:
            conv[id(X)] = (X, np.asfortranarray(X, np.float64) if lib else np.asarray(X, np.float64))
        return conv[id(X)][1]

    def offset(j):
        o = j.get("offset")
        return None if o is None else np.ascontiguousarray(o, np.float64)

    args = [(np.ascontiguousarray(j["y"], np.float64), covariates(j["X"]), fams[j["trait"]], offset(j))
            for j in jobs]
    # one after another on this thread: on the GPU box's host, four library
    # starts on worker threads reached the first GPU step 3 ms later
    # (21.8 ms against 18.7 ms, medians of four rounds)
    fit0s = [start(*a) for a in args]
    handles = []
    for j, a, fit0 in zip(jobs, args, fit0s):
        n = len(j["y"])
        probes = RademacherProbes.shared(n, probe_seed)
        gen = fit_co(j["trait"], a[0], a[1], probes, opt, offset=a[3], fit0=fit0, lib=lib)
        handles.append(sch.submit(gen, j.get("subset"), j.get("tag")))
    try:
        sch.run()
    finally:
        sch.close()
    return handles, sch
