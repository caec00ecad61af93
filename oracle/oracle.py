"""CPU ORACLE — test infrastructure only.

Python side of the oracle for SAIGE step 1's hot path.  Two layers:

* ctypes bindings to liboracle.so (saige_oracle.c), the FP32 restatement of
  the genotype decode/QC/standardisation, GRM matvec and PCG of
  src/SAIGE/src/SAIGE_fitGLMM_fast.cpp (FG.cpp);
* an FP64 numpy restatement of the solver-level functions and of the R
  drivers in src/SAIGE/R/SAIGE_fitGLMM_fast.R (FG.R), evaluated with a dense
  FP64 GRM: getCoefficients, GetTrace[_q], getAIScore[_q],
  fitglmmaiRPCG[_q], Get_Coef, glmmkin.ai_PCG_Rcpp_{Binary,Quantitative},
  the variance-ratio core of extractVarianceRatio, and the leave-one-
  chromosome-out GRM and loop (setStartEndIndex, Get_Coef_LOCO).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product path never does.

Pins (tests/test_oracle_*.py): decode against the reference's .acount/.frq
fixtures and a brute-force PLINK decoder; every FP64 solver function against
dense linear algebra (numpy.linalg.solve / explicit projection matrix P), the
tau update against its closed form, and the glm start values against the
logistic score equations.  R's RNG streams (probe vectors, marker sampling)
cannot be reproduced without R: probes are inputs here, shared with the
device path ("parity unpinned" for the reference's own random draws).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class MarkerStat(ctypes.Structure):
    _fields_ = [
        ("allele_count", ctypes.c_int32),
        ("num_missing", ctypes.c_int32),
        ("mac", ctypes.c_int32),
        ("fill", ctypes.c_int32),
        ("freq", ctypes.c_float),
        ("missing_rate", ctypes.c_float),
        ("inv_std", ctypes.c_float),
        ("pass_qc", ctypes.c_uint8),
        ("pass_vr", ctypes.c_uint8),
        ("pad", ctypes.c_uint8 * 2),
    ]


STAT_DTYPE = np.dtype([
    ("allele_count", np.int32), ("num_missing", np.int32), ("mac", np.int32),
    ("fill", np.int32), ("freq", np.float32), ("missing_rate", np.float32),
    ("inv_std", np.float32), ("pass_qc", np.uint8), ("pass_vr", np.uint8),
    ("pad", np.uint8, (2,)),
])
assert STAT_DTYPE.itemsize == ctypes.sizeof(MarkerStat)


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ------------------------------------------------------------------ genotypes

def setgeno(bed_body, N, M, subidx=None, indicator=None, min_maf=0.0, max_missing=1.0,
            vr_mode=0, min_mac_vr=20.0, max_mac_vr=-1.0, randset=None):
    """genoClass::setGenoObj (FG.cpp:739-1024) over an in-memory .bed body.

    subidx: 1-based sample indices (subSampleInGeno); indicator: N flags.
    Returns a dict with per-marker stats (all M), the QC'd arrays in marker
    order (freq, inv_std, mac) and the repacked genotypes."""
    L = lib()
    bed_body = np.ascontiguousarray(bed_body, dtype=np.uint8)
    if subidx is None:
        subidx = np.arange(1, N + 1, dtype=np.int32)
    subidx = np.ascontiguousarray(subidx, dtype=np.int32)
    if indicator is None:
        indicator = np.zeros(N, np.uint8)
        indicator[subidx - 1] = 1
    indicator = np.ascontiguousarray(indicator, dtype=np.uint8)
    Nnm = len(subidx)
    bpm_nm = (Nnm + 3) // 4
    stats = np.zeros(M, STAT_DTYPE)
    packed = np.zeros((M, bpm_nm), np.uint8)
    packed_vr = np.zeros((M, bpm_nm), np.uint8)
    rs = None if randset is None else np.ascontiguousarray(randset, dtype=np.uint8)
    nvr = ctypes.c_int64(0)
    L.orc_setgeno.restype = ctypes.c_int64
    mqc = L.orc_setgeno(_p(bed_body), ctypes.c_int64(N), ctypes.c_int64(M), _p(indicator),
                        _p(subidx), ctypes.c_int64(Nnm), ctypes.c_float(min_maf),
                        ctypes.c_float(max_missing), ctypes.c_int(vr_mode),
                        ctypes.c_float(min_mac_vr), ctypes.c_float(max_mac_vr), _p(rs),
                        _p(stats), _p(packed), _p(packed_vr), ctypes.byref(nvr))
    qc = stats["pass_qc"].astype(bool)
    vr = stats["pass_vr"].astype(bool)
    return {
        "N": N, "M": M, "Nnm": Nnm, "Mqc": int(mqc), "stats": stats,
        "qc_mask": qc, "vr_mask": vr,
        "freq": stats["freq"][qc].copy(), "inv_std": stats["inv_std"][qc].copy(),
        "mac": stats["mac"][qc].copy(),
        "packed": packed[: int(mqc)].copy(),
        "freq_vr": stats["freq"][vr].copy(), "inv_std_vr": stats["inv_std"][vr].copy(),
        "mac_vr": stats["mac"][vr].copy(), "index_vr": np.nonzero(vr)[0].astype(np.int32),
        "packed_vr": packed_vr[: int(nvr.value)].copy(),
    }


def std_lookup(freq, inv_std):
    out = np.zeros(3, np.float32)
    lib().orc_std_lookup(ctypes.c_float(freq), ctypes.c_float(inv_std), _p(out))
    return out


def get_geno(packed_row, Nnm):
    """Get_OneSNP_Geno (FG.cpp:241-281) of one repacked marker."""
    out = np.zeros(Nnm, np.int32)
    lib().orc_get_geno(_p(np.ascontiguousarray(packed_row)), ctypes.c_int64(Nnm), _p(out))
    return out


def std_geno(packed_row, Nnm, freq, inv_std):
    out = np.zeros(Nnm, np.float32)
    lib().orc_std_geno(_p(np.ascontiguousarray(packed_row)), ctypes.c_int64(Nnm),
                       ctypes.c_float(freq), ctypes.c_float(inv_std), _p(out))
    return out


def std_matrix(g):
    """Z (Nnm x Mqc) FP32, column j = Get_OneSNP_StdGeno(j)."""
    Z = np.zeros((g["Mqc"], g["Nnm"]), np.float32)
    lib().orc_std_matrix(_p(g["packed"]), ctypes.c_int64(g["Mqc"]), ctypes.c_int64(g["Nnm"]),
                         _p(g["freq"]), _p(g["inv_std"]), _p(Z))
    return Z.T


def diag_std(g):
    out = np.zeros(g["Nnm"], np.float32)
    lib().orc_diag_std(_p(g["packed"]), ctypes.c_int64(g["Mqc"]), ctypes.c_int64(g["Nnm"]),
                       _p(g["freq"]), _p(g["inv_std"]), _p(out))
    return out


def crossprod(g, b, nthreads=1):
    """parallelCrossProd (FG.cpp:1576-1598, 1669-1708): K b, FP32."""
    b = np.ascontiguousarray(b, dtype=np.float32)
    out = np.zeros(g["Nnm"], np.float32)
    lib().orc_crossprod(_p(g["packed"]), ctypes.c_int64(g["Mqc"]), ctypes.c_int64(g["Nnm"]),
                        _p(g["freq"]), _p(g["inv_std"]), _p(b), _p(out), ctypes.c_int(nthreads))
    return out


def crossprod_raw(packed, Mqc, Nnm, freq, inv_std, b, nthreads=1):
    """Same as crossprod() on explicit arrays (used for the CPU baseline)."""
    out = np.zeros(Nnm, np.float32)
    lib().orc_crossprod(_p(packed), ctypes.c_int64(Mqc), ctypes.c_int64(Nnm), _p(freq),
                        _p(inv_std), _p(np.ascontiguousarray(b, dtype=np.float32)), _p(out),
                        ctypes.c_int(nthreads))
    return out


def grm_f64(g):
    """Unscaled S = Z Z^T in FP64 (K = S / Mqc)."""
    S = np.zeros((g["Nnm"], g["Nnm"]), np.float64)
    lib().orc_grm_f64(_p(g["packed"]), ctypes.c_int64(g["Mqc"]), ctypes.c_int64(g["Nnm"]),
                      _p(g["freq"]), _p(g["inv_std"]), _p(S))
    return S


def grm_f64_blas(g):
    """grm_f64 by BLAS: the same FP32 standardised values (std_matrix)
    promoted to FP64, S = Z Z^T summed by the BLAS in its own order (equal to
    grm_f64 to FP64 rounding; seconds instead of minutes at N = 2000)."""
    Z = std_matrix(g).astype(np.float64)
    return Z @ Z.T


def marker_subset(g, idx):
    """The genotype dict restricted to QC'd markers idx (in that order)."""
    idx = np.asarray(idx, np.int64)
    return dict(g, packed=np.ascontiguousarray(g["packed"][idx]), freq=g["freq"][idx].copy(),
                inv_std=g["inv_std"][idx].copy(), mac=g["mac"][idx].copy(), Mqc=len(idx))


def loco_diag_std(g, start, end, diag=None):
    """set_Diagof_StdGeno_LOCO (FG.cpp:4934-4958) for one range: the range's
    sum of z^2 accumulated from zero in marker order (FP32), subtracted from
    the full diagonal.  Returns (diag_loco, count)."""
    full = diag_std(g) if diag is None else np.asarray(diag, np.float32)
    part = diag_std(marker_subset(g, np.arange(start, end + 1)))
    return (full - part).astype(np.float32), end - start + 1


def loco_grm(g, start, end):
    """The leave-out GRM of setStartEndIndex (FG.cpp:3067-3082) with
    parallelCrossProd_LOCO (FG.cpp:1790-1851) and getDiagOfSigma_LOCO
    (FG.cpp:2368-2392): K_loco = Z_out Z_out^T / (Mqc - (end-start+1)), the
    FP64 product formed from the markers outside [start, end] directly (the
    reference subtracts the range's product from the full one; the two are
    equal in exact arithmetic), diagonal from loco_diag_std."""
    out = np.r_[np.arange(0, start), np.arange(end + 1, g["Mqc"])]
    S = grm_f64(marker_subset(g, out))
    dl, cnt = loco_diag_std(g, start, end)
    return DenseGRM(S, g["Mqc"] - cnt, diag_std=dl)


def diag_sigma_f32(diag, Mqc, w, tau, diag_one=False):
    """getDiagOfSigma (FG.cpp:2322-2365), FP32."""
    N = len(w)
    out = np.zeros(N, np.float32)
    lib().orc_diag_sigma(_p(np.ascontiguousarray(diag, np.float32)), ctypes.c_int64(N),
                         ctypes.c_int64(Mqc), ctypes.c_int(int(diag_one)),
                         _p(np.ascontiguousarray(w, np.float32)), ctypes.c_float(tau[0]),
                         ctypes.c_float(tau[1]), _p(out))
    return out


def sigma_mul_f32(b, kb, w, tau):
    """getCrossprod (FG.cpp:2397-2425) given kb = K b, FP32."""
    N = len(b)
    out = np.zeros(N, np.float32)
    lib().orc_sigma_mul(_p(np.ascontiguousarray(b, np.float32)),
                        _p(np.ascontiguousarray(kb, np.float32)),
                        _p(np.ascontiguousarray(w, np.float32)), ctypes.c_int64(N),
                        ctypes.c_float(tau[0]), ctypes.c_float(tau[1]), _p(out))
    return out


def pcg_f32(g, diag, w, tau, b, maxiter=500, tol=1e-5, Kdense=None, diag_one=False, nthreads=1):
    """getPCG1ofSigmaAndVector (FG.cpp:2593-2809), FP32 throughout.  With
    Kdense (FP32, already / Mqc) the GRM product is the dense one; else the
    per-marker product of parallelCrossProdOpenMP on nthreads threads."""
    N = len(b)
    x = np.zeros(N, np.float32)
    Kd = None if Kdense is None else np.ascontiguousarray(Kdense, np.float32)
    it = lib().orc_pcg(_p(g["packed"]) if Kd is None else None, ctypes.c_int64(g["Mqc"]),
                       _p(g["freq"]), _p(g["inv_std"]), _p(Kd),
                       _p(np.ascontiguousarray(diag, np.float32)), ctypes.c_int(int(diag_one)),
                       ctypes.c_int64(N), _p(np.ascontiguousarray(w, np.float32)),
                       ctypes.c_float(tau[0]), ctypes.c_float(tau[1]),
                       _p(np.ascontiguousarray(b, np.float32)), ctypes.c_int(maxiter),
                       ctypes.c_float(tol), _p(x), ctypes.c_int(nthreads))
    return x, it


# -------------------------------------------------------------- FP64 layer

class DenseGRM:
    """Dense FP64 GRM K = S / Mqc with its own diagonal (diagStd / Mqc or 1)."""

    def __init__(self, S, Mqc, diag_std=None, diag_one=False):
        self.K = np.asarray(S, np.float64) / float(Mqc)
        self.N = self.K.shape[0]
        if diag_one:
            self.kdiag = np.ones(self.N)
        elif diag_std is not None:
            self.kdiag = np.asarray(diag_std, np.float64) / float(Mqc)
        else:
            self.kdiag = np.diag(self.K).copy()

    def matvec(self, b):
        return self.K @ b

    def sigma(self, w, tau):
        """Dense Sigma = tau0 diag(1/w) + tau1 K (FG.cpp:2320 comment)."""
        return tau[0] * np.diag(1.0 / w) + tau[1] * self.K

    def sigma_mul(self, b, w, tau):
        """getCrossprod (FG.cpp:2397-2425) in FP64."""
        if tau[1] == 0:
            return tau[0] * (b / w[:, None] if b.ndim == 2 else b / w)
        wv = w[:, None] if b.ndim == 2 else w
        return tau[0] * (b / wv) + tau[1] * (self.K @ b)

    def diag_sigma(self, w, tau):
        """getDiagOfSigma (FG.cpp:2322-2365) in FP64."""
        d = tau[1] * self.kdiag + tau[0] / w
        return np.maximum(d, 1e-4)

    def pcg(self, w, tau, b, maxiter=500, tol=1e-5):
        """getPCG1ofSigmaAndVector (FG.cpp:2593-2809) in FP64.  b may be N x k
        (columns solved independently with the same stopping rule)."""
        w = np.asarray(w, np.float64)
        B = np.asarray(b, np.float64)
        single = B.ndim == 1
        if single:
            B = B[:, None]
        minv = 1.0 / self.diag_sigma(w, tau)
        X = np.zeros_like(B)
        iters = np.zeros(B.shape[1], np.int64)
        for c in range(B.shape[1]):
            x = np.zeros(self.N)
            r = B[:, c].copy()
            z = minv * r
            p = z.copy()
            sumr2 = r @ r
            it = 0
            while sumr2 > tol and it < maxiter:
                it += 1
                Ap = self.sigma_mul(p, w, tau)
                a = (r @ z) / (p @ Ap)
                x = x + a * p
                r1 = r - a * Ap
                z1 = minv * r1
                bet = (z1 @ r1) / (z @ r)
                p = z1 + bet * p
                z, r = z1, r1
                sumr2 = r @ r
            X[:, c] = x
            iters[c] = it
        return (X[:, 0], int(iters[0])) if single else (X, iters)


def cal_cv(x):
    """calCV (FG.cpp:3104-3110): (sd/mean)/n with the n-1 standard deviation."""
    x = np.asarray(x, np.float64)
    return (np.std(x, ddof=1) / np.mean(x)) / len(x)


def get_coefficients(grm, Y, X, w, tau, maxiter=500, tol=1e-5):
    """getCoefficients (FG.cpp:3167-3199)."""
    Sigma_iY, _ = grm.pcg(w, tau, Y, maxiter, tol)
    Sigma_iX, _ = grm.pcg(w, tau, X, maxiter, tol)
    XtSX = X.T @ Sigma_iX
    XtSX = np.triu(XtSX) + np.triu(XtSX, 1).T  # symmatu
    try:
        np.linalg.cholesky(XtSX)  # inv_sympd succeeds only on a positive-definite matrix
        cov = np.linalg.inv(XtSX)
    except np.linalg.LinAlgError:
        cov = np.linalg.pinv(XtSX)
    alpha = cov @ (Sigma_iX.T @ Y)
    eta = Y - tau[0] * (Sigma_iY - Sigma_iX @ alpha) / w
    return {"Sigma_iY": Sigma_iY, "Sigma_iX": Sigma_iX, "cov": cov, "alpha": alpha, "eta": eta}


def _project(v, Sigma_iv, Sigma_iX, cov):
    """P v = Sigma^-1 v - Sigma^-1 X cov (Sigma^-1 X)^T v (FG.cpp:3139, 3285)."""
    return Sigma_iv - Sigma_iX @ (cov @ (Sigma_iX.T @ v))


def get_trace(grm, Sigma_iX, X, w, tau, cov, nrun, probes, maxiter=500, tol=1e-5,
              cv_cutoff=0.0025, quantitative=False):
    """GetTrace (FG.cpp:3113-3160) / GetTrace_q (FG.cpp:3409-3472).

    probes(i) returns the i-th Rademacher vector (the reference draws them from
    R's rbinom after comm.set.seed(200) on every call, FG.cpp:3114, 3133-3137)."""
    t1, t0 = [], []
    start, end = 0, nrun
    while True:
        for i in range(start, end):
            u = np.asarray(probes(i), np.float64)
            Sigma_iu, _ = grm.pcg(w, tau, u, maxiter, tol)
            Pu = _project(u, Sigma_iu, Sigma_iX, cov)
            Au = grm.matvec(u)
            t1.append(Au @ Pu)
            t0.append(u @ Pu)
        cv = cal_cv(t1)
        cv0 = cal_cv(t0) if quantitative else 0.0
        if cv > cv_cutoff or (quantitative and cv0 > cv_cutoff):
            start, end = end, end + 10
            continue
        break
    if quantitative:
        return np.array([np.mean(t0), np.mean(t1)]), len(t1)
    return float(np.mean(t1)), len(t1)


def get_ai_score(grm, Y, X, w, tau, Sigma_iY, Sigma_iX, cov, nrun, probes, maxiter=500,
                 tol=1e-5, cv_cutoff=0.0025):
    """getAIScore (FG.cpp:3279-3295)."""
    PY = Sigma_iY - Sigma_iX @ (cov @ (Sigma_iX.T @ Y))
    APY = grm.matvec(PY)
    YPAPY = PY @ APY
    Trace, nprobe = get_trace(grm, Sigma_iX, X, w, tau, cov, nrun, probes, maxiter, tol, cv_cutoff)
    PAPY_1, _ = grm.pcg(w, tau, APY, maxiter, tol)
    PAPY = _project(APY, PAPY_1, Sigma_iX, cov)
    AI = APY @ PAPY
    return {"YPAPY": YPAPY, "Trace": Trace, "PY": PY, "AI": AI, "nprobe": nprobe}


def fitglmmai_rpcg(grm, Y, X, w, tau, Sigma_iY, Sigma_iX, cov, nrun, probes, maxiter=500,
                   tolPCG=1e-5, tol=0.02, cv_cutoff=0.0025):
    """fitglmmaiRPCG (FG.cpp:3302-3341): 1-D AI update of tau[1]."""
    re = get_ai_score(grm, Y, X, w, tau, Sigma_iY, Sigma_iX, cov, nrun, probes, maxiter,
                      tolPCG, cv_cutoff)
    Dtau = (re["YPAPY"] - re["Trace"]) / re["AI"]
    tau0 = np.array(tau, np.float64)
    tau = tau0.copy()
    tau[1] = tau0[1] + Dtau
    tau[tau < tol] = 0.0
    step = 1.0
    while tau[1] < 0.0:
        step *= 0.5
        tau[1] = tau0[1] + step * Dtau
    tau[tau < tol] = 0.0
    return tau, re


def get_ai_score_q(grm, Y, X, w, tau, Sigma_iY, Sigma_iX, cov, nrun, probes, maxiter=500,
                   tol=1e-5, cv_cutoff=0.0025):
    """getAIScore_q (FG.cpp:3479-3531): cov is recomputed from Sigma_iX."""
    XtSX = X.T @ Sigma_iX
    XtSX = np.triu(XtSX) + np.triu(XtSX, 1).T
    try:
        np.linalg.cholesky(XtSX)
        cov1 = np.linalg.inv(XtSX)
    except np.linalg.LinAlgError:
        cov1 = np.linalg.pinv(XtSX)
    PY = Sigma_iY - Sigma_iX @ (cov1 @ (Sigma_iX.T @ Y))
    APY = grm.matvec(PY)
    YPAPY = PY @ APY
    A0PY = PY
    YPA0PY = PY @ A0PY
    Trace, nprobe = get_trace(grm, Sigma_iX, X, w, tau, cov1, nrun, probes, maxiter, tol,
                              cv_cutoff, quantitative=True)
    PA0PY_1, _ = grm.pcg(w, tau, A0PY, maxiter, tol)
    PA0PY = _project(A0PY, PA0PY_1, Sigma_iX, cov1)
    PAPY_1, _ = grm.pcg(w, tau, APY, maxiter, tol)
    PAPY = _project(APY, PAPY_1, Sigma_iX, cov1)
    AI = np.array([[A0PY @ PA0PY, A0PY @ PAPY], [A0PY @ PAPY, APY @ PAPY]])
    return {"YPAPY": YPAPY, "YPA0PY": YPA0PY, "Trace": Trace, "PY": PY, "AI": AI,
            "nprobe": nprobe}


def fitglmmai_rpcg_q(grm, Y, X, w, tau, Sigma_iY, Sigma_iX, cov, nrun, probes, maxiter=500,
                     tolPCG=1e-5, tol=0.02, cv_cutoff=0.0025):
    """fitglmmaiRPCG_q (FG.cpp:3610-3662): 2x2 AI update with GMMAT zeroing."""
    tau = np.array(tau, np.float64)
    zero = tau < tol
    re = get_ai_score_q(grm, Y, X, w, tau, Sigma_iY, Sigma_iX, cov, nrun, probes, maxiter,
                        tolPCG, cv_cutoff)
    score = np.array([re["YPA0PY"] - re["Trace"][0], re["YPAPY"] - re["Trace"][1]])
    Dtau = np.linalg.solve(re["AI"], score)
    tau0 = tau.copy()
    tau = tau0 + Dtau
    tau[zero & (tau < tol)] = 0.0
    step = 1.0
    while tau[0] < 0.0 or tau[1] < 0.0:
        step *= 0.5
        tau = tau0 + step * Dtau
        tau[zero & (tau < tol)] = 0.0
    tau[tau < tol] = 0.0
    return tau, re


# ------------------------------------------------------------ R drivers

def logistic(eta):
    return 1.0 / (1.0 + np.exp(-eta))


_EPS = np.finfo(np.float64).eps


class Binomial:
    """R binomial(link='logit') family functions used by FG.R.  R's
    logit_linkinv takes exp(eta) as DBL_EPSILON below eta = -30 and as
    1/DBL_EPSILON above 30; logit_mu_eta is DBL_EPSILON beyond |eta| > 30."""
    name = "binomial"

    @staticmethod
    def linkinv(eta):
        eta = np.asarray(eta, np.float64)
        t = np.where(eta < -30.0, _EPS, np.where(eta > 30.0, 1.0 / _EPS, np.exp(np.clip(eta, -30.0, 30.0))))
        return t / (1.0 + t)

    @staticmethod
    def mu_eta(eta):
        eta = np.asarray(eta, np.float64)
        t = np.exp(np.clip(eta, -30.0, 30.0))
        return np.where(np.abs(eta) > 30.0, _EPS, t / ((1.0 + t) * (1.0 + t)))

    @staticmethod
    def variance(mu):
        return mu * (1.0 - mu)


class Gaussian:
    """R gaussian(link='identity') family functions."""
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


def glm_fit(y, X, family, offset=None, maxit=25, epsilon=1e-8):
    """R glm.fit IRLS (the fit0 = glm(...) start values, FG.R:1119, 1370):
    deviance-based stopping as in R's glm.control defaults."""
    n = len(y)
    off = np.zeros(n) if offset is None else offset
    if family.name == "binomial":
        mu = (y + 0.5) / 2.0
        eta = np.log(mu / (1 - mu))
    else:
        mu = y.astype(np.float64).copy()
        eta = mu.copy()

    def deviance(mu):
        if family.name == "binomial":
            with np.errstate(divide="ignore", invalid="ignore"):
                t1 = np.where(y > 0, y * np.log(y / mu), 0.0)
                t2 = np.where(y < 1, (1 - y) * np.log((1 - y) / (1 - mu)), 0.0)
            return 2 * np.sum(t1 + t2)
        return np.sum((y - mu) ** 2)

    dev_old = deviance(mu)
    beta = np.zeros(X.shape[1])
    for _ in range(maxit):
        mu_eta = family.mu_eta(eta)
        z = (eta - off) + (y - mu) / mu_eta
        wt = mu_eta ** 2 / family.variance(mu)
        sw = np.sqrt(wt)
        beta, *_ = np.linalg.lstsq(X * sw[:, None], z * sw, rcond=None)
        eta = X @ beta + off
        mu = family.linkinv(eta)
        dev = deviance(mu)
        if abs(dev - dev_old) / (abs(dev) + 0.1) < epsilon:
            break
        dev_old = dev
    return {"coef": beta, "eta": eta, "mu": mu, "offset": off}


def get_coef(grm, y, X, tau, family, alpha0, eta0, offset, maxiterPCG=500, tolPCG=1e-5,
             maxiter=20):
    """Get_Coef (FG.R:2-35): IRLS for the fixed effects at fixed tau."""
    tol_coef = 0.1
    mu = family.linkinv(eta0)
    mu_eta = family.mu_eta(eta0)
    Y = eta0 - offset + (y - mu) / mu_eta
    W = (mu_eta / np.sqrt(family.variance(mu))) ** 2
    for _ in range(maxiter):
        rc = get_coefficients(grm, Y, X, W, tau, maxiterPCG, tolPCG)
        alpha = rc["alpha"]
        eta = rc["eta"] + offset
        mu = family.linkinv(eta)
        mu_eta = family.mu_eta(eta)
        Y = eta - offset + (y - mu) / mu_eta
        W = (mu_eta / np.sqrt(family.variance(mu))) ** 2
        if np.max(np.abs(alpha - alpha0) / (np.abs(alpha) + np.abs(alpha0) + tol_coef)) < tol_coef:
            break
        alpha0 = alpha
    return {"Y": Y, "alpha": alpha, "eta": eta, "W": W, "cov": rc["cov"],
            "Sigma_iY": rc["Sigma_iY"], "Sigma_iX": rc["Sigma_iX"], "mu": mu}


def loco_chromosome(grm_loco, fit, y, X, trait, offset=None, maxiterPCG=500, tolPCG=1e-5, maxiter=20):
    """One chromosome of the LOCO loop of fitNULLGLMM (FG.R:1240-1272 binary,
    FG.R:1486-1512 quantitative): Get_Coef_LOCO (FG.R:42-73) started at the
    full fit's theta, coefficients and linear predictors, then the score-test
    null object with mu2 = mu(1-mu) (binary) or 1/tau[0] (quantitative)."""
    y = np.asarray(y, np.float64)
    offset = np.zeros(len(y)) if offset is None else offset
    fam = Binomial if trait == "binary" else Gaussian
    tau = np.asarray(fit["theta"], np.float64)
    rc = get_coef(grm_loco, y, X, tau, fam, fit["coefficients"], fit["linear_predictors"], offset,
                  maxiterPCG, tolPCG, maxiter)
    mu = rc["mu"]
    mu2 = mu * (1 - mu) if trait == "binary" else np.full(len(y), 1.0 / tau[0])
    return {"coefficients": rc["alpha"], "linear_predictors": rc["eta"], "fitted_values": mu,
            "Y": rc["Y"], "residuals": y - mu, "cov": rc["cov"],
            "obj_noK": score_test_null(mu, mu2, y, X)}


def glmmkin_binary(grm, y, X, probes, offset=None, fit0=None, tau_init=(0.0, 0.0), maxiter=20,
                   tol=0.02, nrun=30, tolPCG=1e-5, maxiterPCG=500, cv_cutoff=0.0025):
    """glmmkin.ai_PCG_Rcpp_Binary (FG.R:79-304) without LOCO."""
    n = len(y)
    offset = np.zeros(n) if offset is None else offset
    fam = Binomial
    if fit0 is None:
        fit0 = glm_fit(y, X, fam, offset)
    eta = fit0["eta"]
    alpha0 = fit0["coef"]
    eta0 = eta
    tau = np.array([1.0, 0.0])
    tau[1] = 0.1 if tau_init[1] == 0 else tau_init[1]
    tau0 = tau.copy()
    rc = get_coef(grm, y, X, tau, fam, alpha0, eta0, offset, maxiterPCG, tolPCG, maxiter)
    re = get_ai_score(grm, rc["Y"], X, rc["W"], tau, rc["Sigma_iY"], rc["Sigma_iX"], rc["cov"],
                      nrun, probes, maxiterPCG, tolPCG, cv_cutoff)
    tau[1] = max(0.0, tau0[1] + tau0[1] ** 2 * (re["YPAPY"] - re["Trace"]) / n)
    history = [tau.copy()]
    it = 0
    for it in range(1, maxiter + 1):
        alpha0 = rc["alpha"]
        tau0 = tau.copy()
        eta0 = eta
        rc = get_coef(grm, y, X, tau, fam, alpha0, eta0, offset, maxiterPCG, tolPCG, maxiter)
        tau, _ = fitglmmai_rpcg(grm, rc["Y"], X, rc["W"], tau, rc["Sigma_iY"], rc["Sigma_iX"],
                                rc["cov"], nrun, probes, maxiterPCG, tolPCG, tol, cv_cutoff)
        eta = rc["eta"]
        history.append(tau.copy())
        if tau[1] == 0:
            break
        if np.max(np.abs(tau - tau0) / (np.abs(tau) + np.abs(tau0) + tol)) < tol:
            break
        if np.max(tau) > tol ** (-2):
            it = maxiter
            break
    rc = get_coef(grm, y, X, tau, fam, rc["alpha"], eta, offset, maxiterPCG, tolPCG, maxiter)
    return {"theta": tau, "coefficients": rc["alpha"], "linear_predictors": rc["eta"],
            "fitted_values": rc["mu"], "Y": rc["Y"], "cov": rc["cov"],
            "converged": it < maxiter, "history": history, "iterations": it}


def glmmkin_quantitative(grm, y, X, probes, offset=None, fit0=None, maxiter=20, tol=0.02, nrun=30,
                         tolPCG=1e-5, maxiterPCG=500, cv_cutoff=0.0025):
    """glmmkin.ai_PCG_Rcpp_Quantitative (FG.R:309-549) without LOCO."""
    n = len(y)
    offset = np.zeros(n) if offset is None else offset
    fam = Gaussian
    if fit0 is None:
        fit0 = glm_fit(y, X, fam, offset)
    eta = fit0["eta"]
    alpha = fit0["coef"]
    alpha0 = alpha
    eta0 = eta
    tau = np.array([1.0, 0.0])
    tau0 = tau.copy()
    rc = get_coef(grm, y, X, tau, fam, alpha0, eta0, offset, maxiterPCG, tolPCG, maxiter)
    re = get_ai_score_q(grm, rc["Y"], X, rc["W"], tau, rc["Sigma_iY"], rc["Sigma_iX"], rc["cov"],
                        nrun, probes, maxiterPCG, tolPCG, cv_cutoff)
    tau[1] = max(0.0, tau0[1] + tau0[1] ** 2 * (re["YPAPY"] - re["Trace"][1]) / n)
    tau[0] = max(0.0, tau0[0] + tau0[0] ** 2 * (re["YPA0PY"] - re["Trace"][0]) / n)
    history = [tau.copy()]
    it = 0
    for it in range(1, maxiter + 1):
        alpha0 = alpha
        tau0 = tau.copy()
        eta0 = eta
        rc = get_coef(grm, y, X, tau, fam, alpha0, eta0, offset, maxiterPCG, tolPCG, maxiter)
        tau, _ = fitglmmai_rpcg_q(grm, rc["Y"], X, rc["W"], tau, rc["Sigma_iY"], rc["Sigma_iX"],
                                  rc["cov"], nrun, probes, maxiterPCG, tolPCG, tol, cv_cutoff)
        alpha = rc["alpha"]
        eta = rc["eta"]
        history.append(tau.copy())
        if tau[0] <= 0:
            raise RuntimeError("The first variance component parameter estimate is 0")
        if tau[1] <= 0:
            break
        if np.max(np.abs(tau - tau0) / (np.abs(tau) + np.abs(tau0) + tol)) < tol:
            break
        if np.max(tau) > tol ** (-2):
            it = maxiter
            break
    rc = get_coef(grm, y, X, tau, fam, alpha, eta, offset, maxiterPCG, tolPCG, maxiter)
    return {"theta": tau, "coefficients": rc["alpha"], "linear_predictors": rc["eta"],
            "fitted_values": rc["mu"], "Y": rc["Y"], "cov": rc["cov"],
            "converged": it < maxiter, "history": history, "iterations": it}


def variance_ratio_marker(grm, G0, X, y, mu, w, tau, Sigma_iX, XV, XXVX_inv, trait="binary",
                          maxiterPCG=500, tolPCG=1e-2):
    """The per-marker core of extractVarianceRatio (FG.R:2309-2368):
    var1 / var2 for one marker's integer genotype vector G0."""
    G0 = np.asarray(G0, np.float64)
    N = len(G0)
    if G0.sum() / (2 * N) > 0.5:
        G0 = 2 - G0
    AC = G0.sum()
    G = G0 - XXVX_inv @ (XV @ G0)
    g = G / np.sqrt(AC)
    Sigma_iG, _ = grm.pcg(w, tau, G, maxiterPCG, tolPCG)
    var1a = G @ Sigma_iG - G @ Sigma_iX @ np.linalg.solve(X.T @ Sigma_iX, X.T @ Sigma_iG)
    var1 = var1a / AC
    if trait == "binary":
        var2 = np.sum(mu * (1 - mu) * g * g)
    else:
        var2 = g @ g
    return var1 / var2


def score_test_null(mu, mu2, y, X):
    """ScoreTest_NULL_Model (FG.R:579-591)."""
    V = mu2
    XV = (X * V[:, None]).T
    XVX = X.T @ XV.T
    XVX_inv = np.linalg.inv(XVX)
    XXVX_inv = X @ XVX_inv
    return {"XV": XV, "XVX": XVX, "XXVX_inv": XXVX_inv, "XVX_inv": XVX_inv,
            "S_a": (X * (y - mu)[:, None]).sum(0), "XVX_inv_XV": XXVX_inv * V[:, None], "V": V}
