"""The null-model object fitNULLGLMM saves for step 2 (`modglmm`,
FG.R:239-243 binary, FG.R:495-499 quantitative, FG.R:1195-1297 for the
offset / LOCO fields), built from this package's fit results and written
as an R data file with saige_gpu.rdata.

Field order, matrix shapes and attributes follow the reference's own saved
models (extdata/output/example_*_fullGRM.rda): N x 1 matrices for the
per-sample vectors, a p x 1 coefficient matrix, model-matrix dimnames and
"assign", obj.noK of class "SA_NULL", one LOCOResult entry per autosome."""
import numpy as np

from . import rdata
from .driver import score_test_null


def _col(v):
    return np.asarray(v, np.float64).reshape(-1, 1)


def _noK(mu, mu2, y, X, rows, cols):
    """ScoreTest_NULL_Model (FG.R:579-591) with R's dimnames."""
    o = score_test_null(mu, mu2, y, X)
    V = np.asarray(o["V"], np.float64)
    assign = np.arange(X.shape[1], dtype=np.int32)
    return rdata.RObj({
        "XV": rdata.RObj(np.asarray(o["XV"]), {"dimnames": [cols, rows], "assign": assign}),
        "XVX": rdata.RObj(np.asarray(o["XVX"]), {"dimnames": [cols, cols]}),
        "XXVX_inv": rdata.RObj(np.asarray(o["XXVX_inv"]), {"dimnames": [rows, cols]}),
        "XVX_inv": rdata.RObj(np.asarray(o["XVX_inv"]), {"dimnames": [cols, cols]}),
        "S_a": rdata.RObj(np.asarray(o["S_a"], np.float64), {"names": cols}),
        "XVX_inv_XV": rdata.RObj(np.asarray(o["XVX_inv_XV"]), {"dimnames": [rows, cols]}),
        "V": V,
    }, {"class": rdata.RStr(["SA_NULL"])})


def _mu2(trait, mu, tau):
    return mu * (1 - mu) if trait == "binary" else np.full(len(mu), 1.0 / tau[0])


def modglmm(fit, y, X, trait, sample_ids, colnames=None, loco=None, offset=None,
            is_covariate_offset=False):
    """The list saved as `modglmm`.  fit: a fit result (theta, coefficients,
    linear_predictors, fitted_values, Y, cov, converged); loco: the 22
    per-chromosome results of driver.loco for this fit, or None (LOCO off)."""
    y = np.asarray(y, np.float64)
    X = np.asarray(X, np.float64)
    N, p = X.shape
    rows = rdata.RStr(str(i + 1) for i in range(N))
    cols = rdata.RStr(colnames if colnames is not None else
                      ["(Intercept)"] + [f"x{j}" for j in range(1, p)])
    tau = np.asarray(fit["theta"], np.float64)
    mu = np.asarray(fit["fitted_values"], np.float64)
    noK = _noK(mu, _mu2(trait, mu, tau), y, X, rows, cols)
    out = {
        "theta": tau,
        "coefficients": _col(fit["coefficients"]),
        "linear.predictors": _col(fit["linear_predictors"]),
        "fitted.values": _col(mu),
        "Y": _col(fit["Y"]),
        "residuals": _col(y - mu),
        "cov": np.atleast_2d(np.asarray(fit["cov"], np.float64)),
        "converged": np.array([bool(fit.get("converged", True))]),
        "sampleID": rdata.RStr(str(s) for s in sample_ids),
        "obj.noK": noK,
        "y": rdata.RObj(y, {"names": rows}),
        "X": rdata.RObj(X, {"dimnames": [rows, cols], "assign": np.arange(p, dtype=np.int32)}),
        "traitType": rdata.RStr([trait]),
        "isCovariateOffset": np.array([bool(is_covariate_offset)]),
        "LOCO": np.array([loco is not None]),
    }
    if loco is not None:
        res = []
        for r in loco:
            if not r.get("isLOCO"):
                res.append({"isLOCO": np.array([False])})
                continue
            m = np.asarray(r["fitted_values"], np.float64)
            res.append({
                "isLOCO": np.array([True]),
                "coefficients": _col(r["coefficients"]),
                "linear.predictors": _col(r["linear_predictors"]),
                "fitted.values": _col(m),
                "Y": _col(r["Y"]),
                "residuals": _col(y - m),
                "cov": np.atleast_2d(np.asarray(r["cov"], np.float64)),
                "obj.noK": _noK(m, _mu2(trait, m, tau), y, X, rows, cols),
            })
        out["LOCOResult"] = res
    else:
        out["LOCOResult"] = None
    out["obj.glm.null"] = {"model": rdata.RObj({}, {"class": rdata.RStr(["data.frame"]),
                                                    "row.names": np.zeros(0, np.int32)})}
    out["offset"] = rdata.RObj(_col(np.zeros(N) if offset is None else offset), {"dimnames": [rows, None]})
    out["useSparseGRMtoFitNULL"] = np.array([False])
    return out


def save_null_model(path, model):
    """save(modglmm, file = path) (FG.R:1297)."""
    rdata.save(path, {"modglmm": model})
