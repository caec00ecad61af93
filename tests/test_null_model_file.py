"""The step-2 null-model file: saige_gpu.rdata (R serialization v3, XDR)
read against the reference's own saved models (extdata/output/
example_{binary,quantitative}_fullGRM.rda, copied to tests/golden/), and the
product's modglmm writer (saige_gpu.null_model) checked for the same
structure and a lossless round trip.

Pins: the theta values the reference printed for these runs; the algebraic
relations R used to fill the object (residuals = y - mu, V = mu(1-mu),
ScoreTest_NULL_Model FG.R:579-591), which fail if a matrix were read
transposed or mis-shaped; the product's score_test_null against the
reference's stored obj.noK."""
import os

import numpy as np
import pytest

from saige_gpu import driver as D
from saige_gpu import null_model, rdata

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _v(x):
    return x.value if isinstance(x, rdata.RObj) else x


@pytest.fixture(scope="module")
def ref_models():
    return {t: rdata.load(os.path.join(GOLD, f"example_{t}_fullGRM.rda"))["modglmm"]
            for t in ("binary", "quantitative")}


def test_reader_reproduces_reference_theta(ref_models):
    b, q = ref_models["binary"], ref_models["quantitative"]
    assert np.array_equal(b["theta"], [1.0, 0.3349943459033966])
    assert np.array_equal(q["theta"], [0.2014368176460266, 0.4834878444671631])
    assert list(b)[:10] == ["theta", "coefficients", "linear.predictors", "fitted.values", "Y", "residuals",
                            "cov", "converged", "sampleID", "obj.noK"]
    assert b["traitType"] == ["binary"] and q["traitType"] == ["quantitative"]
    assert len(b["sampleID"]) == 1000 and b["sampleID"][0] == "1a1"
    assert _v(b["obj.noK"]) and b["obj.noK"].attrs["class"] == ["SA_NULL"]


@pytest.mark.parametrize("trait", ["binary", "quantitative"])
def test_reference_model_algebra_and_score_object(ref_models, trait):
    m = ref_models[trait]
    y = _v(m["y"])
    X = _v(m["X"])
    mu = m["fitted.values"][:, 0]
    assert X.shape == (1000, X.shape[1]) and y.shape == (1000,)
    assert np.array_equal(m["residuals"][:, 0], y - mu)
    noK = _v(m["obj.noK"])
    V = noK["V"]
    if trait == "binary":
        assert np.array_equal(V, mu * (1 - mu))
    else:
        assert np.allclose(V, 1.0 / m["theta"][0], rtol=1e-15)
    # the product's ScoreTest_NULL_Model restatement reproduces the stored object
    o = D.score_test_null(mu, V, y, X)
    for k in ("XV", "XVX", "XXVX_inv", "XVX_inv", "S_a", "XVX_inv_XV"):
        ref = _v(noK[k])
        assert ref.shape == o[k].shape, k
        assert np.allclose(o[k], ref, rtol=1e-9, atol=1e-12 * np.max(np.abs(ref))), k
    # LOCO results carry the same relations per chromosome
    for r in m["LOCOResult"]:
        if r["isLOCO"][0] != 1:
            continue
        mj = r["fitted.values"][:, 0]
        assert np.array_equal(r["residuals"][:, 0], y - mj)
        o = D.score_test_null(mj, _v(r["obj.noK"])["V"], y, X)
        assert np.allclose(o["XVX"], _v(_v(r["obj.noK"])["XVX"]), rtol=1e-9)


def test_round_trip_of_reference_model(ref_models):
    """Writing what was read and reading it back loses nothing."""
    m = ref_models["binary"]
    back = rdata.loads(rdata.dumps({"modglmm": m}))["modglmm"]

    def same(a, b):
        if isinstance(a, rdata.RObj):
            assert isinstance(b, rdata.RObj) and list(a.attrs) == list(b.attrs)
            for k in a.attrs:
                same(a.attrs[k], b.attrs[k])
            return same(a.value, b.value)
        if isinstance(a, dict):
            assert isinstance(b, dict) and list(a) == list(b)
            for k in a:
                same(a[k], b[k])
        elif isinstance(a, list):
            assert type(a) is type(b) and len(a) == len(b)
            for x, z in zip(a, b):
                same(x, z)
        elif isinstance(a, np.ndarray):
            assert a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a, b)
        else:
            assert a == b

    same(m, back)


def test_writer_matches_reference_structure(ref_models):
    """The product's modglmm has the reference's fields, order, shapes and
    attributes, and survives a file round trip."""
    ref = ref_models["binary"]
    rng = np.random.default_rng(3)
    N, p = 1000, 3
    X = np.column_stack([np.ones(N), rng.normal(0, 1, (N, 2))])
    y = (rng.random(N) < 0.1).astype(float)
    eta = np.full(N, -2.2)
    mu = 1 / (1 + np.exp(-eta))
    fit = {"theta": np.array([1.0, 0.33]), "coefficients": np.array([-2.2, 0.0, 0.0]), "linear_predictors": eta,
           "fitted_values": mu, "Y": eta + (y - mu) / (mu * (1 - mu)), "cov": np.eye(3) * 0.01, "converged": True}
    loco = [{"isLOCO": True, "coefficients": fit["coefficients"], "linear_predictors": eta, "fitted_values": mu,
             "Y": fit["Y"], "cov": fit["cov"]} for _ in range(21)] + [{"isLOCO": False}]
    m = null_model.modglmm(fit, y, X, "binary", [f"s{i}" for i in range(N)], loco=loco)
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "saige_null_model_test.rda")
    null_model.save_null_model(path, m)
    got = rdata.load(path)["modglmm"]
    os.remove(path)
    assert list(got) == list(ref)
    for k in ref:
        a, b = _v(ref[k]), _v(got[k])
        if isinstance(a, np.ndarray):
            assert a.ndim == b.ndim and a.dtype == b.dtype, k
        if isinstance(ref[k], rdata.RObj):
            assert list(ref[k].attrs) == list(got[k].attrs), k
    assert got["coefficients"].shape == (3, 1) and got["linear.predictors"].shape == (N, 1)
    assert list(_v(got["obj.noK"])) == list(_v(ref["obj.noK"]))
    assert list(got["LOCOResult"][0]) == list(ref["LOCOResult"][0])
    assert list(got["LOCOResult"][21]) == ["isLOCO"] and got["LOCOResult"][21]["isLOCO"][0] == 0
    assert np.array_equal(got["fitted.values"][:, 0], mu) and got["sampleID"][5] == "s5"
    assert np.allclose(_v(_v(got["obj.noK"])["XVX"]), X.T @ (X * (mu * (1 - mu))[:, None]))


def test_variance_ratio_file_format():
    """The reference's variance-ratio file: one "value kind category" row per
    line, values printed to 15 significant digits; the product's writer emits
    the same row format."""
    rows = [ln.split() for ln in open(os.path.join(GOLD, "example_binary_fullGRM.varianceRatio.txt")).read()
            .strip().split("\n")]
    assert [r[1:] for r in rows] == [["sparse", "1"], ["null", "1"], ["sparse", "2"], ["null", "2"]]
    assert [float(r[0]) for r in rows if r[1] == "null"] == [0.945648350425262, 0.939961056366588]
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "saige_vr_test.txt")
    D.write_variance_ratio(path, [(0.945648350425262, "null", 1), (0.939961056366588, "null", 2)])
    mine = [ln.split() for ln in open(path).read().strip().split("\n")]
    os.remove(path)
    assert mine == [r for r in rows if r[1] == "null"]
