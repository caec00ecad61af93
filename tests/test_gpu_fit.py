"""End-to-end null fits on the GPU against the FP64 oracle driver
(FG.R:79-304, 309-549 restated in oracle/oracle.py), with identical probe
vectors on both sides.

Tolerances: tau within 1e-3 relative at every AI
iteration, same iteration count; batched within 1e-3 of plain; a fit's
result in a mix is bit-identical to the same fit alone."""
import numpy as np
import pytest

from oracle import oracle as O
from saige_gpu import driver as D
from saige_gpu import plink, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cohort(session):
    c = synth.small_cohort(1500, 4000, maf_lo=0.05, maf_hi=0.5, miss_rate=0.005, seed=synth.config_seed(0),
                           ncausal=400)
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setminMAC_VarianceRatio(20, -1, False)
    session.setgeno_bytes(c["bed"], c["N"], c["M"])
    session.build_grm("fp32")
    og = O.setgeno(c["bed"], c["N"], c["M"], min_maf=0.01, max_missing=0.15)
    grm = O.DenseGRM(O.grm_f64(og), og["Mqc"], O.diag_std(og))
    X, _, _ = plink.covariate_transform(c["X"])
    g = synth.polygenic(c["Zc"], 2.0, c["seed"])   # logit-scale variance: tau1 well away from 0
    phenos = {
        "b20": ("binary", synth.binary_phenotype(c["X"], g, 0.2, c["seed"])),
        "b10": ("binary", synth.binary_phenotype(c["X"], g, 0.1, c["seed"] + 1)),
        "b30": ("binary", synth.binary_phenotype(c["X"], g, 0.3, c["seed"] + 2)),
        "q40": ("quantitative", synth.quantitative_phenotype(c["X"], synth.polygenic(c["Zc"], 0.4, 9), 0.4,
                                                           c["seed"])),
    }
    return c, grm, X, phenos


def _oracle_fit(grm, trait, y, X, opt):
    probes = D.RademacherProbes(len(y), 200)
    kw = dict(nrun=opt.nrun, tolPCG=opt.tolPCG, maxiterPCG=opt.maxiterPCG, cv_cutoff=opt.traceCVcutoff,
              tol=opt.tol, maxiter=opt.maxiter)
    if trait == "binary":
        return O.glmmkin_binary(grm, y, X, probes, **kw)
    return O.glmmkin_quantitative(grm, y, X, probes, **kw)


@pytest.mark.parametrize("name", ["b20", "q40"])
def test_plain_fit_matches_fp64_oracle(session, cohort, name):
    c, grm, X, phenos = cohort
    trait, y = phenos[name]
    print(f"case: plain fit {name}", flush=True)
    opt = D.Options()
    res = D.fit_plain(session, trait, y, X, opt)
    ref = _oracle_fit(grm, trait, y, X, opt)
    print("  gpu theta", res["theta"], "oracle theta", ref["theta"], "iters", res["iterations"])
    assert res["iterations"] == ref["iterations"]
    for a, b in zip(res["history"], ref["history"]):
        assert np.allclose(a, b, rtol=1e-3, atol=1e-6)
    assert np.allclose(res["theta"], ref["theta"], rtol=1e-3, atol=1e-6)
    assert np.allclose(res["coefficients"], ref["coefficients"], rtol=1e-3, atol=1e-5)
    assert res["theta"][1] > 0


def test_batched_matches_plain_and_is_mix_invariant(session, cohort):
    c, grm, X, phenos = cohort
    print("case: batched fits", flush=True)
    opt = D.Options()
    jobs = [{"trait": t, "y": y, "X": X, "tag": k} for k, (t, y) in phenos.items()]
    handles, sch = D.fit_batched(session, jobs, opt)
    assert all(h.error is None for h in handles), [h.error for h in handles]
    for h in handles:
        trait, y = phenos[h.tag]
        plain = D.fit_plain(session, trait, y, X, opt)
        assert np.allclose(h.result["theta"], plain["theta"], rtol=1e-3, atol=1e-6), h.tag
    alone, _ = D.fit_batched(session, jobs[:1], opt)
    assert np.array_equal(alone[0].result["theta"], handles[0].result["theta"])
    assert np.array_equal(alone[0].result["coefficients"], handles[0].result["coefficients"])
    assert sch.n_iterations > 0 and sch.busy_slot_iters > sch.n_iterations


def test_subset_fit_equals_restricted_problem(session, cohort):
    c, grm, X, phenos = cohort
    print("case: subset fit", flush=True)
    trait, y = phenos["b20"]
    rng = np.random.default_rng(4)
    sub = np.sort(rng.choice(c["N"], 1100, replace=False))
    Xs, _, _ = plink.covariate_transform(c["X"][sub])
    opt = D.Options()
    handles, _ = D.fit_batched(session, [{"trait": trait, "y": y[sub], "X": Xs, "subset": sub}], opt)
    gsub = O.DenseGRM(grm.K[np.ix_(sub, sub)] * 1.0, 1)
    gsub.kdiag = grm.kdiag[sub]
    ref = _oracle_fit(gsub, trait, y[sub], Xs, opt)
    assert np.allclose(handles[0].result["theta"], ref["theta"], rtol=1e-3, atol=1e-6)
