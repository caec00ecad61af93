"""Null fits at the two CPU-sized configurations against the FP64 oracle
(oracle/oracle.py: glmmkin_binary, FG.R:79-304), with identical probe
vectors on both sides and the FP32 GRM of the plain method:

* the reference's own quick test (docs/installation.md:122):
  plinkforGRM_1000samples_10kMarkers with pheno_1000samples.txt, binary y,
  covariates x1, x2, minMAFforGRM 0.01, and the step-1 CLI defaults
  (step1_fitNULLGLMM.R:39-117: nrun 30, tolPCG 1e-5, maxiterPCG 500,
  tol 0.02, maxiter 20, traceCVcutoff 0.0025, isCovariateOffset TRUE);
* C0: N = 2,000, M = 10,000, p ~ U(0.05, 0.5), 0.5% missing calls,
  intercept + age + sex, binary trait with logit-scale h2 = 0.3 on 1,000
  causal markers at prevalence 0.2, the same solver settings.

Each fit runs through the three GPU paths an R caller can reach: the plain
single-RHS path (fit_plain: GEMV per PCG iteration), the batched scheduler
and the library's C++ drivers (saige_fit_null_batch).  Tolerances: the same
AI iteration count as the oracle; tau within 1e-3 relative at every
iteration; coefficients within 1e-3.  LOCO and the
variance ratios of the quick test are left out: their ranges and marker
draws come from the .bim before QC and from R's RNG."""
import os

import numpy as np
import pytest

from configs import SEED, covariate_offset
from fixtures import GOLDEN, load_bed, load_pheno
from oracle import oracle as O
from saige_gpu import driver as D
from saige_gpu import plink, synth

pytestmark = pytest.mark.gpu


def _oracle(grm, y, X, offset, opt):
    return O.glmmkin_binary(grm, y, X, D.RademacherProbes(len(y), 200), offset=offset, nrun=opt.nrun,
                            tolPCG=opt.tolPCG, maxiterPCG=opt.maxiterPCG, cv_cutoff=opt.traceCVcutoff,
                            tol=opt.tol, maxiter=opt.maxiter)


def _check(name, res, ref, coef=True):
    print(f"  {name}: theta {res['theta']} oracle {ref['theta']} iterations {res['iterations']} / "
          f"{ref['iterations']}", flush=True)
    assert res["iterations"] == ref["iterations"], name
    if "history" in res:
        assert len(res["history"]) == len(ref["history"]), name
        for a, b in zip(res["history"], ref["history"]):
            assert np.allclose(a, b, rtol=1e-3, atol=1e-6), name
    assert np.allclose(res["theta"], ref["theta"], rtol=1e-3, atol=1e-6), name
    if coef:
        assert np.allclose(res["coefficients"], ref["coefficients"], rtol=1e-3, atol=1e-5), name


def _three_paths(session, y, X, offset, opt):
    plain = D.fit_plain(session, "binary", y, X, opt, offset=offset)
    job = {"trait": "binary", "y": y, "X": X, "offset": offset}
    hs, _ = D.fit_batched(session, [job], opt)
    assert hs[0].error is None, hs[0].error
    nat = session.fit_null_batch([job], opt)[0]
    assert nat["status"] == 0
    return {"plain": plain, "batched": hs[0].result, "native": nat}


@pytest.fixture(scope="module")
def quicktest(session):
    body, ids = load_bed("grm1000x10k", 1000)
    fam = [l.split()[1] for l in open(os.path.join(GOLDEN, "grm1000x10k.fam"))]
    y, x1, x2, iid = load_pheno()
    # subSampleInGeno: the phenotyped samples in .fam order (FG.R:975-976)
    pos = {s: i for i, s in enumerate(iid)}
    order = np.array([pos[s] for s in fam])
    y, x1, x2 = y[order], x1[order], x2[order]
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setminMAC_VarianceRatio(20, -1, False)
    session.setgeno_bytes(body, 1000, len(ids))
    session.build_grm("fp32")
    og = O.setgeno(body, 1000, len(ids), min_maf=0.01, max_missing=0.15)
    assert session.getMsub_MAFge_minMAFtoConstructGRM() == og["Mqc"] == 9650
    grm = O.DenseGRM(O.grm_f64_blas(og), og["Mqc"], O.diag_std(og))
    Xt, _, _ = plink.covariate_transform(np.column_stack([np.ones(1000), x1, x2]))
    X1, off = covariate_offset(y, Xt, "binary")
    return y, Xt, X1, off, grm


def test_quicktest_example_matches_fp64_oracle(session, quicktest):
    y, Xt, X1, off, grm = quicktest
    print("case: reference quick test, isCovariateOffset TRUE", flush=True)
    opt = D.Options()
    ref = _oracle(grm, y, X1, off, opt)
    assert ref["theta"][1] > 0.02
    for name, res in _three_paths(session, y, X1, off, opt).items():
        _check(name, res, ref)
    # and with the covariates estimated (isCovariateOffset FALSE)
    print("case: reference quick test, isCovariateOffset FALSE", flush=True)
    ref = _oracle(grm, y, Xt, None, opt)
    for name, res in _three_paths(session, y, Xt, None, opt).items():
        _check(name, res, ref)


@pytest.fixture(scope="module")
def c0(session):
    N, M = 2000, 10000
    c = synth.small_cohort(N, M, maf_lo=0.05, maf_hi=0.5, miss_rate=0.005, seed=SEED[0], ncausal=1000)
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setminMAC_VarianceRatio(20, -1, False)
    session.setgeno_bytes(c["bed"], N, M)
    session.build_grm("fp32")
    og = O.setgeno(c["bed"], N, M, min_maf=0.01, max_missing=0.15)
    assert session.getMsub_MAFge_minMAFtoConstructGRM() == og["Mqc"]
    grm = O.DenseGRM(O.grm_f64_blas(og), og["Mqc"], O.diag_std(og))
    # the GPU's FP32 GRM against the FP64 one: an FP32 sum of Mqc terms
    K = session.grm_block(0, 0, N, N)
    assert np.max(np.abs(K - grm.K)) <= 4.0 * np.sqrt(og["Mqc"]) * 2.0 ** -24 * np.max(np.abs(grm.K))
    Xt, _, _ = plink.covariate_transform(c["X"])
    y = synth.binary_phenotype(c["X"], synth.polygenic(c["Zc"], 0.3, c["seed"]), 0.2, c["seed"])
    return y, Xt, grm


def test_c0_binary_fit_matches_fp64_oracle(session, c0):
    y, Xt, grm = c0
    opt = D.Options(nrun=30)
    assert 0.17 < y.mean() < 0.23
    print("case: C0, intercept + age + sex estimated", flush=True)
    ref = _oracle(grm, y, Xt, None, opt)
    assert ref["theta"][1] > 0.02
    for name, res in _three_paths(session, y, Xt, None, opt).items():
        _check(name, res, ref)
    print("case: C0, isCovariateOffset TRUE", flush=True)
    X1, off = covariate_offset(y, Xt, "binary")
    ref = _oracle(grm, y, X1, off, opt)
    for name, res in _three_paths(session, y, X1, off, opt).items():
        _check(name, res, ref)
