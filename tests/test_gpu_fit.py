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


@pytest.mark.parametrize("mode", ["bf16x3", "bf16x3_tri"])
def test_batched_split_modes(session, cohort, mode):
    """A batched round (four phenotypes and a 1100-sample subset) with the
    block GEMM in a split mode: tau within 1e-4 of the same round on the
    FP32 MFMA and within the north-star bound of the FP64 oracle, the
    subset's columns masked as in FP32 mode, and a fit's bits the same alone
    and in the mix."""
    c, grm, X, phenos = cohort
    print(f"case: batched fits, GEMM {mode}", flush=True)
    opt = D.Options()
    rng = np.random.default_rng(4)
    sub = np.sort(rng.choice(c["N"], 1100, replace=False))
    Xs, _, _ = plink.covariate_transform(c["X"][sub])
    jobs = [{"trait": t, "y": y, "X": X, "tag": k} for k, (t, y) in phenos.items()]
    jobs.append({"trait": "binary", "y": phenos["b20"][1][sub], "X": Xs, "subset": sub, "tag": "b20sub"})
    ref, _ = D.fit_batched(session, jobs, opt)
    session.set_gemm_mode(mode)
    try:
        session.autotune(reps=1)
        got, _ = D.fit_batched(session, jobs, opt)
        alone, _ = D.fit_batched(session, jobs[-1:], opt)
    finally:
        session.set_gemm_mode("fp32")
        session.set_gemm_config(1, 1, 4)
    for g, r in zip(got, ref):
        assert g.error is None, (g.tag, g.error)
        print(f"  {g.tag}: {mode} theta {g.result['theta']} fp32 {r.result['theta']}", flush=True)
        assert np.allclose(g.result["theta"], r.result["theta"], rtol=1e-4, atol=1e-7), g.tag
    for name in ("b20", "q40"):
        i = list(phenos).index(name)
        o = _oracle_fit(grm, phenos[name][0], phenos[name][1], X, opt)
        assert np.allclose(got[i].result["theta"], o["theta"], rtol=1e-3, atol=1e-6), name
    assert np.array_equal(alone[0].result["theta"], got[-1].result["theta"])
    assert np.array_equal(alone[0].result["coefficients"], got[-1].result["coefficients"])


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


def test_native_driver_batch(session, cohort):
    """saige_fit_null_batch (the C++ drivers and scheduler) on a mix of
    binary and quantitative fits plus a sample subset: same tau and
    iteration count as the Python batch scheduler (both run the same
    engine columns; host FP64 arithmetic differs in summation order only),
    within the north-star bound of the FP64 oracle, and each fit's bits the
    same alone and in the mix."""
    c, grm, X, phenos = cohort
    print("case: native C++ batched fits", flush=True)
    opt = D.Options()
    rng = np.random.default_rng(4)
    sub = np.sort(rng.choice(c["N"], 1100, replace=False))
    Xs, _, _ = plink.covariate_transform(c["X"][sub])
    jobs = [{"trait": t, "y": y, "X": X, "tag": k} for k, (t, y) in phenos.items()]
    jobs.append({"trait": "binary", "y": phenos["b20"][1][sub], "X": Xs, "subset": sub, "tag": "b20sub"})
    nat = session.fit_null_batch(jobs, opt)
    assert all(r["status"] == 0 for r in nat), [r["status"] for r in nat]
    handles, _ = D.fit_batched(session, jobs, opt)
    for r, h, j in zip(nat, handles, jobs):
        ref = h.result
        print(f"  {j['tag']}: native theta {r['theta']} python {ref['theta']} iters {r['iterations']}", flush=True)
        assert r["iterations"] == ref["iterations"], j["tag"]
        assert np.allclose(r["theta"], ref["theta"], rtol=1e-5, atol=1e-8), j["tag"]
        assert np.allclose(r["coefficients"], ref["coefficients"], rtol=1e-5, atol=1e-7), j["tag"]
        assert np.allclose(r["fitted_values"], ref["fitted_values"], rtol=1e-5, atol=1e-8), j["tag"]
    for name in ("b20", "q40"):
        i = list(phenos).index(name)
        ref = _oracle_fit(grm, phenos[name][0], phenos[name][1], X, opt)
        assert np.allclose(nat[i]["theta"], ref["theta"], rtol=1e-3, atol=1e-6), name
    assert nat[0]["theta"][1] > 0
    alone = session.fit_null_batch(jobs[:1], opt)
    assert np.array_equal(alone[0]["theta"], nat[0]["theta"])
    assert np.array_equal(alone[0]["coefficients"], nat[0]["coefficients"])


def test_server_ticket_polled_from_two_threads(session, cohort):
    """saige_poll of one ticket from two threads at once: one collects the
    result (the same fit as saige_fit_null_batch gives), the other is told
    the ticket is gone, and the server stops cleanly."""
    import ctypes
    import threading
    from saige_gpu import _lib, api
    c, grm, X, phenos = cohort
    print("case: server, one ticket polled from two threads", flush=True)
    opt = D.Options()
    job = {"trait": "binary", "y": phenos["b20"][1], "X": X}
    ref = session.fit_null_batch([job], opt)[0]
    L = session._L
    with session.server(opt, max_concurrent=4, hbm_reserve_bytes=0) as srv:
        ticket = srv.submit(job)
        rcs, outs = [None, None], [None, None]

        def poll(i):
            res = _lib.FitResult()
            keep = api._fit_result(job, res)
            rcs[i] = L.saige_poll(srv.h, ctypes.c_int64(ticket), -1, ctypes.byref(res), None)
            outs[i] = (res, keep)

        th = [threading.Thread(target=poll, args=(i,)) for i in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=120)
        assert not any(t.is_alive() for t in th)
        print(f"  poll return codes {rcs}", flush=True)
        assert sorted(rcs) == [_lib.SAIGE_E_ARG, _lib.SAIGE_OK]
        i = rcs.index(_lib.SAIGE_OK)
        got = api._result_dict(outs[i][0], outs[i][1], "binary")
        assert np.array_equal(got["theta"], ref["theta"])
        assert np.array_equal(got["coefficients"], ref["coefficients"])
        srv._jobs.pop(ticket, None)


def test_variance_ratio_vs_oracle(session):
    """Null fit + variance ratio (FG.R:2152-2423) with variance-ratio markers
    held out of the GRM (FG.cpp:545-547), against the FP64 oracle on the
    same held-out genotypes and the same null model."""
    print("case: variance ratio", flush=True)
    c = synth.small_cohort(1200, 3000, maf_lo=0.02, maf_hi=0.5, miss_rate=0.005, seed=synth.config_seed(4),
                           ncausal=300)
    rng = np.random.default_rng(11)
    pool = np.sort(rng.choice(c["M"], 120, replace=False)).astype(np.int32)
    flags = np.zeros(c["M"], np.uint8)
    flags[pool] = 1
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setminMAC_VarianceRatio(20, -1, True)
    session.set_vr_random_markers(pool)
    try:
        session.setgeno_bytes(c["bed"], c["N"], c["M"])
        session.build_grm("fp32")
        og = O.setgeno(c["bed"], c["N"], c["M"], min_maf=0.01, max_missing=0.15, vr_mode=2, min_mac_vr=20,
                       randset=flags)
        assert np.array_equal(session.getIndexVec_forVarRatio(), og["index_vr"])
        grm = O.DenseGRM(O.grm_f64(og), og["Mqc"], O.diag_std(og))
        X, _, _ = plink.covariate_transform(c["X"])
        y = synth.binary_phenotype(c["X"], synth.polygenic(c["Zc"], 2.0, 5), 0.25, 5)
        opt = D.Options()
        fit = D.fit_plain(session, "binary", y, X, opt)
        nvr = len(og["index_vr"])
        order = rng.permutation(nvr)
        cats = [list(order)]
        out = D.variance_ratio(session, fit, y, X, "binary", cats, opt, vr_pool=True, num_markers=30)
        mu = fit["fitted_values"]
        w = mu * (1 - mu)
        Sigma_iX, _ = grm.pcg(w, fit["theta"], X, 500, 1e-5)
        noK = O.score_test_null(mu, mu * (1 - mu), y, X)
        ref = [O.variance_ratio_marker(grm, O.get_geno(og["packed_vr"][i], c["N"]), X, y, mu, w, fit["theta"],
                                       Sigma_iX, noK["XV"], noK["XXVX_inv"], "binary", 500, 1e-5)
               for i in order[:len(out["ratios"][0])]]
        print("  ratio", out["table"][0][0], "oracle", np.mean(ref), "markers", len(ref))
        assert len(ref) >= 30
        assert np.allclose(out["ratios"][0], ref, rtol=2e-3)
        assert 0.5 < out["table"][0][0] < 1.5
    finally:
        session.setminMAC_VarianceRatio(20, -1, False)
        session.set_vr_random_markers(np.zeros(0, np.int32))
