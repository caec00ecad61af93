"""C2 and C3 at their full size on one resident GRM: N = 100,000 samples,
M = 150,000 markers, MAF U(0.01, 0.5), 1% missing calls; the GRM built
from BF16 standardised genotypes on the MFMA (FP32 accumulation, FP32
storage), then rebuilt split (each standardised value as two BF16 pieces,
three MFMA products), which the fits use; the 64-column block PCG in the
default arithmetic (FP32 operands as three BF16 pieces over the GRM's lower
triangle), autotuned.

C2 (precision and batching):
* the GRM of a 2,048-sample principal subblock against FP64 from those
  samples' genotypes over all markers: single BF16 pass within 4e-3
  max|K64|, the split GRM within the FP32 sum bound 4 sqrt(Mqc) 2^-24
  plus 2^-16 (it is stored at 16 significant bits);
* a 64-column block solve (60 Rademacher probes and 4 phenotype working
  vectors) at full N: ||Sigma x - b||^2 <= 1.5 tolPCG in FP64 on the GPU's
  own GRM for a sample of the columns;
* tau against the FP64 oracle: null fits on the subblock's samples (binary
  and quantitative) through the batched scheduler on the resident BF16 GRM,
  against the oracle's driver on the FP64 subblock GRM: tau within 1e-3,
  the same iterations; the subblock's block solves against dense FP64;
* four binary phenotypes (prevalences 0.05-0.3, nrun 15: one 64-column
  block) at full N, and one of them alone: bitwise the same fit.

C3 (serving): requests on random 20-100% sample subsets (70% binary with
prevalence U(0.01, 0.3), 30% quantitative with h2 U(0.1, 0.6)) arrive as a
Poisson stream while others are on the device; each completes, and a
request rerun alone gives bitwise the same fit."""
import numpy as np
import pytest

from configs import (SEED, causal_std, causal_std_device, device_cohort, engine_solve, fp64_grm_subblock,
                     fp64_residual2, grm_offdiag_error, grm_rel_error, subblock_genotypes)
from fixtures import on_16bit_grid
from oracle import oracle as O
from saige_gpu import driver as D
from saige_gpu import plink, synth

pytestmark = pytest.mark.gpu

N, M, N0 = 100_000, 150_000, 2048


@pytest.fixture(scope="module")
def c2():
    from saige_gpu import Session
    print(f"case: C2 cohort N={N} M={M}", flush=True)
    s = Session(0)
    try:
        dev = device_cohort(s, N, M, 0.01, 0.5, 0.01, 0.0, SEED[2])
        s.setminMAFforGRM(0.01)
        s.setmaxMissingRateforGRM(0.15)
        s.setminMAC_VarianceRatio(20, -1, False)
        s.setgeno_device(dev, N, M)
        G, miss = subblock_genotypes(s, dev, N, M, N0)
        Zc = causal_std_device(s, dev, N, M, 1000, SEED[2])
        s.free_device(dev)
        stats, mask = s.marker_stats(), s.getQCdMarkerIndex().astype(bool)
        Mqc = s.getMsub_MAFge_minMAFtoConstructGRM()
        grm_s = s.build_grm("bf16")
        print(f"  BF16 GRM {grm_s:.2f} s ({N * (N + 1) * Mqc / grm_s / 1e12:.0f} TFLOP/s)", flush=True)
        K_bf16 = s.grm_block(0, 0, N0, N0)
        split_s = s.build_grm("bf16_split")
        print(f"  split BF16 GRM {split_s:.2f} s", flush=True)
        s.set_gemm_mode("bf16x3_tri")
        tune = s.autotune(reps=1)
        print(f"  autotune: rt{tune['rt']}/pd{tune['pd']}/s{tune['nsplit']} {tune['ms']:.2f} ms", flush=True)
        K64 = fp64_grm_subblock(G, miss, stats, mask)
        Xraw = synth.covariates(N, SEED[2])
        Xt, _, _ = plink.covariate_transform(Xraw)
        yield {"s": s, "G": G, "miss": miss, "K64": K64, "K_bf16": K_bf16, "Mqc": Mqc, "Xraw": Xraw, "Xt": Xt,
               "Zc": Zc}
    finally:
        s.close()


def test_c2_grm_subblock_bf16(c2):
    err = grm_rel_error(c2["K_bf16"], c2["K64"])
    off = grm_offdiag_error(c2["K_bf16"], c2["K64"])
    print(f"  C2 BF16 GRM subblock max|K - K64| / max|K64| = {err:.2e}, off-diagonal {off:.2e}", flush=True)
    assert err <= 4e-3


def test_c2_grm_subblock_split(c2):
    s = c2["s"]
    assert s.grm_precision() == "bf16_split"
    K = s.grm_block(0, 0, N0, N0)
    err = grm_rel_error(K, c2["K64"])
    off = grm_offdiag_error(K, c2["K64"])
    off_bf16 = grm_offdiag_error(c2["K_bf16"], c2["K64"])
    print(f"  C2 split GRM subblock max|K - K64| / max|K64| = {err:.2e}, off-diagonal {off:.2e}", flush=True)
    assert err <= 4.0 * np.sqrt(c2["Mqc"]) * 2.0 ** -24 + 2.0 ** -16    # + storage at 16 significant bits
    assert on_16bit_grid(K, c2["Mqc"])
    assert np.array_equal(K, K.T)
    # A single BF16 pass leaves each product a random ~2^-9 relative error:
    # off-diagonal error ~2^-9 / sqrt(Mqc) against entries ~1 / sqrt(Mqc).
    # The split leaves the dropped lo.lo' term, whose mean over a marker's
    # genotypes is a small common shift of every off-diagonal entry (taken up
    # by the intercept), several times below the single pass.
    assert off < 0.25 * off_bf16


def test_c2_block_pcg_64_columns_residual(c2):
    s, Xraw, Zc = c2["s"], c2["Xraw"], c2["Zc"]
    rng = np.random.default_rng(21)
    mu = rng.uniform(0.05, 0.3, N)
    w = mu * (1 - mu)
    tau = np.array([1.0, 0.1])
    B = np.empty((N, 64))
    B[:, :60] = np.column_stack([D.probe_vector(200, i, N) for i in range(60)])
    for k, prev in enumerate((0.05, 0.1, 0.2, 0.3)):
        y = synth.binary_phenotype(Xraw, synth.polygenic(Zc, 0.3, SEED[2] + 10 * k), prev, SEED[2] + 10 * k)
        B[:, 60 + k] = (y - mu) / w                              # a working residual
    X, iters = s.pcg_block(w, tau, B, 500, 1e-5)
    cols = [0, 17, 41, 59, 60, 61, 62, 63]
    r2 = fp64_residual2(s, w, tau, X[:, cols], B[:, cols])
    print(f"  C2 64-column block PCG: iterations {iters.min()}-{iters.max()}, FP64 ||Sigma x - b||^2 of "
          f"columns {cols}: max {r2.max():.3e}", flush=True)
    assert np.all(iters < 500)
    assert np.all(r2 <= 1.5e-5)


def test_c2_subblock_fits_and_solves_vs_fp64_oracle(c2):
    s, K64, G, miss, Xraw = c2["s"], c2["K64"], c2["G"], c2["miss"], c2["Xraw"]
    sub = np.arange(N0)
    Xs, _, _ = plink.covariate_transform(Xraw[sub])
    Zs = causal_std(G, miss, 500, SEED[2])
    opt = D.Options(nrun=30)
    grm = O.DenseGRM(K64, 1)
    probes = D.RademacherProbes(N0, 200)
    kw = dict(nrun=opt.nrun, tolPCG=opt.tolPCG, maxiterPCG=opt.maxiterPCG, cv_cutoff=opt.traceCVcutoff,
              tol=opt.tol, maxiter=opt.maxiter)
    # On 2,048 unrelated samples PQL leaves tau1 at its boundary 0 for about
    # half of the binary draws; the comparison wants an interior optimum, so
    # the phenotype is the first draw whose FP64 oracle fit has tau1 > 0.02.
    refs = {}
    for k in range(8):
        yb = synth.binary_phenotype(Xraw[sub], synth.polygenic(Zs, 2.0, SEED[2] + 1 + 10 * k), 0.2,
                                    SEED[2] + 1 + 10 * k)
        refs["binary"] = O.glmmkin_binary(grm, yb, Xs, probes, **kw)
        if refs["binary"]["theta"][1] > 0.02:
            break
    yq = synth.quantitative_phenotype(Xraw[sub], synth.polygenic(Zs, 0.6, SEED[2] + 2), 0.6, SEED[2] + 2)
    refs["quantitative"] = O.glmmkin_quantitative(grm, yq, Xs, probes, **kw)
    jobs = [{"trait": "binary", "y": yb, "X": Xs, "subset": sub, "tag": "binary"},
            {"trait": "quantitative", "y": yq, "X": Xs, "subset": sub, "tag": "quantitative"}]
    hs, _ = D.fit_batched(s, jobs, opt)
    for h in hs:
        assert h.error is None, h.error
        ref = refs[h.tag]
        rel = np.abs(h.result["theta"] - ref["theta"]) / np.abs(ref["theta"])
        print(f"  C2 subblock {h.tag} fit: theta {h.result['theta']} FP64 oracle {ref['theta']} (rel error "
              f"{rel}), iterations {h.result['iterations']} / {ref['iterations']}", flush=True)
        assert ref["theta"][1] > 0.02
        assert h.result["iterations"] == ref["iterations"]
        assert np.allclose(h.result["theta"], ref["theta"], rtol=1e-3, atol=1e-6)
    # block solves on the subblock (64 columns, the BF16-split GEMM on K_SS) against dense FP64
    mask = np.zeros(N, np.uint8)
    mask[:N0] = 1
    rng = np.random.default_rng(22)
    w = np.ones(N)
    w[:N0] = rng.uniform(0.1, 0.25, N0)
    tau = np.array([1.0, 0.3])
    B = np.zeros((N, 64))
    B[:N0] = rng.standard_normal((N0, 64))
    Xsol, it = engine_solve(s, w, tau, B, mask, 500, 1e-5)
    exact = np.linalg.solve(tau[0] * np.diag(1 / w[:N0]) + tau[1] * K64, B[:N0])
    err = float(np.max(np.abs(Xsol[:N0] - exact)) / np.max(np.abs(exact)))
    print(f"  C2 subblock 64-column solve: rel error vs FP64 {err:.2e}, iterations {it.min()}-{it.max()}",
          flush=True)
    assert err < 2e-3
    assert np.all(Xsol[N0:] == 0)


def test_c2_fit_bitwise_alone_and_in_a_64_column_mix(c2):
    s, Xraw, Xt, Zc = c2["s"], c2["Xraw"], c2["Xt"], c2["Zc"]
    opt = D.Options(nrun=15)
    jobs = []
    for k, prev in enumerate((0.05, 0.1, 0.2, 0.3)):
        y = synth.binary_phenotype(Xraw, synth.polygenic(Zc, 0.3, SEED[2] + 10 * k), prev, SEED[2] + 10 * k)
        jobs.append({"trait": "binary", "y": y, "X": Xt, "tag": f"prev{prev}"})
    hs, sch = D.fit_batched(s, jobs, opt)
    assert all(h.error is None for h in hs), [h.error for h in hs]
    assert sch.busy_slot_iters > 60 * 3                       # full 64-column iterations ran
    alone, _ = D.fit_batched(s, jobs[2:3], opt)
    print(f"  C2 mix: theta {[list(h.result['theta']) for h in hs]}; alone {alone[0].result['theta']}", flush=True)
    assert np.array_equal(alone[0].result["theta"], hs[2].result["theta"])
    assert np.array_equal(alone[0].result["coefficients"], hs[2].result["coefficients"])
    assert np.array_equal(alone[0].result["linear_predictors"], hs[2].result["linear_predictors"])


def test_c3_serving_poisson_arrivals(c2):
    s, Xraw, Zc = c2["s"], c2["Xraw"], c2["Zc"]
    rng = np.random.default_rng(SEED[3])
    opt = D.Options(nrun=15)
    reqs = []
    t = 0.0
    for r in range(14):
        t += rng.exponential(1.0 / 8.0)
        sub = np.sort(rng.choice(N, int(rng.uniform(0.2, 1.0) * N), replace=False))
        Xs, _, _ = plink.covariate_transform(Xraw[sub])
        if rng.random() < 0.7:
            y = synth.binary_phenotype(Xraw[sub], synth.polygenic(Zc[sub], 0.3, SEED[3] + r),
                                       rng.uniform(0.01, 0.3), SEED[3] + r)
            trait = "binary"
        else:
            h2 = rng.uniform(0.1, 0.6)
            y = synth.quantitative_phenotype(Xraw[sub], synth.polygenic(Zc[sub], h2, SEED[3] + r), h2, SEED[3] + r)
            trait = "quantitative"
        reqs.append((t, trait, y, Xs, sub))
    arrivals = [(t, D.fit_co(tr, y, X, D.RademacherProbes(len(sub), 200), opt, lib=True), sub, k)
                for k, (t, tr, y, X, sub) in enumerate(reqs)]
    with D.BatchScheduler(s, opt, width=64, n_systems=16) as sch:
        done = sch.run(arrivals)
    assert len(done) == len(reqs) and all(f.error is None for f in done), [f.error for f in done]
    lat = sorted(f.t_done - f.t_submit for f in done)
    overlap = sum(1 for f in done for g in done if f is not g and g.t_admit < f.t_done and f.t_admit < g.t_done)
    print(f"  C3 serving: {len(done)} fits, latency p50 {lat[len(lat) // 2]:.2f}s max {lat[-1]:.2f}s, "
          f"{overlap} overlapping pairs", flush=True)
    assert overlap > 0                                          # requests shared the device
    by_tag = {f.tag: f for f in done}
    for k in (0, 5, 11):
        t_, tr, y, X, sub = reqs[k]
        hs, _ = D.fit_batched(s, [{"trait": tr, "y": y, "X": X, "subset": sub}], opt)
        assert np.array_equal(hs[0].result["theta"], by_tag[k].result["theta"]), k
        assert np.array_equal(hs[0].result["coefficients"], by_tag[k].result["coefficients"]), k


def test_c3_serving_through_the_c_abi(c2):
    """The same stream through saige_submit / saige_poll: the server's
    scheduler thread admits requests as systems and HBM free up while the
    caller keeps submitting; each result is bitwise the same request run
    alone (saige_fit_null_batch)."""
    import time
    s, Xraw, Zc = c2["s"], c2["Xraw"], c2["Zc"]
    rng = np.random.default_rng(SEED[3] + 1)
    opt = D.Options(nrun=15)
    jobs, offsets = [], []
    t = 0.0
    for r in range(16):
        t += rng.exponential(1.0 / 10.0)
        sub = np.sort(rng.choice(N, int(rng.uniform(0.2, 1.0) * N), replace=False))
        Xs, _, _ = plink.covariate_transform(Xraw[sub])
        if rng.random() < 0.7:
            y = synth.binary_phenotype(Xraw[sub], synth.polygenic(Zc[sub], 0.3, SEED[3] + 50 + r),
                                       rng.uniform(0.01, 0.3), SEED[3] + 50 + r)
            jobs.append({"trait": "binary", "y": y, "X": Xs, "subset": sub})
        else:
            h2 = rng.uniform(0.1, 0.6)
            y = synth.quantitative_phenotype(Xraw[sub], synth.polygenic(Zc[sub], h2, SEED[3] + 50 + r), h2,
                                             SEED[3] + 50 + r)
            jobs.append({"trait": "quantitative", "y": y, "X": Xs, "subset": sub})
        offsets.append(t)
    results, t_sub = {}, {}
    with s.server(opt, max_concurrent=16, hbm_reserve_bytes=1 << 30) as srv:
        t0 = time.perf_counter()
        tickets = []
        for k, (job, toff) in enumerate(zip(jobs, offsets)):
            while time.perf_counter() - t0 < toff:
                for tk in tickets:                      # collect whatever finished meanwhile
                    if tk not in results:
                        r = srv.poll(tk, 0)
                        if r is not None:
                            results[tk] = r
                time.sleep(0.002)
            t_sub[k] = time.perf_counter() - t0
            tickets.append(srv.submit(job))
        for tk in tickets:
            if tk not in results:
                results[tk] = srv.poll(tk, -1)
    res = [results[tk] for tk in tickets]
    assert all(r["status"] == 0 for r in res), [r["status"] for r in res]
    span = [(t_sub[k] + r["t_admit"], t_sub[k] + r["t_done"]) for k, r in enumerate(res)]
    overlap = sum(1 for a in span for b in span if a is not b and b[0] < a[1] and a[0] < b[1])
    lat = sorted(r["t_done"] for r in res)
    print(f"  C3 C-ABI serving: {len(res)} fits, latency p50 {lat[len(lat) // 2]:.2f}s max {lat[-1]:.2f}s, "
          f"{overlap} overlapping pairs", flush=True)
    assert overlap > 0
    for k in (0, 6, 13):
        alone = s.fit_null_batch([jobs[k]], opt)[0]
        assert alone["status"] == 0
        assert np.array_equal(alone["theta"], res[k]["theta"]), k
        assert np.array_equal(alone["coefficients"], res[k]["coefficients"]), k
        assert np.array_equal(alone["linear_predictors"], res[k]["linear_predictors"]), k
