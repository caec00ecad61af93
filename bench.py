"""Benchmark: null-GLMM fits/hour at N=100k, batched vs plain, plus GRM
TFLOP/s, on one MI355X.

Workload (default): N = 100,000 samples, M = 150,000 markers, MAF ~
U(0.01, 0.5), 1% missing calls, synthetic PLINK .bed generated in HBM;
intercept + age + sex covariates (QR-transformed as fitNULLGLMM does);
four binary phenotypes with prevalences 0.05, 0.1, 0.2, 0.3.  The GRM is
built once from BF16 standardised genotypes on MFMA with FP32 accumulation
(a single pass, timed for GRM TFLOP/s, then the split build the fits use:
each value as two BF16 pieces, three products) and held in HBM as FP32
(the split GRM's values rounded to 16 significant bits).

A step = one batched round of null fits of the four phenotypes through the
library's C++ drivers (saige_fit_null_batch): AI-REML with nrun = 15
Rademacher probes per phenotype, so each trace phase is a block PCG over
4 x (15 probes + 1) = 64 right-hand sides (FP32 vectors; GRM products at
FP32 accuracy from BF16 pieces on the MFMA, three of each vector and two of
the GRM's 16-bit values, read from the GRM's lower triangle; FP64 dot
products).  `value` = fits completed per hour of
step time.  The Python drivers over the same engine and the plain path (one
phenotype, one right-hand side at a time, GEMV per PCG iteration) are timed
beside it.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "saige-gpu_amd")]

HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK = 157.3
# HBM bytes per 64-column block-GEMM launch (all its kernels, FETCH_SIZE and
# WRITE_SIZE scaled by a 1 GiB copy), measured with rocprofv3; keyed by
# (mode, N, rt, pd, nsplit).  Sources: PROFILE_FILES
PROFILED_GEMM_TRAFFIC = {("fp32", 100_000, 1, 3, 8): (42.10 + 0.205 + 0.108 + 0.026) * 1e9,
                         ("bf16x3", 100_000, 1, 3, 8): (0.064 + 41.579 + 0.152) * 1e9,
                         # lower triangle on 16x16 tiles, on the split GRM (two pieces of
                         # S): all four kernels of a launch (pd 3 spills)
                         ("bf16x3_tri", 100_000, 1, 1, 8): (25.27 + 4.08) * 1e9,
                         ("bf16x3_tri", 100_000, 2, 1, 8): (26.63 + 5.54) * 1e9,
                         ("bf16x3_tri", 100_000, 1, 3, 8): (46.77 + 6.12) * 1e9,
                         ("bf16x3_tri", 100_000, 2, 3, 8): (29.99 + 9.17) * 1e9}
PROFILE_FILES = {"fp32": "profiles/kernels_N100k_M150k_bf16.md", "bf16x3": "profiles/gemm_bf16x3_N100k_M150k.md",
                 "bf16x3_tri": "profiles/kernels_N100k_M150k_split_grm.md"}
BF16_MFMA_PEAK = 2500.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(sess, N, M, maf_lo, maf_hi, miss, seed, prevalences, ncausal=1000, h2=0.3, keep=None):
    import numpy as np
    from saige_gpu import plink, synth

    bpm = (N + 3) // 4
    dev = sess.alloc_device(M * bpm)
    sess.synth_bed_device(dev, N, M, maf_lo, maf_hi, miss, 0.0, seed)
    rng = np.random.default_rng(seed)
    causal = np.sort(rng.choice(M, ncausal, replace=False))
    rows = np.stack([sess.copy_device_to_host(dev, bpm, offset=int(j) * bpm) for j in causal])
    G, mm = synth.decode_bed_rows(rows, N)
    Zc = synth.standardized(G, mm)
    Xraw = synth.covariates(N, seed)
    X, _, _ = plink.covariate_transform(Xraw)
    phenos = []
    for k, prev in enumerate(prevalences):
        g = synth.polygenic(Zc, h2, seed + 10 * k)
        phenos.append(synth.binary_phenotype(Xraw, g, prev, seed + 10 * k))
    if keep is not None:
        keep["Xraw"] = Xraw
        keep["Zc"] = Zc
    return dev, X, phenos


def _phenotype(trait, Xraw, g, param, rng):
    """A phenotype on one request's samples: binary with logit-scale genetic
    part g and prevalence param (intercept by bisection on up to 4096 of the
    samples), or quantitative with h2 = param (g scaled to variance h2)."""
    import numpy as np
    lin = Xraw[:, 1:] @ np.array([0.02, 0.3]) + g
    lin = lin - lin.mean()
    if trait == "binary":
        ref = lin[rng.choice(len(lin), min(4096, len(lin)), replace=False)]
        lo, hi = -30.0, 30.0
        for _ in range(60):
            a = 0.5 * (lo + hi)
            lo, hi = (a, hi) if np.mean(1.0 / (1.0 + np.exp(-(a + ref)))) < param else (lo, a)
        return (rng.random(len(lin)) < 1.0 / (1.0 + np.exp(-(a + lin)))).astype(np.float64)
    gs = g * np.sqrt(param) / max(g.std(), 1e-12)
    return Xraw[:, 1:] @ np.array([0.02, 0.3]) + gs + rng.normal(0.0, np.sqrt(1.0 - param), len(g))


def serve_leg(sess, N, Xraw, Zc, n_req, rate, seed, opt):
    """Continuous-batching serving through the library's request server
    (saige_submit / saige_poll): a Poisson stream of null-fit requests on
    the resident GRM, each on a random 20-100% sample subset, 70% binary
    (prevalence U(0.01, 0.3)) and 30% quantitative (h2 U(0.1, 0.6)).  The
    genetic parts come from a pool of 16 polygenic vectors over all samples;
    a request's phenotype is drawn at its submission.  Latency = the
    server's own clock from submission to completion."""
    import numpy as np
    from saige_gpu import plink

    rng = np.random.default_rng(seed)
    pool = [Zc @ rng.normal(0.0, 1.0, Zc.shape[1]) for _ in range(16)]
    pool = [(g - g.mean()) / g.std() for g in pool]
    t_arr = np.cumsum(rng.exponential(1.0 / rate, n_req))
    results, tickets = {}, []
    with sess.server(opt, max_concurrent=16, hbm_reserve_bytes=2 << 30) as srv:
        t0 = time.perf_counter()
        for r in range(n_req):
            sub = np.sort(rng.choice(N, int(rng.uniform(0.2, 1.0) * N), replace=False))
            Xs, _, _ = plink.covariate_transform(Xraw[sub])
            g = pool[r % 16][sub]
            if rng.random() < 0.7:
                job = {"trait": "binary", "X": Xs, "subset": sub,
                       "y": _phenotype("binary", Xraw[sub], np.sqrt(0.3) * g, rng.uniform(0.01, 0.3), rng)}
            else:
                job = {"trait": "quantitative", "X": Xs, "subset": sub,
                       "y": _phenotype("quantitative", Xraw[sub], g, rng.uniform(0.1, 0.6), rng)}
            while time.perf_counter() - t0 < t_arr[r]:
                for tk in tickets:
                    if tk not in results:
                        out = srv.poll(tk, 0)
                        if out is not None:
                            results[tk] = out
                time.sleep(0.001)
            tickets.append(srv.submit(job))
            if (r + 1) % 200 == 0:
                log(f"serving: {r + 1} of {n_req} requests submitted, {len(results)} done")
        for tk in tickets:
            if tk not in results:
                results[tk] = srv.poll(tk, -1)
        wall = time.perf_counter() - t0
    res = [results[tk] for tk in tickets]
    ok = [r for r in res if r["status"] == 0]
    lat = np.array([r["t_done"] for r in ok])
    return {"path": "saige_submit / saige_poll (C++ drivers and scheduler)", "requests": n_req,
            "completed": len(ok), "errors": len(res) - len(ok), "offered_rate_per_s": rate, "wall_s": wall,
            "fits_per_hour": len(ok) / wall * 3600.0, "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "max_latency_s": float(lat.max()),
            "mean_queue_s": float(np.mean([r["t_admit"] for r in ok])), "max_concurrent_fits": 16}


class DenseFP64:
    """Exact FP64 answers for the fit coroutines (driver.fit_co) on a dense
    GRM: Sigma^-1 B by a Cholesky solve of Sigma = tau0 diag(1/w) + tau1 K,
    K B by a dense product; the trace rounds as their products and solves."""

    def __init__(self, K):
        self.K = K

    def _solve(self, w, tau, B):
        import numpy as np
        from scipy.linalg import cho_factor, cho_solve
        Sig = tau[1] * self.K + np.diag(tau[0] / np.asarray(w, np.float64))
        return cho_solve(cho_factor(Sig), np.asarray(B, np.float64))

    def run(self, gen):
        from saige_gpu import driver as D
        value = None
        while True:
            try:
                req = gen.send(value)
            except StopIteration as e:
                return e.value
            if req[0] == "solve":
                value = self._solve(*req[1:])
            elif req[0] == "matvec":
                value = self.K @ req[1]
            else:
                value = D.trace_round_host(req, lambda B: self.K @ B, self._solve)


def subblock_genotypes(sess, dev, N, M, n0, chunk=8192):
    """A1 counts and missing mask of samples 0..n0-1 from the device .bed."""
    import numpy as np
    from saige_gpu import synth
    bpm = (N + 3) // 4
    G = np.empty((n0, M), np.int8)
    miss = np.empty((n0, M), bool)
    for m0 in range(0, M, chunk):
        k = min(chunk, M - m0)
        raw = sess.copy_device_to_host(dev, k * bpm, offset=m0 * bpm).reshape(k, bpm)[:, : n0 // 4]
        G[:, m0:m0 + k], miss[:, m0:m0 + k] = synth.decode_bed_rows(raw, n0)
    return G, miss


def _offdiag_error(K, K64):
    import numpy as np
    D = K - K64
    np.fill_diagonal(D, 0.0)
    R = K64.copy()
    np.fill_diagonal(R, 0.0)
    return float(np.linalg.norm(D) / np.linalg.norm(R))


def fp64_check_leg(sess, G, miss, Xraw, opt, seed, K_single=None):
    """Error of the device path against FP64 on the principal subblock of
    the first n0 samples: the GRM (FP64 from the samples' genotypes over all
    GRM markers, imputed as FG.cpp:447-454, frequencies of the whole cohort;
    the resident GRM, and K_single -- the same subblock of a single-pass BF16
    build -- when given; max error over max|K64| and the off-diagonal
    relative Frobenius error, the relatedness signal tau1 rests on),
    a 64-column block solve on the subblock (engine with the sample mask)
    against a dense FP64 solve, and tau of a binary and a quantitative null
    fit on the subblock's samples (the batched device path on the resident
    GRM) against the same fits with exact FP64 solves on the FP64 GRM."""
    import numpy as np
    from saige_gpu import driver as D
    from saige_gpu import plink, synth

    n0 = G.shape[0]
    N = sess.getNnomissingOut()
    st, mask = sess.marker_stats(), sess.getQCdMarkerIndex().astype(bool)
    cols = np.nonzero(mask)[0]
    S = np.zeros((n0, n0))
    for c0 in range(0, len(cols), 16384):
        c = cols[c0:c0 + 16384]
        f = st["freq"][c].astype(np.float64)
        g = np.where(miss[:, c], st["fill"][c][None, :], G[:, c]).astype(np.float64)
        Z = (g - 2.0 * f) / np.sqrt(2.0 * f * (1.0 - f))
        S += Z @ Z.T
    K64 = S / len(cols)
    K = sess.grm_block(0, 0, n0, n0).astype(np.float64)
    grm_err = float(np.max(np.abs(K - K64)) / np.max(np.abs(K64)))
    grm_off = _offdiag_error(K, K64)
    single = None
    if K_single is not None:
        Ks = K_single.astype(np.float64)
        single = {"grm_max_rel_error": float(np.max(np.abs(Ks - K64)) / np.max(np.abs(K64))),
                  "grm_offdiag_rel_error": _offdiag_error(Ks, K64)}
    # a 64-column block solve on the subblock
    rng = np.random.default_rng(seed)
    w = np.ones(N)
    w[:n0] = rng.uniform(0.1, 0.25, n0)
    tau = np.array([1.0, 0.3])
    B = np.zeros((N, 64))
    B[:n0] = rng.standard_normal((n0, 64))
    msk = np.zeros(N, np.uint8)
    msk[:n0] = 1
    sess.engine_reserve(64, 1)
    sess.engine_set_system(0, w, tau, msk)
    sess.engine_add_batch(list(range(64)), [0] * 64, B, opt.maxiterPCG, opt.tolPCG)
    while not np.all(sess.engine_step(1, 64)[0][:64]):
        pass
    X = sess.engine_take_batch(list(range(64)))[:n0].astype(np.float64)
    exact = DenseFP64(K64)._solve(w[:n0], tau, B[:n0])
    sol_err = np.max(np.abs(X - exact), axis=0) / np.max(np.abs(exact), axis=0)
    # tau of two null fits on the subblock's samples
    sub = np.arange(n0)
    Xs, _, _ = plink.covariate_transform(Xraw[sub])
    causal = np.sort(rng.choice(G.shape[1], 500, replace=False))
    Zs = synth.standardized(G[:, causal], miss[:, causal])
    jobs = [{"trait": "binary", "X": Xs, "subset": sub,
             "y": synth.binary_phenotype(Xraw[sub], synth.polygenic(Zs, 2.0, seed), 0.2, seed)},
            {"trait": "quantitative", "X": Xs, "subset": sub,
             "y": synth.quantitative_phenotype(Xraw[sub], synth.polygenic(Zs, 0.6, seed + 1), 0.6, seed + 1)}]
    hs, _ = D.fit_batched(sess, jobs, opt)
    fits = {}
    for h, j in zip(hs, jobs):
        if h.error is not None:
            raise h.error
        ref = DenseFP64(K64).run(D.fit_co(j["trait"], j["y"], Xs, D.RademacherProbes(n0, 200), opt))
        fits[j["trait"]] = {"theta": list(map(float, h.result["theta"])),
                            "theta_fp64": list(map(float, ref["theta"])),
                            "rel_error": [float(abs(a - b) / abs(b)) if b != 0 else float(abs(a - b))
                                          for a, b in zip(h.result["theta"], ref["theta"])],
                            "iterations": [int(h.result["iterations"]), int(ref["iterations"])]}
    return {"samples": n0, "grm": sess.grm_precision(), "grm_max_rel_error": grm_err,
            "grm_offdiag_rel_error": grm_off, "single_pass_bf16": single,
            "block_solve_rel_error": {"max": float(sol_err.max()), "median": float(np.median(sol_err))},
            "tau": fits,
            "reference": "FP64 GRM of the subblock from its genotypes over all GRM markers; exact (Cholesky) "
                         "FP64 solves; the same driver and probes"}


class CpuReference:
    """The reference's CPU path, one right-hand side at a time:
    getPCG1ofSigmaAndVector in FP32 (FG.cpp:2593-2809) over
    parallelCrossProdOpenMP (FG.cpp:1576-1598: every marker decoded on every
    product) -- the oracle's restatement -- driving the fit coroutines."""

    def __init__(self, O, g, diag, threads, opt):
        self.O, self.g, self.diag, self.threads, self.opt = O, g, diag, threads, opt

    def _solve(self, w, tau, B):
        import numpy as np
        B = np.asarray(B, np.float64).reshape(len(w), -1)
        w32, t32 = np.asarray(w, np.float32), np.asarray(tau, np.float32)
        return np.column_stack([self.O.pcg_f32(self.g, self.diag, w32, t32, B[:, c].astype(np.float32),
                                               self.opt.maxiterPCG, self.opt.tolPCG, nthreads=self.threads)[0]
                                for c in range(B.shape[1])]).astype(np.float64)

    def _matvec(self, B):
        import numpy as np
        B = np.asarray(B, np.float64).reshape(self.g["Nnm"], -1)
        return np.column_stack([self.O.crossprod(self.g, B[:, c], nthreads=self.threads)
                                for c in range(B.shape[1])]).astype(np.float64)

    def run(self, gen):
        from saige_gpu import driver as D
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
                value = D.trace_round_host(req, self._matvec, self._solve)


def cpu_baseline(sess, dev, N, M, n_matvec_fit, threads, sample_markers, min_seconds=10.0):
    """Oracle restatement of the reference CPU matvec (parallelCrossProdOpenMP,
    FG.cpp:1576-1598: decode every marker on every product, FP32 dot+axpy)
    timed on a marker sample and scaled to the full marker count."""
    import numpy as np
    from oracle import oracle as O

    bpm = (N + 3) // 4
    k = min(sample_markers, M)
    body = sess.copy_device_to_host(dev, k * bpm)
    og = O.setgeno(body, N, k, min_maf=0.01, max_missing=0.15)
    b = np.random.default_rng(1).standard_normal(N).astype(np.float32)
    reps = 0
    t0 = time.perf_counter()
    while True:
        O.crossprod_raw(og["packed"], og["Mqc"], N, og["freq"], og["inv_std"], b, threads)
        reps += 1
        if time.perf_counter() - t0 >= min_seconds:
            break
    per_marker = (time.perf_counter() - t0) / (reps * og["Mqc"])
    return per_marker, og["Mqc"], reps


def cpu_c0_fit(threads, opt):
    """C0 (N = 2,000, M = 10,000, binary, h2 0.3 on 1,000 causal markers,
    prevalence 0.2, intercept + age + sex) fitted end to end on the CPU
    path of CpuReference.  Returns (seconds, fit)."""
    from oracle import oracle as O
    from saige_gpu import driver as D
    from saige_gpu import plink, synth
    c = synth.small_cohort(2000, 10000, maf_lo=0.05, maf_hi=0.5, miss_rate=0.005, seed=synth.config_seed(0),
                           ncausal=1000)
    og = O.setgeno(c["bed"], 2000, 10000, min_maf=0.01, max_missing=0.15)
    X, _, _ = plink.covariate_transform(c["X"])
    y = synth.binary_phenotype(c["X"], synth.polygenic(c["Zc"], 0.3, c["seed"]), 0.2, c["seed"])
    ex = CpuReference(O, og, O.diag_std(og), threads, opt)
    t0 = time.perf_counter()
    fit = ex.run(D.fit_co("binary", y, X, D.RademacherProbes(2000, 200), opt))
    return time.perf_counter() - t0, fit


def c4_leg(args):
    """Large related cohort: N=200,000, M=500,000, MAF U(0.01, 0.5), 2% missing,
    10% of samples in sibling pairs; FP32 GRM (~160 GB) resident; binary trait
    (prevalence 0.01, h2 0.3); null fit, then variance ratios on 100 markers
    in each of 4 MAF bins, held out of the GRM and solved as block-PCG
    right-hand sides."""
    import numpy as np
    from saige_gpu import Session, plink, synth
    from saige_gpu import driver as D

    N, M = args.N, args.M
    seed = 20261202 + 1000 * 4
    sess = Session(0)
    bpm = (N + 3) // 4
    dev = sess.alloc_device(M * bpm)
    sess.synth_bed_device(dev, N, M, 0.01, 0.5, 0.02, 0.10, seed)
    sess.setminMAFforGRM(0.01)
    sess.setmaxMissingRateforGRM(0.15)
    sess.setminMAC_VarianceRatio(20, -1, False)
    sess.setgeno_device(dev, N, M)
    st = sess.marker_stats()
    qc = (st["flags"] & 1).astype(bool)
    maf = np.minimum(st["freq"], 1 - st["freq"])
    rng = np.random.default_rng(seed)
    bins = [(0.01, 0.05), (0.05, 0.1), (0.1, 0.25), (0.25, 0.5001)]
    pool, cats_m = [], []
    for lo, hi in bins:
        cand = np.nonzero(qc & (maf >= lo) & (maf < hi) & (st["mac"] >= 20))[0]
        pick = np.sort(rng.choice(cand, 100, replace=False))
        pool.append(pick)
        cats_m.append(pick)
    pool = np.sort(np.concatenate(pool)).astype(np.int32)
    sess.setminMAC_VarianceRatio(20, -1, True)
    sess.set_vr_random_markers(pool)
    sess.setgeno_device(dev, N, M)              # VR markers leave the GRM (FG.cpp:545-547)
    Mqc = sess.getMsub_MAFge_minMAFtoConstructGRM()
    if Mqc < 0.9 * (M - len(pool)):
        raise RuntimeError(f"synthetic cohort malformed: only {Mqc} of {M} markers pass QC")
    vr_index = sess.getIndexVec_forVarRatio()
    pos = {int(m): i for i, m in enumerate(vr_index)}
    categories = [[pos[int(m)] for m in rng.permutation(c) if int(m) in pos] for c in cats_m]
    causal = np.sort(rng.choice(M, 1000, replace=False))
    rows = np.stack([sess.copy_device_to_host(dev, bpm, offset=int(j) * bpm) for j in causal])
    sess.free_device(dev)                       # the .bed body is no longer needed: 25 GB back
    G, mm = synth.decode_bed_rows(rows, N)
    Zc = synth.standardized(G, mm)
    Xraw = synth.covariates(N, seed)
    X, _, _ = plink.covariate_transform(Xraw)
    y = synth.binary_phenotype(Xraw, synth.polygenic(Zc, 0.3, seed), 0.01, seed)
    del G, mm, Zc
    grm_s = sess.build_grm(args.grm)
    grm_tflops = N * (N + 1) * Mqc / grm_s / 1e12
    log(f"C4: Mqc={Mqc}, VR markers {len(vr_index)}, GRM {args.grm} {grm_s:.2f}s {grm_tflops:.0f} TFLOP/s")
    sess.set_gemm_mode(args.gemm_mode)
    tune = sess.autotune(reps=3)
    log(f"C4 autotune: rt{tune['rt']}/pd{tune['pd']}/s{tune['nsplit']} {tune['ms']:.2f} ms")
    opt = D.Options(nrun=30)
    g0, n0 = sess.engine_stats("gemm_full")
    t0 = time.perf_counter()
    handles, sch = D.fit_batched(sess, [{"trait": "binary", "y": y, "X": X}], opt)
    if handles[0].error is not None:
        raise handles[0].error
    fit = handles[0].result
    t1 = time.perf_counter()
    vr = D.variance_ratio(sess, fit, y, X, "binary", categories, opt, vr_pool=True, num_markers=100)
    t2 = time.perf_counter()
    g1, n1 = sess.engine_stats("gemm_full")
    gemm_avg = (g1 - g0) / max(1, n1 - n0)
    out = {"config": "C4", "N": N, "M": M, "Mqc": Mqc, "grm": args.grm, "grm_seconds": grm_s,
           "grm_tflops": grm_tflops if args.grm != "bf16_split" else None,
           "grm_GB": 4.0 * (N + 255) // 256 * 256 * ((N + 255) // 256 * 256) / 1e9,
           "fit_seconds": t1 - t0, "theta": list(map(float, fit["theta"])), "vr_seconds": t2 - t1,
           "variance_ratios": [row[0] for row in vr["table"]], "vr_markers": [len(r) for r in vr["ratios"]],
           "gemm_avg_ms": gemm_avg * 1e3, "gemm_tflops": 2.0 * N * N * 64 / gemm_avg / 1e12,
           # bytes of S a product must read: its lower triangle in that mode
           "gemm_hbm_GBs": (2.0 if args.gemm_mode == "bf16x3_tri" else 4.0) * N * N / gemm_avg / 1e9,
           "gemm_mode": args.gemm_mode, "gemm_autotune": tune}
    log(f"C4: fit {t1 - t0:.1f}s theta {fit['theta']}, VR {t2 - t1:.1f}s ratios {out['variance_ratios']}")
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=100_000)
    ap.add_argument("--M", type=int, default=150_000)
    ap.add_argument("--nrun", type=int, default=15,
                    help="probes per trace round: 15 makes the C2 block of 60 probes + 4 phenotypes")
    ap.add_argument("--grm", default="bf16_split", choices=["bf16_split", "bf16", "fp32"],
                    help="GRM the fits use (bf16_split: FP32-accurate from three BF16 MFMA passes)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-nrun30", action="store_true")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--cpu-sample-markers", type=int, default=4000)
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--config", default="c2", choices=["c2", "c4"])
    ap.add_argument("--gemm-mode", default="bf16x3_tri", choices=["bf16x3_tri", "bf16x3", "fp32"],
                    help="block-GEMM arithmetic: FP32 operands split into three BF16 pieces on the BF16 "
                         "MFMA (FP32-accurate), the same over the GRM's lower triangle only, or FP32 MFMA")
    ap.add_argument("--loco-chromosomes", type=int, default=22,
                    help="leave-one-chromosome-out fits for this many autosomes (0: skip)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    if args.config == "c4":
        if args.N == 100_000:
            args.N = 200_000
        if args.M == 150_000:
            args.M = 500_000
        c4_leg(args)
        return

    import numpy as np
    from saige_gpu import Session
    from saige_gpu import driver as D

    N, M = args.N, args.M
    seed = 20261202 + 1000 * 2
    prevalences = [0.05, 0.1, 0.2, 0.3]
    sess = Session(0)
    t = time.perf_counter()
    kept = {}
    dev, X, phenos = make_workload(sess, N, M, 0.01, 0.5, 0.01, seed, prevalences, keep=kept)
    log(f"workload generated in {time.perf_counter() - t:.1f}s")
    sess.setminMAFforGRM(0.01)
    sess.setmaxMissingRateforGRM(0.15)
    t = time.perf_counter()
    sess.setgeno_device(dev, N, M)
    setgeno_s = time.perf_counter() - t
    Mqc = sess.getMsub_MAFge_minMAFtoConstructGRM()
    log(f"setgeno (decode/QC/repack/diag) {setgeno_s:.3f}s, Mqc={Mqc}")
    n0 = 2048
    G0 = miss0 = None
    if not args.no_fp64_check:
        G0, miss0 = subblock_genotypes(sess, dev, N, M, n0)
    # GRM TFLOP/s is the single-pass build (N(N+1)Mqc flop of the SYRK's
    # lower triangle); the split build the fits use does three times the MFMA work
    grm_s = sess.build_grm("bf16" if args.grm == "bf16_split" else args.grm)
    grm_tflops = N * (N + 1) * Mqc / grm_s / 1e12
    log(f"GRM {sess.grm_precision()}: {grm_s:.3f}s device time, {grm_tflops:.1f} TFLOP/s")
    K_single, split = None, None
    if args.grm == "bf16_split":
        if G0 is not None:
            K_single = sess.grm_block(0, 0, n0, n0)
        split_s = sess.build_grm("bf16_split")
        split = {"seconds": split_s, "fp32_equivalent_tflops": N * (N + 1) * Mqc / split_s / 1e12,
                 "bf16_piece_tflops": 3 * N * (N + 1) * Mqc / split_s / 1e12,
                 "note": "each standardised value as two BF16 pieces; hi.hi + hi.lo + lo.hi on the MFMA"}
        log(f"GRM bf16_split: {split_s:.3f}s device time ({split['bf16_piece_tflops']:.0f} TFLOP/s of BF16 "
            f"piece products)")
    sess.set_gemm_mode(args.gemm_mode)
    tune = sess.autotune(reps=3)
    log(f"autotune block GEMM ({args.gemm_mode}): " + ", ".join(f"rt{a}/pd{b}/s{c} {t:.2f}ms" for a, b, c, t in tune["table"]) +
        f" -> rt{tune['rt']}/pd{tune['pd']}/s{tune['nsplit']}; one half: " +
        ", ".join(f"rt{a}/pd{b} {t:.2f}ms" for a, b, _, t in tune["half_table"]) +
        f" -> rt{tune['half']['rt']}/pd{tune['half']['pd']}" +
        ("; 16 columns: " + ", ".join(f"rt{a}/pd{b} {t:.2f}ms" for a, b, _, t in tune["quarter_table"]) +
         f" -> rt{tune['quarter']['rt']}/pd{tune['quarter']['pd']}" if tune["quarter_table"] else ""))

    opt = D.Options(nrun=args.nrun)
    jobs = [{"trait": "binary", "y": y, "X": X, "tag": f"prev{p}"} for y, p in zip(phenos, prevalences)]

    def step(o=opt):
        handles, sch = D.fit_batched(sess, jobs, o, width=64, n_systems=len(jobs))
        for h in handles:
            if h.error is not None:
                raise h.error
        return handles, sch

    def native_step(o=opt):
        res = sess.fit_null_batch(jobs, o)
        if any(r["status"] != 0 for r in res):
            raise RuntimeError(f"native driver failed: {[r['status'] for r in res]}")
        return res

    # the timed steps: the library's C++ drivers and scheduler
    # (saige_fit_null_batch), the path an R or C++ caller reaches
    for _ in range(args.warmup):
        native_step()
    g0, n0_ = sess.engine_stats("gemm_full")
    h0, m0 = sess.engine_stats("gemm_half")
    q0_, k0 = sess.engine_stats("gemm_quarter")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nat = native_step()
    t1 = time.perf_counter()
    g1, n1 = sess.engine_stats("gemm_full")
    h1, m1 = sess.engine_stats("gemm_half")
    q1_, k1 = sess.engine_stats("gemm_quarter")
    half_launches = m1 - m0
    half_avg = (h1 - h0) / max(1, half_launches)
    quarter_launches = k1 - k0
    quarter_avg = (q1_ - q0_) / max(1, quarter_launches)
    step_s = (t1 - t0) / args.steps
    fits = len(jobs) * args.steps
    fits_per_hour = fits / (t1 - t0) * 3600.0
    gemm_launches = n1 - n0_
    gemm_avg = (g1 - g0) / max(1, gemm_launches)
    gemm_flop = 2.0 * N * N * 64
    gemm_tflops = gemm_flop / gemm_avg / 1e12
    nat_s = step_s
    log(f"C++ drivers (saige_fit_null_batch): {step_s:.3f}s/step, {fits_per_hour:.1f} fits/h, {gemm_launches} "
        f"64-column GEMMs avg {gemm_avg * 1e3:.2f} ms ({gemm_tflops:.1f} TFLOP/s), {half_launches} 32-column GEMMs "
        f"avg {half_avg * 1e3:.2f} ms, {quarter_launches} 16-column GEMMs avg {quarter_avg * 1e3:.2f} ms "
        f"({(g1 - g0 + h1 - h0 + q1_ - q0_) / args.steps:.3f}s of GEMM per step), theta "
        f"{[list(np.round(r['theta'], 4)) for r in nat]}")

    # the same rounds through the Python drivers and scheduler (the
    # library's test harness), timed over the same number of steps
    step()
    st0 = [sess.engine_stats(k) for k in ("gemm_full", "gemm_half")]
    t0 = time.perf_counter()
    for _ in range(args.steps):
        handles, sch = step()
    py_s = (time.perf_counter() - t0) / args.steps
    st1 = [sess.engine_stats(k) for k in ("gemm_full", "gemm_half")]
    lat = sorted(h.t_done - h.t_submit for h in handles)
    python_driver = {
        "fits_per_hour": len(jobs) / py_s * 3600.0, "seconds_per_step": py_s, "steps": args.steps,
        "max_rel_theta_diff_vs_cpp_driver": max(
            float(np.max(np.abs(h.result["theta"] - r["theta"]) / np.maximum(np.abs(r["theta"]), 1e-12)))
            for r, h in zip(nat, handles)),
        "gemm_full_launches_per_step": (st1[0][1] - st0[0][1]) / args.steps,
        "gemm_half_launches_per_step": (st1[1][1] - st0[1][1]) / args.steps,
        "gemm_device_seconds_per_step": ((st1[0][0] - st0[0][0]) + (st1[1][0] - st0[1][0])) / args.steps}
    log("Python scheduler step split (s): " + ", ".join(f"{k} {v:.3f}" for k, v in sorted(sch.timing.items())))
    log(f"Python drivers: {py_s:.3f}s/step, {python_driver['fits_per_hour']:.0f} fits/h, max rel theta diff vs the "
        f"C++ drivers {python_driver['max_rel_theta_diff_vs_cpp_driver']:.1e}")

    # the same round with the block GEMM on the FP32 MFMA, for comparison:
    # tuned in that mode, one step timed, results checked against the step above
    fp32_mode = None
    if args.gemm_mode != "fp32":
        sess.set_gemm_mode("fp32")
        t32 = sess.autotune(reps=3)
        g0, n0_ = sess.engine_stats("gemm_full")
        h0, m0 = sess.engine_stats("gemm_half")
        t0 = time.perf_counter()
        r32 = native_step()
        s32 = time.perf_counter() - t0
        g1, n1 = sess.engine_stats("gemm_full")
        h1, m1 = sess.engine_stats("gemm_half")
        dtheta = max(float(np.max(np.abs(np.asarray(a["theta"]) - np.asarray(b["theta"]))
                                  / np.maximum(np.abs(np.asarray(a["theta"])), 1e-12)))
                     for a, b in zip(nat, r32))
        fp32_mode = {"fits_per_hour": len(jobs) / s32 * 3600.0, "seconds_per_step": s32,
                     "gemm_tuned": {k: t32[k] for k in ("rt", "pd", "nsplit", "ms")},
                     "gemm_full_avg_ms": (g1 - g0) / max(1, n1 - n0_) * 1e3,
                     "gemm_half_avg_ms": (h1 - h0) / max(1, m1 - m0) * 1e3,
                     "max_rel_theta_diff_vs_split_mode": dtheta}
        log(f"FP32-MFMA mode: {s32:.2f}s/step, {fp32_mode['fits_per_hour']:.0f} fits/h, GEMM full "
            f"{fp32_mode['gemm_full_avg_ms']:.2f} ms, half {fp32_mode['gemm_half_avg_ms']:.2f} ms, "
            f"max rel theta diff vs {args.gemm_mode} {dtheta:.2e}")
        sess.set_gemm_mode(args.gemm_mode)
        sess.autotune(reps=3)                      # full and one-half shapes of this mode again

    plain = None
    if not args.no_plain:
        g2, n2 = sess.engine_stats("gemv")
        tp = time.perf_counter()
        pr = D.fit_plain(sess, "binary", phenos[2], X, opt)
        plain_s = time.perf_counter() - tp
        g3, n3 = sess.engine_stats("gemv")
        gemv_avg = (g3 - g2) / max(1, n3 - n2)
        # GB/s of S the product needs: its lower triangle (2 N^2 bytes) in the
        # default mode, all of it (4 N^2) in the "full" one
        gemv_bytes = (2.0 if sess.gemv_mode() == "tri" else 4.0) * N * N
        plain = {"seconds": plain_s, "fits_per_hour": 3600.0 / plain_s, "pcg_iters": pr["pcg_iters"],
                 "grm_products": pr["pcg_iters"] + pr["n_products"],
                 "n_solves": pr["n_solves"], "gemv_ms": gemv_avg * 1e3, "gemv_mode": sess.gemv_mode(),
                 "gemv_GBps": gemv_bytes / gemv_avg / 1e9, "theta": list(map(float, pr["theta"]))}
        log(f"plain: {plain_s:.2f}s per fit, {plain['fits_per_hour']:.1f} fits/h, GEMV {gemv_avg * 1e3:.2f} ms "
            f"({plain['gemv_GBps']:.0f} GB/s), theta {pr['theta']}")

    # the reference's default of 30 probes per trace round (step1_fitNULLGLMM.R:67)
    nrun30 = None
    if not args.no_nrun30:
        o30 = D.Options(nrun=30)
        t0 = time.perf_counter()
        native_step(o30)
        n30 = time.perf_counter() - t0
        t0 = time.perf_counter()
        step(o30)
        b30 = time.perf_counter() - t0
        tp = time.perf_counter()
        D.fit_plain(sess, "binary", phenos[2], X, o30)
        p30 = time.perf_counter() - tp
        nrun30 = {"fits_per_hour": len(jobs) / n30 * 3600.0, "python_fits_per_hour": len(jobs) / b30 * 3600.0,
                  "plain_fits_per_hour": 3600.0 / p30, "batched_vs_plain": p30 * len(jobs) / n30,
                  "python_vs_plain": p30 * len(jobs) / b30}
        log(f"nrun 30: C++ {nrun30['fits_per_hour']:.0f}, Python {nrun30['python_fits_per_hour']:.0f}, "
            f"plain {nrun30['plain_fits_per_hour']:.0f} fits/h")

    serving = saturation = None
    if args.serve_requests > 0:
        # saturation: a burst at four times the batch path's capacity gives the
        # server's own throughput on this request mix; the latency leg then
        # offers the C3 stream of requests at 0.8 of it, so p50/p99 measure
        # service and modest queueing rather than a backlog that grows with
        # the length of the run
        capacity = len(jobs) / nat_s
        n_sat = max(1, args.serve_requests * 3 // 10)
        saturation = serve_leg(sess, N, kept["Xraw"], kept["Zc"], n_sat, 4.0 * capacity, seed + 6, opt)
        log(f"serving saturation: {saturation['completed']}/{saturation['requests']} fits offered at "
            f"{saturation['offered_rate_per_s']:.2f}/s: {saturation['fits_per_hour']:.0f} fits/h")
        rate = 0.8 * saturation["fits_per_hour"] / 3600.0
        serving = serve_leg(sess, N, kept["Xraw"], kept["Zc"], args.serve_requests, rate, seed + 7, opt)
        serving["load_vs_saturation"] = 0.8
        log(f"serving: {serving['completed']}/{serving['requests']} fits at offered {serving['offered_rate_per_s']:.2f}/s: "
            f"{serving['fits_per_hour']:.0f} fits/h, p50 {serving['p50_latency_s']:.2f}s, "
            f"p99 {serving['p99_latency_s']:.2f}s")

    fp64 = None
    if G0 is not None:
        fp64 = fp64_check_leg(sess, G0, miss0, kept["Xraw"], D.Options(nrun=30), seed + 11, K_single)
        log(f"FP64 subblock check ({n0} samples): GRM {fp64['grm_max_rel_error']:.2e} (off-diagonal "
            f"{fp64['grm_offdiag_rel_error']:.2e}), block solve "
            f"{fp64['block_solve_rel_error']['max']:.2e}, tau rel error " +
            ", ".join(f"{k} {v['rel_error']}" for k, v in fp64["tau"].items()))
        del G0, miss0

    loco = None
    if args.loco_chromosomes > 0:
        # the QC'd markers cut into autosomes in proportion to their lengths (Mb)
        lengths = np.array([248, 242, 198, 190, 181, 171, 159, 145, 138, 134, 135, 133, 114, 107, 102, 90,
                            83, 80, 59, 64, 47, 51], np.float64)
        bounds = np.round(np.cumsum(lengths) / lengths.sum() * Mqc).astype(np.int64)
        chr_vec = np.repeat(np.arange(1, 23), np.diff(np.r_[0, bounds]))
        _, starts, ends = D.update_chr_start_end_index(chr_vec)
        nc = args.loco_chromosomes
        starts = starts[:nc] + [None] * (22 - nc)
        ends = ends[:nc] + [None] * (22 - nc)
        ljobs = [{"fit": h.result, "trait": "binary", "y": j["y"], "X": j["X"]} for h, j in zip(handles, jobs)]
        import collections
        tm = collections.Counter()
        tl = time.perf_counter()
        lres = D.loco(sess, ljobs, starts, ends, opt, batched=True, n_systems=len(ljobs), timing=tm)
        loco_s = time.perf_counter() - tl
        loco = {"chromosomes": nc, "fits": len(ljobs), "seconds": loco_s,
                "seconds_per_chromosome": loco_s / nc,
                "set_range_s_per_chromosome": tm["set_range"] / nc,
                "coef_s_per_chromosome": tm["coef"] / nc,
                "loco_fits_per_hour": len(ljobs) * nc / loco_s * 3600.0,
                "max_abs_coef_shift": float(max(np.max(np.abs(r["coefficients"] - h.result["coefficients"]))
                                                for rr, h in zip(lres, handles) for r in rr if r["isLOCO"]))}
        log(f"LOCO: {nc} chromosomes x {len(ljobs)} fits in {loco_s:.1f}s "
            f"({tm['set_range'] / nc:.3f}s range build + {tm['coef'] / nc:.3f}s Get_Coef_LOCO per chromosome)")

    cpu = None
    if not args.no_cpu:
        # the box's CPU share: the cpuset this process may run on
        threads = len(os.sched_getaffinity(0))
        n_matvec = plain["grm_products"] if plain else 3000
        legs = {}
        for nt in (threads, 1):
            per_marker, ms, reps = cpu_baseline(sess, dev, N, M, n_matvec, nt, args.cpu_sample_markers,
                                                min_seconds=8.0)
            t_mv = per_marker * Mqc
            legs[nt] = (t_mv, 3600.0 / (n_matvec * t_mv), ms, reps)
            log(f"cpu baseline: {t_mv:.2f}s per K.v with {nt} threads -> {legs[nt][1]:.4f} fits/h (extrapolated)")
        c0 = {}
        for nt in (threads, 1):
            secs, fit = cpu_c0_fit(nt, D.Options(nrun=30))
            c0[str(nt)] = {"threads": nt, "seconds": secs, "fits_per_hour": 3600.0 / secs,
                           "theta": list(map(float, fit["theta"])), "iterations": int(fit["iterations"])}
            log(f"cpu C0 end to end ({nt} threads): {secs:.1f}s, theta {fit['theta']}")
        t_mv, fph, ms, reps = legs[threads]
        cpu = {"value": fph, "unit": "fits/hour", "cores": threads, "kind": "port",
               "sample": (f"oracle parallelCrossProdOpenMP restatement (per-marker 2-bit decode + FP32 "
                          f"dot/axpy) over {ms} of the {Mqc} QC'd markers at N={N}, {reps} reps, "
                          f"{threads} threads (the process's CPU set; os.cpu_count() {os.cpu_count()}): "
                          f"{t_mv:.2f}s per K.v, x {n_matvec} GRM products of the plain GPU fit (extrapolated)"),
               "one_thread": {"value": legs[1][1], "seconds_per_matvec": legs[1][0],
                              "note": "OMP_NUM_THREADS=1 as the reference's Dockerfile ships it"},
               "c0_end_to_end": c0}

    sess.free_device(dev)
    traffic = PROFILED_GEMM_TRAFFIC.get((args.gemm_mode, N, tune["rt"], tune["pd"], tune["nsplit"]))
    common = {"traffic": traffic,
              "traffic_source": PROFILE_FILES[args.gemm_mode] + " (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)",
              "avg_launch_ms": gemm_avg * 1e3, "launches": gemm_launches}
    # the algorithmic unit: 2 N^2 k flop of an FP32 product with k = 64
    # right-hand sides, against the FP32 matrix peak; the split modes reach
    # it through BF16 pieces, whose MFMA count is reported beside it
    def narrow(avg, launches, k):
        # a launch over k of the 64 columns: 2 N^2 k flop
        if launches == 0:
            return {"launches": 0}
        return {"avg_ms": avg * 1e3, "launches": launches, "achieved": 2.0 * N * N * k / avg / 1e12,
                "unit": "TFLOP/s", "peak": FP32_MFMA_PEAK, "frac": 2.0 * N * N * k / avg / 1e12 / FP32_MFMA_PEAK,
                "algorithmic": f"2*N^2*{k} = {2.0 * N * N * k:.3e} flop per launch"}

    roofline = {"bound": "mfma", "achieved": gemm_tflops, "peak": FP32_MFMA_PEAK, "unit": "TFLOP/s",
                "frac": gemm_tflops / FP32_MFMA_PEAK,
                "algorithmic": f"2*N^2*64 = {gemm_flop:.3e} FP32 flop per launch (64 right-hand sides)",
                "half_width": narrow(half_avg, half_launches, 32),
                "quarter_width": narrow(quarter_avg, quarter_launches, 16), **common}
    if args.gemm_mode == "fp32":
        roofline["kernel"] = "k_gemm_part + k_gemm_epi (block-PCG GRM x 64 RHS, FP32 MFMA)"
        roofline["bytes"] = f"4*N^2 = {4.0 * N * N:.3e} B of S per launch"
    elif args.gemm_mode == "bf16x3_tri":
        # on the split GRM (stored at 16 significant bits) S has two pieces:
        # five piece products per FP32 product instead of six
        npc = 5 if sess.grm_precision() == "bf16_split" else 6
        roofline["kernel"] = ("k_split_p + k_gemm_t16 + k_tri_reduce + k_gemm_epi (block-PCG GRM x 64 RHS from "
                              "the lower triangle on 16x16x32 MFMA tiles, P as three BF16 pieces, S as " +
                              ("two (the split GRM's 16-bit values)" if npc == 5 else "three") + ")")
        roofline["bytes"] = f"2*N^2 = {2.0 * N * N:.3e} B of S's lower triangle per launch"
        roofline["hbm_GBs_of_triangle"] = 2.0 * N * N / gemm_avg / 1e9
        roofline["bf16_piece_products"] = {"tflops": npc * gemm_flop / gemm_avg / 1e12, "peak": BF16_MFMA_PEAK,
                                           "frac": npc * gemm_flop / gemm_avg / 1e12 / BF16_MFMA_PEAK,
                                           "note": f"{npc} BF16 piece products per FP32 product (the emulation's "
                                                   "cost)"}
        roofline["note"] = ("FP32-accurate products from BF16 piece products on the BF16 MFMA: the FP32-flop rate "
                            "may exceed the FP32 MFMA peak (frac > 1); the matrix unit the kernel runs on is the "
                            "BF16 one (bf16_piece_products)")
    else:
        roofline["kernel"] = ("k_split_p + k_gemm_x3w (rt 1) or k_gemm_x3 (rt 2) + k_gemm_epi (block-PCG GRM x 64 "
                              "RHS, FP32 operands as three BF16 pieces on the BF16 MFMA)")
        roofline["bytes"] = f"4*N^2 = {4.0 * N * N:.3e} B of S per launch"
        roofline["hbm_GBs_of_S"] = 4.0 * N * N / gemm_avg / 1e9
        roofline["bf16_piece_products"] = {"tflops": 6.0 * gemm_flop / gemm_avg / 1e12, "peak": BF16_MFMA_PEAK,
                                           "frac": 6.0 * gemm_flop / gemm_avg / 1e12 / BF16_MFMA_PEAK}
    out = {
        "metric": "null-GLMM fits/hour at N=100k batched vs plain; GRM TFLOP/s; p99 fit latency",
        "value": fits_per_hour,
        "unit": "fits/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device-generated PLINK .bed, simulated phenotypes)",
        "config": {"workload": f"batched binary null fits through the C++ drivers (saige_fit_null_batch), N={N}, "
                               f"M={M}, 4 phenotypes x (nrun {args.nrun} + 1) = 64-RHS block PCG, {args.grm} GRM "
                               f"(FP32 storage)",
                   "N": N, "M": M, "Mqc": Mqc, "phenotypes": len(jobs), "nrun": args.nrun, "grm": args.grm,
                   "gemm_mode": args.gemm_mode},
        "grm_tflops": grm_tflops,
        "grm_seconds": grm_s,
        "grm_tflops_pass": "single BF16 pass" if args.grm != "fp32" else "FP32",
        "grm_split": split,
        "setgeno_seconds": setgeno_s,
        "gemm_autotune": tune,
        "python_driver": python_driver,
        "fp32_mode": fp32_mode,
        "plain": plain,
        "batched_vs_plain": (fits_per_hour / plain["fits_per_hour"]) if plain else None,
        "python_vs_plain": (python_driver["fits_per_hour"] / plain["fits_per_hour"]) if plain else None,
        "nrun30": nrun30,
        "p50_fit_latency_s": serving["p50_latency_s"] if serving else float(np.percentile(lat, 50)),
        "p99_fit_latency_s": serving["p99_latency_s"] if serving else float(np.percentile(lat, 99)),
        "serving": serving,
        "serving_saturation": saturation,
        "fp64_subblock_check": fp64,
        "loco": loco,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
