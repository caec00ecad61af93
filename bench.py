"""Benchmark: null-GLMM fits/hour at N=100k, batched vs plain, plus GRM
TFLOP/s, on one MI355X.

Workload (default): N = 100,000 samples, M = 150,000 markers, MAF ~
U(0.01, 0.5), 1% missing calls, synthetic PLINK .bed generated in HBM;
intercept + age + sex covariates (QR-transformed as fitNULLGLMM does);
four binary phenotypes with prevalences 0.05, 0.1, 0.2, 0.3.  The GRM is
built once from BF16 standardised genotypes on MFMA with FP32 accumulation
and held in HBM as FP32.

A step = one batched round of null fits of the four phenotypes: AI-REML
with nrun = 15 Rademacher probes per phenotype, so each trace phase is a
block PCG over 4 x (15 probes + 1) = 64 right-hand sides (FP32 vectors and
FP32 MFMA GRM products, FP64 dot products).  `value` = fits completed per
hour of step time.  The plain path (one phenotype, one right-hand side at a
time, GEMV per PCG iteration) is timed once beside it.

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
BF16_MFMA_PEAK = 2500.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(sess, N, M, maf_lo, maf_hi, miss, seed, prevalences, ncausal=1000, h2=0.3):
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
    return dev, X, phenos


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=100_000)
    ap.add_argument("--M", type=int, default=150_000)
    ap.add_argument("--nrun", type=int, default=15)
    ap.add_argument("--grm", default="bf16", choices=["bf16", "fp32"])
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample-markers", type=int, default=4000)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from saige_gpu import Session
    from saige_gpu import driver as D

    N, M = args.N, args.M
    seed = 20261202 + 1000 * 2
    prevalences = [0.05, 0.1, 0.2, 0.3]
    sess = Session(0)
    t = time.perf_counter()
    dev, X, phenos = make_workload(sess, N, M, 0.01, 0.5, 0.01, seed, prevalences)
    log(f"workload generated in {time.perf_counter() - t:.1f}s")
    sess.setminMAFforGRM(0.01)
    sess.setmaxMissingRateforGRM(0.15)
    t = time.perf_counter()
    sess.setgeno_device(dev, N, M)
    log(f"setgeno (decode/QC/repack/diag) {time.perf_counter() - t:.2f}s, Mqc={sess.getMsub_MAFge_minMAFtoConstructGRM()}")
    Mqc = sess.getMsub_MAFge_minMAFtoConstructGRM()
    grm_s = sess.build_grm(args.grm)
    grm_tflops = N * (N + 1) * Mqc / grm_s / 1e12
    log(f"GRM {args.grm}: {grm_s:.3f}s device time, {grm_tflops:.1f} TFLOP/s")

    opt = D.Options(nrun=args.nrun)
    jobs = [{"trait": "binary", "y": y, "X": X, "tag": f"prev{p}"} for y, p in zip(phenos, prevalences)]

    def step():
        handles, sch = D.fit_batched(sess, jobs, opt, width=64, n_systems=len(jobs))
        for h in handles:
            if h.error is not None:
                raise h.error
        return handles, sch

    for _ in range(args.warmup):
        step()
    g0, n0 = sess.engine_stats()
    t0 = time.perf_counter()
    results = []
    for _ in range(args.steps):
        results.append(step())
    t1 = time.perf_counter()
    g1, n1 = sess.engine_stats()
    step_s = (t1 - t0) / args.steps
    fits = len(jobs) * args.steps
    fits_per_hour = fits / (t1 - t0) * 3600.0
    handles, sch = results[-1]
    gemm_launches = n1 - n0
    gemm_avg = (g1 - g0) / max(1, gemm_launches)
    gemm_flop = 2.0 * N * N * 64
    gemm_tflops = gemm_flop / gemm_avg / 1e12
    lat = sorted(h.t_done - h.t_submit for h in handles)
    log("batched step split (s): " + ", ".join(f"{k} {v:.3f}" for k, v in sorted(sch.timing.items())))
    log(f"batched: {step_s:.2f}s/step, {fits_per_hour:.1f} fits/h, {gemm_launches} GEMMs avg "
        f"{gemm_avg * 1e3:.2f} ms ({gemm_tflops:.1f} TFLOP/s, {4.0 * N * N / gemm_avg / 1e9:.0f} GB/s), "
        f"theta {[list(np.round(h.result['theta'], 4)) for h in handles]}")

    plain = None
    if not args.no_plain:
        g2, n2 = sess.engine_stats()
        tp = time.perf_counter()
        pr = D.fit_plain(sess, "binary", phenos[2], X, opt)
        plain_s = time.perf_counter() - tp
        g3, n3 = sess.engine_stats()
        gemv_avg = (g3 - g2) / max(1, n3 - n2)
        plain = {"seconds": plain_s, "fits_per_hour": 3600.0 / plain_s, "pcg_iters": pr["pcg_iters"],
                 "grm_products": pr["pcg_iters"] + pr["n_products"],
                 "n_solves": pr["n_solves"], "gemv_ms": gemv_avg * 1e3,
                 "gemv_GBps": 4.0 * N * N / gemv_avg / 1e9, "theta": list(map(float, pr["theta"]))}
        log(f"plain: {plain_s:.1f}s per fit, {plain['fits_per_hour']:.1f} fits/h, GEMV {gemv_avg * 1e3:.2f} ms "
            f"({plain['gemv_GBps']:.0f} GB/s), theta {pr['theta']}")

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        n_matvec = plain["grm_products"] if plain else 3000
        per_marker, ms, reps = cpu_baseline(sess, dev, N, M, n_matvec, threads, args.cpu_sample_markers)
        t_mv = per_marker * Mqc
        cpu_fph = 3600.0 / (n_matvec * t_mv)
        cpu = {"value": cpu_fph, "unit": "fits/hour", "cores": threads, "kind": "port",
               "sample": (f"oracle parallelCrossProdOpenMP restatement (per-marker 2-bit decode + FP32 "
                          f"dot/axpy) over {ms} of the {Mqc} QC'd markers at N={N}, {reps} reps, "
                          f"{threads} threads: {t_mv:.2f}s per K.v, x {n_matvec} GRM products of the plain "
                          f"GPU fit (extrapolated)")}
        log(f"cpu baseline: {t_mv:.2f}s per K.v with {threads} threads -> {cpu_fph:.4f} fits/h (extrapolated)")

    sess.free_device(dev)
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
        "config": {"workload": f"batched binary null fits, N={N}, M={M}, 4 phenotypes x (nrun {args.nrun} + 1) "
                               f"= 64-RHS block PCG, {args.grm.upper()} GRM (FP32 storage)",
                   "N": N, "M": M, "Mqc": Mqc, "phenotypes": len(jobs), "nrun": args.nrun, "grm": args.grm},
        "grm_tflops": grm_tflops,
        "grm_seconds": grm_s,
        "plain": plain,
        "batched_vs_plain": (fits_per_hour / plain["fits_per_hour"]) if plain else None,
        "p50_fit_latency_s": float(np.percentile(lat, 50)),
        "p99_fit_latency_s": float(np.percentile(lat, 99)),
        "roofline": {"bound": "mfma", "kernel": "k_gemm (block-PCG GRM x 64 RHS, FP32 MFMA)",
                     "achieved": gemm_tflops, "peak": FP32_MFMA_PEAK, "unit": "TFLOP/s",
                     "frac": gemm_tflops / FP32_MFMA_PEAK, "traffic": None,
                     "algorithmic": f"2*N^2*64 = {gemm_flop:.3e} flop per launch; 4*N^2 = {4.0 * N * N:.3e} B",
                     "avg_launch_ms": gemm_avg * 1e3, "launches": gemm_launches},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
