"""Workloads of the five benchmark configurations (C0-C4) for the GPU tests,
and the FP64 references they are checked against where the full problem is
too large for a dense FP64 computation:

* the GRM of a principal subblock (samples 0..n0-1) computed in FP64 from
  those samples' genotypes over every GRM marker, with the cohort's allele
  frequencies and the imputation of FG.cpp:447-454 (BLAS on the host);
* the residual ||Sigma x - b||^2 of a device solve, evaluated in FP64 on the
  GPU's own GRM read back in row blocks (the stopping rule of
  FG.cpp:2696-2792 bounds the recurrence's residual by tolPCG);
* dense FP64 fits on the subblock through the oracle's drivers.

Seeds follow 20261202 + 1000 * config index; every stream derives from it.
"""
import numpy as np

from saige_gpu import synth

SEED = {k: synth.config_seed(k) for k in range(5)}


def device_cohort(sess, N, M, maf_lo, maf_hi, miss, pair_frac, seed):
    """A PLINK .bed body generated in HBM (Binomial(2, p_j) calls, p_j ~
    U(maf_lo, maf_hi), i.i.d. missing calls, the first 2*pair_frac*N/2
    samples in sibling pairs).  Returns the device pointer."""
    bpm = (N + 3) // 4
    dev = sess.alloc_device(M * bpm)
    sess.synth_bed_device(dev, N, M, maf_lo, maf_hi, miss, pair_frac, seed)
    return dev


def subblock_genotypes(sess, dev, N, M, n0, chunk=8192):
    """A1 counts (n0 x M, int8) and the missing mask of samples 0..n0-1,
    decoded on the host from the device .bed (n0 a multiple of 4)."""
    assert n0 % 4 == 0 and n0 <= N
    bpm = (N + 3) // 4
    G = np.empty((n0, M), np.int8)
    miss = np.empty((n0, M), bool)
    for m0 in range(0, M, chunk):
        k = min(chunk, M - m0)
        raw = sess.copy_device_to_host(dev, k * bpm, offset=m0 * bpm).reshape(k, bpm)[:, : n0 // 4]
        g, mm = synth.decode_bed_rows(raw, n0)
        G[:, m0:m0 + k] = g
        miss[:, m0:m0 + k] = mm
    return G, miss


def fp64_grm_subblock(G, miss, stats, grm_mask, chunk=16384):
    """K64 of the subblock: z = (g - 2f) / sqrt(2f(1-f)) in FP64 over the GRM
    markers, missing calls imputed to round(2 p-hat) (the stats' fill),
    f the marker's frequency after imputation; K = Z Z^T / Mqc."""
    cols = np.nonzero(grm_mask)[0]
    n0 = G.shape[0]
    S = np.zeros((n0, n0))
    for c0 in range(0, len(cols), chunk):
        c = cols[c0:c0 + chunk]
        f = stats["freq"][c].astype(np.float64)
        g = np.where(miss[:, c], stats["fill"][c][None, :], G[:, c]).astype(np.float64)
        Z = (g - 2.0 * f) / np.sqrt(2.0 * f * (1.0 - f))
        S += Z @ Z.T
    return S / len(cols)


def grm_rel_error(K, K64):
    return float(np.max(np.abs(K.astype(np.float64) - K64)) / np.max(np.abs(K64)))


def grm_offdiag_error(K, K64):
    """||offdiag(K - K64)||_F / ||offdiag(K64)||_F: the relatedness signal
    that tau1 is estimated from (about 1/sqrt(Mqc) for unrelated samples,
    far below the diagonal's 1)."""
    D = K.astype(np.float64) - K64
    np.fill_diagonal(D, 0.0)
    R = K64.copy()
    np.fill_diagonal(R, 0.0)
    return float(np.linalg.norm(D) / np.linalg.norm(R))


def fp64_residual2(sess, w, tau, X, B, rows=2048):
    """||Sigma x_c - b_c||^2 per column in FP64, Sigma = tau0 diag(1/w) +
    tau1 K with K the GPU's own GRM (FP32, read back in row blocks); w and
    tau rounded to FP32 as the device holds them."""
    w = np.asarray(w, np.float32).astype(np.float64)
    t0, t1 = (float(np.float32(t)) for t in tau)
    X = np.asarray(X, np.float64).reshape(len(w), -1)
    R = t0 * X / w[:, None] - np.asarray(B, np.float64).reshape(len(w), -1)
    N = len(w)
    for r0 in range(0, N, rows):
        nr = min(rows, N - r0)
        R[r0:r0 + nr] += t1 * (sess.grm_block(r0, 0, nr, N).astype(np.float64) @ X)
    return np.sum(R * R, axis=0)


def causal_std(G, miss, n_causal, seed):
    """Standardised genotypes of n_causal random markers (for phenotype
    simulation only)."""
    rng = np.random.default_rng(seed + 5)
    causal = np.sort(rng.choice(G.shape[1], n_causal, replace=False))
    return synth.standardized(G[:, causal], miss[:, causal])


def covariate_offset(y, Xt, trait):
    """fitNULLGLMM with isCovariateOffset = TRUE, the step-1 CLI default
    (FG.R:1043-1065, 1118-1127; step1_fitNULLGLMM.R:117): glm() on the
    transformed covariates, their non-intercept part X[, -1] beta[-1] as the
    offset, and an intercept-only GLMM.  Returns (X1, offset)."""
    from oracle import oracle as O
    fam = O.Binomial if trait == "binary" else O.Gaussian
    beta = O.glm_fit(y, Xt, fam)["coef"]
    return np.ones((len(y), 1)), Xt[:, 1:] @ beta[1:]


def causal_std_device(sess, dev, N, M, n_causal, seed):
    """Standardised genotypes of all N samples at n_causal random markers,
    read from the device .bed (for phenotype simulation only)."""
    bpm = (N + 3) // 4
    rng = np.random.default_rng(seed)
    causal = np.sort(rng.choice(M, n_causal, replace=False))
    rows = np.stack([sess.copy_device_to_host(dev, bpm, offset=int(j) * bpm) for j in causal])
    G, mm = synth.decode_bed_rows(rows, N)
    return synth.standardized(G, mm)


def engine_solve(sess, w, tau, B, mask=None, maxiter=500, tol=1e-5):
    """Sigma^-1 B through the block engine (64 slots, the session's block
    GEMM) on the samples of mask (K_SS of the resident GRM; w and B are
    full-length, zero outside the mask), k <= 64 columns.  Returns
    (X, iterations)."""
    B = np.asarray(B, np.float64).reshape(len(w), -1)
    k = B.shape[1]
    assert k <= 64
    sess.engine_reserve(64, 1)
    sess.engine_set_system(0, w, tau, mask)
    sess.engine_add_batch(list(range(k)), [0] * k, B, maxiter, tol)
    while True:
        done, iters = sess.engine_step(1)
        if np.all(done[:k]):
            break
    return sess.engine_take_batch(list(range(k))).astype(np.float64), np.asarray(iters[:k])
