"""Synthetic SAIGE step-1 inputs: genotypes g ~ Binomial(2, p_j) with
missing calls, covariates (intercept, age, sex) and binary / quantitative
phenotypes with a polygenic effect.  Small cohorts are drawn with numpy;
large ones on the device (saige_synth_bed_device) so the .bed body never
leaves HBM."""
import numpy as np

from .plink import pack_bed

SEED_BASE = 20261202


def config_seed(index):
    return SEED_BASE + 1000 * int(index)


def genotypes(N, M, maf_lo, maf_hi, miss_rate, seed):
    """N x M A1 counts, missing mask and allele frequencies (numpy)."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(maf_lo, maf_hi, M)
    G = rng.binomial(2, p[None, :], size=(N, M)).astype(np.int8)
    miss = rng.random((N, M)) < miss_rate
    return G, miss, p


def covariates(N, seed):
    rng = np.random.default_rng(seed + 1)
    age = rng.normal(50.0, 10.0, N)
    sex = (rng.random(N) < 0.5).astype(np.float64)
    return np.column_stack([np.ones(N), age, sex])


def standardized(G, miss):
    """Column-standardised genotypes with mean imputation (for simulating the
    polygenic component only)."""
    Gf = np.where(miss, np.nan, G.astype(np.float64))
    mu = np.nanmean(Gf, axis=0)
    Gf = np.where(np.isnan(Gf), mu[None, :], Gf)
    sd = Gf.std(axis=0)
    sd[sd == 0] = 1.0
    return (Gf - mu) / sd


def polygenic(Zc, h2, seed):
    """g = Z_c b, b ~ N(0, h2 / ncausal), rescaled to var(g) = h2."""
    rng = np.random.default_rng(seed + 2)
    b = rng.normal(0.0, np.sqrt(h2 / Zc.shape[1]), Zc.shape[1])
    g = Zc @ b
    g = g - g.mean()
    sd = g.std()
    return g * (np.sqrt(h2) / sd) if sd > 0 else g


def binary_phenotype(X, g, prevalence, seed, beta=(0.0, 0.02, 0.3)):
    """Logistic model logit P(y=1) = a + X[:,1:] beta + g with the intercept a
    chosen by bisection so the expected prevalence matches."""
    rng = np.random.default_rng(seed + 3)
    lin = X[:, 1:] @ np.asarray(beta[1:X.shape[1]]) + g
    lin = lin - lin.mean()
    lo, hi = -30.0, 30.0
    for _ in range(100):
        a = 0.5 * (lo + hi)
        if np.mean(1.0 / (1.0 + np.exp(-(a + lin)))) < prevalence:
            lo = a
        else:
            hi = a
    pr = 1.0 / (1.0 + np.exp(-(a + lin)))
    return (rng.random(len(pr)) < pr).astype(np.float64)


def quantitative_phenotype(X, g, h2, seed, beta=(0.0, 0.02, 0.3)):
    rng = np.random.default_rng(seed + 4)
    e = rng.normal(0.0, np.sqrt(1.0 - h2), X.shape[0])
    return X[:, 1:] @ np.asarray(beta[1:X.shape[1]]) + g + e


def small_cohort(N, M, maf_lo=0.05, maf_hi=0.5, miss_rate=0.005, seed=SEED_BASE, ncausal=None):
    """A complete small cohort in host memory: .bed body, covariates and the
    standardised causal genotypes for phenotype simulation."""
    G, miss, p = genotypes(N, M, maf_lo, maf_hi, miss_rate, seed)
    body = pack_bed(G, miss)
    X = covariates(N, seed)
    rng = np.random.default_rng(seed + 5)
    nc = min(M, ncausal or max(1, M // 10))
    causal = np.sort(rng.choice(M, nc, replace=False))
    Zc = standardized(G[:, causal], miss[:, causal])
    return {"N": N, "M": M, "bed": body, "G": G, "miss": miss, "p": p, "X": X, "Zc": Zc,
            "causal": causal, "seed": seed}


def decode_bed_rows(body_rows, N):
    """Decode packed PLINK rows (k x ceil(N/4)) to A1 counts (N x k) and a
    missing mask, following the PLINK-1 code table."""
    rows = np.asarray(body_rows, np.uint8)
    k = rows.shape[0]
    codes = np.stack([(rows >> (2 * j)) & 3 for j in range(4)], axis=2).reshape(k, -1)[:, :N]
    lut = np.array([2, 0, 1, 0], np.int8)
    G = lut[codes].T
    miss = (codes == 1).T
    return G, miss
