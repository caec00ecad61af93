/*
 * CPU ORACLE — test infrastructure only.
 *
 * A plain-C restatement of the arithmetic of SAIGE step 1's hot path as the
 * reference implements it in src/SAIGE/src/SAIGE_fitGLMM_fast.cpp (abbreviated
 * FG.cpp below).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library; the product path never does.
 *
 * Every function cites the reference lines it follows.  FP32 arithmetic is
 * kept in the reference's operation order, with no fused multiply-add
 * (built with -ffp-contract=off), so that each rounding step matches.
 *
 * Pins (what checks each function, see tests/test_oracle_*.py):
 *   orc_setgeno / orc_decode_marker : PLINK .acount ALT_CTS and .frq MAF
 *       fixtures held by the reference (bit-exact counts), plus a brute-force
 *       per-sample decoder written from the PLINK-1 spec on inputs with
 *       missing calls and sample subsets.
 *   orc_std_geno, orc_diag_std      : numpy float32 evaluation of the same
 *       closed form (bit-exact) and FP64 closed form (tolerance).
 *   orc_crossprod                   : FP64 dense Z Z^T b / M (tolerance).
 *   orc_diag_sigma, orc_sigma_mul   : numpy float32 elementwise (bit-exact).
 *   orc_pcg                         : dense FP64 solve of Sigma x = b
 *       (tolerance) and the residual invariant |Sigma x - b|^2 <= tol.
 *   parity unpinned: the order of FP32 summation inside BLAS sdot and the
 *       OpenMP reduction (FG.cpp:1582) -- third-party and schedule dependent.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- decode */

/* Per-marker statistics produced by Get_OneSNP_Geno_atBeginning
 * (FG.cpp:326-579). */
typedef struct {
    int32_t allele_count;   /* A1 allele count after imputation (FG.cpp:454) */
    int32_t num_missing;    /* missing calls among phenotyped samples (FG.cpp:376-385) */
    int32_t mac;            /* min(AC, 2*Nnm - AC) (FG.cpp:490) */
    int32_t fill;           /* int(round(2 p_hat)) (FG.cpp:447) */
    float   freq;           /* AC / (2 Nnm) after imputation (FG.cpp:484) */
    float   missing_rate;   /* miss / Nnm (FG.cpp:442) */
    float   inv_std;        /* 1/sqrt(2f(1-f)), 0 if the sd is 0 (FG.cpp:914-921) */
    uint8_t pass_qc;        /* maf >= minMAF && missingRate <= maxMissing, minus VR (FG.cpp:493-547) */
    uint8_t pass_vr;        /* chosen for variance-ratio estimation (FG.cpp:496-543) */
    uint8_t pad[2];
} orc_marker_stat;

/* 2-bit PLINK code -> number of A1 alleles, 3 = missing.
 * b = bit0, a = bit1 (FG.cpp:343-357): (0,0)->2, (1,0)->missing, (0,1)->1, (1,1)->0. */
static int code_to_geno(int code) {
    int b = code & 1, a = (code >> 1) & 1;
    if (b == 1 && a == 0) return 3;
    if (b == 0 && a == 0) return 2;
    if (b == 0 && a == 1) return 1;
    return 0;
}

/* Internal 2-bit code used after repacking (FG.cpp:41-44, 565-571):
 * 2 alleles -> 0b00 (HOM_REF), 1 -> 0b10 (HET), 0 -> 0b11 (HOM_ALT). */
static unsigned internal_code(int g) {
    if (g == 0) return 0x3u;
    if (g == 1) return 0x2u;
    return 0x0u;
}

static float float_min(float a, float b) { return (b < a) ? b : a; } /* std::min */

/*
 * Get_OneSNP_Geno_atBeginning (FG.cpp:326-579) for one marker.
 *   bytes      : ceil(N/4) bytes of the marker, SNP-major PLINK-1 layout
 *   indicator  : N flags, sample has a phenotype (indicatorGenoSamplesWithPheno)
 *   subidx     : Nnm 1-based sample indices into the .fam order (subSampleInGeno)
 *   vr_mode    : 0 no variance-ratio markers; 1 categorical (g_maxMACVarRatio != -1);
 *                2 single category (FG.cpp:497-541)
 *   in_randset : marker index is in the random VR pool g_randMarkerIndforVR
 *   repacked   : ceil(Nnm/4) output bytes in the internal code, or NULL
 */
void orc_decode_marker(const uint8_t* bytes, int64_t N, int64_t Nnm,
                       const uint8_t* indicator, const int32_t* subidx,
                       float minMAF, float maxMissing,
                       int vr_mode, float minMACvr, float maxMACvr, int in_randset,
                       orc_marker_stat* st, uint8_t* repacked)
{
    int alleleCount = 0, numMissing = 0;
    /* FG.cpp:338-420: decode every sample of the marker and count over the
     * phenotyped ones. */
    for (int64_t ind = 0; ind < N; ind++) {
        int code = (bytes[ind >> 2] >> ((ind & 3) * 2)) & 3;
        int g = code_to_geno(code);
        if (indicator[ind]) {
            if (g == 3) numMissing++;
            else alleleCount += g;
        }
    }
    /* FG.cpp:424: altFreq = alleleCount/float((Nnomissing-numMissing) * 2) */
    float altFreq = (float)alleleCount / (float)((size_t)(Nnm - numMissing) * 2);
    /* FG.cpp:442 */
    float missingRate = (float)numMissing / (float)(size_t)Nnm;
    /* FG.cpp:447: int(round(2*altFreq)); an all-missing marker gives NaN here,
     * which the reference converts with undefined behaviour; the oracle maps it
     * to 0 (such a marker always fails the missing-rate test below). */
    float twice = 2.0f * altFreq;
    int fill = isnan(twice) ? 0 : (int)roundf(twice);
    if (numMissing > 0) alleleCount = alleleCount + fill * numMissing; /* FG.cpp:454 */
    altFreq = (float)alleleCount / (float)((size_t)Nnm * 2);         /* FG.cpp:484 */

    float maf = float_min(altFreq, 1.0f - altFreq);                    /* FG.cpp:489 */
    int twoN = (int)Nnm * 2;
    int mac = (alleleCount < twoN - alleleCount) ? alleleCount : twoN - alleleCount; /* FG.cpp:490 */

    int passQC = (maf >= minMAF && missingRate <= maxMissing) ? 1 : 0; /* FG.cpp:493-495 */
    int passVR = 0;
    if (vr_mode == 1) {                                                 /* FG.cpp:497-523 */
        if (mac >= minMACvr && mac < maxMACvr) passVR = 1;
        else if (mac >= maxMACvr) passVR = in_randset ? 1 : 0;
    } else if (vr_mode == 2) {                                          /* FG.cpp:524-541 */
        if (mac >= minMACvr) passVR = in_randset ? 1 : 0;
    }
    if (passVR) passQC = 0;                                             /* FG.cpp:545-547 */

    /* FG.cpp:551-576: repack in subSampleInGeno order with imputation.  The
     * reference leaves the first packed byte's accumulator uninitialised
     * (FG.cpp:486); it is zero here. */
    if (repacked && (passQC || passVR)) {
        unsigned geno2 = 0;
        for (int64_t indx = 0; indx < Nnm; indx++) {
            int u = (int)(indx & 3);
            int64_t src = (int64_t)subidx[indx] - 1;
            int code = (bytes[src >> 2] >> ((src & 3) * 2)) & 3;
            int g = code_to_geno(code);
            if (g == 3) g = fill;
            geno2 |= internal_code(g) << (u << 1);
            if (u == 3 || indx == Nnm - 1) {
                repacked[indx >> 2] = (uint8_t)geno2;
                geno2 = 0;
            }
        }
    }

    /* FG.cpp:914-921: Std = std::sqrt(2*freq*(1-freq)); invStd = 1/Std or 0. */
    float Std = sqrtf(2.0f * altFreq * (1.0f - altFreq));
    float invStd = (Std == 0.0f) ? 0.0f : 1.0f / Std;

    st->allele_count = alleleCount;
    st->num_missing = numMissing;
    st->mac = mac;
    st->fill = fill;
    st->freq = altFreq;
    st->missing_rate = missingRate;
    st->inv_std = invStd;
    st->pass_qc = (uint8_t)passQC;
    st->pass_vr = (uint8_t)passVR;
}

/*
 * genoClass::setGenoObj marker loop (FG.cpp:739-1024) over an in-memory .bed
 * body (the 3 magic bytes already stripped).  Fills the per-marker stats for
 * all M markers and, for QC-passing (and VR) markers in marker order, the
 * repacked genotypes.  Returns the QC-passing marker count
 * (numberofMarkerswithMAFge_minMAFtoConstructGRM).
 */
int64_t orc_setgeno(const uint8_t* bed, int64_t N, int64_t M,
                    const uint8_t* indicator, const int32_t* subidx, int64_t Nnm,
                    float minMAF, float maxMissing,
                    int vr_mode, float minMACvr, float maxMACvr, const uint8_t* randset,
                    orc_marker_stat* stats, uint8_t* packed_qc, uint8_t* packed_vr,
                    int64_t* n_vr_out)
{
    int64_t bpm = (N + 3) / 4, bpm_nm = (Nnm + 3) / 4;
    int64_t mqc = 0, mvr = 0;
    for (int64_t i = 0; i < M; i++) {
        orc_marker_stat st;
        orc_decode_marker(bed + i * bpm, N, Nnm, indicator, subidx, minMAF, maxMissing,
                          vr_mode, minMACvr, maxMACvr, randset ? randset[i] : 0, &st, NULL);
        if (st.pass_qc && packed_qc)
            orc_decode_marker(bed + i * bpm, N, Nnm, indicator, subidx, minMAF, maxMissing,
                              vr_mode, minMACvr, maxMACvr, randset ? randset[i] : 0, &st,
                              packed_qc + mqc * bpm_nm);
        if (st.pass_vr && packed_vr)
            orc_decode_marker(bed + i * bpm, N, Nnm, indicator, subidx, minMAF, maxMissing,
                              vr_mode, minMACvr, maxMACvr, randset ? randset[i] : 0, &st,
                              packed_vr + mvr * bpm_nm);
        if (st.pass_qc) mqc++;
        if (st.pass_vr) mvr++;
        stats[i] = st;
    }
    if (n_vr_out) *n_vr_out = mvr;
    return mqc;
}

/* ------------------------------------------------------ standardisation */

/* setStdGenoLookUpArr (FG.cpp:152-159): L[g] = (g - 2 f) * invStd. */
void orc_std_lookup(float freq, float invStd, float* L3)
{
    float m2 = 2.0f * freq;
    L3[0] = (0.0f - m2) * invStd;
    L3[1] = (1.0f - m2) * invStd;
    L3[2] = (2.0f - m2) * invStd;
}

/* Get_OneSNP_Geno (FG.cpp:241-281): internal code -> 2-(a+b). */
void orc_get_geno(const uint8_t* packed, int64_t Nnm, int32_t* out)
{
    for (int64_t i = 0; i < Nnm; i++) {
        int c = (packed[i >> 2] >> ((i & 3) * 2)) & 3;
        int b = c & 1, a = (c >> 1) & 1;
        out[i] = 2 - (a + b);
    }
}

/* Get_OneSNP_StdGeno (FG.cpp:582-662). */
void orc_std_geno(const uint8_t* packed, int64_t Nnm, float freq, float invStd, float* out)
{
    float L[3];
    orc_std_lookup(freq, invStd, L);
    for (int64_t i = 0; i < Nnm; i++) {
        int c = (packed[i >> 2] >> ((i & 3) * 2)) & 3;
        int b = c & 1, a = (c >> 1) & 1;
        out[i] = L[2 - (a + b)];
    }
}

/* Get_Diagof_StdGeno (FG.cpp:665-704): m_DiagStd += z % z, markers in order. */
void orc_diag_std(const uint8_t* packed, int64_t Mqc, int64_t Nnm,
                  const float* freq, const float* invStd, float* diag)
{
    int64_t bpm = (Nnm + 3) / 4;
    float* z = (float*)malloc(sizeof(float) * (size_t)Nnm);
    memset(diag, 0, sizeof(float) * (size_t)Nnm);
    for (int64_t j = 0; j < Mqc; j++) {
        orc_std_geno(packed + j * bpm, Nnm, freq[j], invStd[j], z);
        for (int64_t i = 0; i < Nnm; i++) {
            float sq = z[i] * z[i];
            diag[i] = diag[i] + sq;
        }
    }
    free(z);
}

/* Dense standardised matrix Z (Nnm x Mqc, column-major like the reference's
 * GPU upload, FG.cpp:1915-1923). */
void orc_std_matrix(const uint8_t* packed, int64_t Mqc, int64_t Nnm,
                    const float* freq, const float* invStd, float* Z)
{
    int64_t bpm = (Nnm + 3) / 4;
    for (int64_t j = 0; j < Mqc; j++)
        orc_std_geno(packed + j * bpm, Nnm, freq[j], invStd[j], Z + j * Nnm);
}

/* ------------------------------------------------------------ GRM matvec */

/*
 * parallelCrossProdOpenMP + parallelCrossProd (FG.cpp:1576-1598, 1669-1708):
 *   out = (sum_j z_j (z_j . b)) / Mqc, decoding every marker on every call.
 * With nthreads == 1 the markers are summed in order with a sequential FP32
 * dot; with more threads each thread keeps its own accumulator and the
 * per-thread vectors are summed in thread order (the reference's OpenMP
 * reduction, whose order is schedule dependent).
 */
void orc_crossprod(const uint8_t* packed, int64_t Mqc, int64_t Nnm,
                   const float* freq, const float* invStd,
                   const float* b, float* out, int nthreads)
{
    int64_t bpm = (Nnm + 3) / 4;
    if (nthreads < 1) nthreads = 1;
    float* acc = (float*)calloc((size_t)Nnm * (size_t)nthreads, sizeof(float));
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        float* my = acc + (size_t)t * (size_t)Nnm;
        float* z = (float*)malloc(sizeof(float) * (size_t)Nnm);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t j = 0; j < Mqc; j++) {
            orc_std_geno(packed + j * bpm, Nnm, freq[j], invStd[j], z);
            float val1 = 0.0f;
            for (int64_t i = 0; i < Nnm; i++) val1 = val1 + z[i] * b[i];
            for (int64_t i = 0; i < Nnm; i++) my[i] = my[i] + val1 * z[i];
        }
        free(z);
    }
    for (int64_t i = 0; i < Nnm; i++) {
        float s = acc[i];
        for (int t = 1; t < nthreads; t++) s = s + acc[(size_t)t * (size_t)Nnm + i];
        out[i] = s / (float)Mqc;
    }
    free(acc);
}

/* ------------------------------------------------------------- Sigma ops */

/* getDiagOfSigma (FG.cpp:2322-2365):
 *   d = tau1 * diagStd / Mqc + tau0 / w   (or tau1 + tau0/w when the kinship
 *   diagonal is set to one), clamped so d >= 1e-4. */
void orc_diag_sigma(const float* diagStd, int64_t N, int64_t Mqc, int diag_one,
                    const float* w, float tau0, float tau1, float* out)
{
    for (int64_t i = 0; i < N; i++) {
        float d;
        if (!diag_one) d = (tau1 * diagStd[i]) / (float)Mqc + tau0 / w[i];
        else d = tau1 + tau0 / w[i];
        if (d < 1e-4f) d = 1e-4f;
        out[i] = d;
    }
}

/* getCrossprod (FG.cpp:2397-2425) given the GRM product kb = K b:
 *   out = tau0 * (b % (1/w)) + tau1 * kb, or only the first term if tau1 == 0. */
void orc_sigma_mul(const float* b, const float* kb, const float* w, int64_t N,
                   float tau0, float tau1, float* out)
{
    for (int64_t i = 0; i < N; i++) {
        float winv = 1.0f / w[i];
        float t = tau0 * (b[i] * winv);
        if (tau1 == 0.0f) out[i] = t;
        else out[i] = t + tau1 * kb[i];
    }
}

/* Dense FP32 K b with K stored row-major (Nnm x Nnm), summed in column
 * order; used by the oracle when the per-marker form is too slow. */
static void dense_matvec(const float* K, int64_t N, const float* b, float* out)
{
    for (int64_t i = 0; i < N; i++) {
        float s = 0.0f;
        const float* row = K + i * N;
        for (int64_t j = 0; j < N; j++) s = s + row[j] * b[j];
        out[i] = s;
    }
}

static float sdot(const float* a, const float* b, int64_t n)
{
    float s = 0.0f;
    for (int64_t i = 0; i < n; i++) s = s + a[i] * b[i];
    return s;
}

/*
 * getPCG1ofSigmaAndVector (FG.cpp:2593-2809), Jacobi preconditioner,
 * absolute ||r||^2 stop, all FP32.  The GRM product is either the
 * per-marker form (packed != NULL) or a dense FP32 K (Kdense != NULL,
 * already divided by Mqc).  The per-marker product runs on nthreads
 * OpenMP threads, as parallelCrossProdOpenMP does (FG.cpp:1576-1598).
 * Returns the iteration count.
 */
int orc_pcg(const uint8_t* packed, int64_t Mqc, const float* freq, const float* invStd,
            const float* Kdense, const float* diagStd, int diag_one,
            int64_t N, const float* w, float tau0, float tau1, const float* bvec,
            int maxiter, float tol, float* x, int nthreads)
{
    float* r = (float*)malloc(sizeof(float) * N);
    float* z = (float*)malloc(sizeof(float) * N);
    float* p = (float*)malloc(sizeof(float) * N);
    float* Ap = (float*)malloc(sizeof(float) * N);
    float* kp = (float*)malloc(sizeof(float) * N);
    float* minv = (float*)malloc(sizeof(float) * N);
    float* r1 = (float*)malloc(sizeof(float) * N);
    float* z1 = (float*)malloc(sizeof(float) * N);

    orc_diag_sigma(diagStd, N, Mqc, diag_one, w, tau0, tau1, minv);
    for (int64_t i = 0; i < N; i++) {
        minv[i] = 1.0f / minv[i];
        x[i] = 0.0f;
        r[i] = bvec[i];
        z[i] = minv[i] * r[i];
        p[i] = z[i];
    }
    float sumr2 = sdot(r, r, N);
    int iter = 0;
    while (sumr2 > tol && iter < maxiter) {
        iter++;
        if (tau1 != 0.0f) {
            if (Kdense) dense_matvec(Kdense, N, p, kp);
            else orc_crossprod(packed, Mqc, N, freq, invStd, p, kp, nthreads > 0 ? nthreads : 1);
        }
        orc_sigma_mul(p, kp, w, N, tau0, tau1, Ap);
        float a = sdot(r, z, N) / sdot(p, Ap, N);
        for (int64_t i = 0; i < N; i++) x[i] = x[i] + a * p[i];
        for (int64_t i = 0; i < N; i++) r1[i] = r[i] - a * Ap[i];
        for (int64_t i = 0; i < N; i++) z1[i] = minv[i] * r1[i];
        float bet = sdot(z1, r1, N) / sdot(z, r, N);
        for (int64_t i = 0; i < N; i++) p[i] = z1[i] + bet * p[i];
        memcpy(z, z1, sizeof(float) * N);
        memcpy(r, r1, sizeof(float) * N);
        sumr2 = sdot(r, r, N);
    }
    free(r); free(z); free(p); free(Ap); free(kp); free(minv); free(r1); free(z1);
    return iter;
}

/* FP64 dense S = Z Z^T (unscaled) from the packed genotypes, for GRM checks
 * at small N.  S is Nnm x Nnm row-major. */
void orc_grm_f64(const uint8_t* packed, int64_t Mqc, int64_t Nnm,
                 const float* freq, const float* invStd, double* S)
{
    int64_t bpm = (Nnm + 3) / 4;
    float* z = (float*)malloc(sizeof(float) * (size_t)Nnm);
    memset(S, 0, sizeof(double) * (size_t)Nnm * (size_t)Nnm);
    for (int64_t j = 0; j < Mqc; j++) {
        orc_std_geno(packed + j * bpm, Nnm, freq[j], invStd[j], z);
        for (int64_t a = 0; a < Nnm; a++) {
            double za = z[a];
            if (za == 0.0) continue;
            double* row = S + a * Nnm;
            for (int64_t c = 0; c < Nnm; c++) row[c] += za * (double)z[c];
        }
    }
    free(z);
}
