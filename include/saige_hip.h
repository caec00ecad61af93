/*
 * saige_hip.h -- C ABI of libsaigehip, the MI355X (gfx950) implementation of
 * SAIGE step 1's hot path: PLINK decode / QC / standardisation, the resident
 * GRM built by an MFMA SYRK, the PCG solve on Sigma = tau0 W^-1 + tau1 K, and
 * the block PCG that the trace, AI-REML and serving loops run on.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative SAIGE_E* code on
 *     failure; saige_last_error() then holds a message.  No C++ exception
 *     crosses this boundary (the reference turns its C++ exceptions into R
 *     errors with BEGIN_RCPP/END_RCPP, src/SAIGE/src/RcppExports.cpp).
 *   - Pointers named *_host are host memory owned by the caller; the library
 *     copies what it needs.  Device memory is owned by the context.
 *   - One context per device; calls on one context are synchronous and must
 *     not overlap (the reference's state is the process-global genoClass
 *     geno, FG.cpp:1188, driven from R's single thread).
 *   - Vectors are FP32, length Nnomissing (the phenotyped-sample count),
 *     in subSampleInGeno order.  Multi-column arguments are column-major
 *     (column c starts at c * Nnomissing), like arma::fmat.
 *
 * FG.cpp below is src/SAIGE/src/SAIGE_fitGLMM_fast.cpp and GPU.hpp is
 * src/SAIGE/src/gpuSymMatMult.hpp of the reference.
 */
#ifndef SAIGE_HIP_H
#define SAIGE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SAIGE_OK 0
#define SAIGE_E_ARG (-1)      /* invalid argument / size mismatch        */
#define SAIGE_E_STATE (-2)    /* call out of order (e.g. no genotypes)   */
#define SAIGE_E_NOMEM (-3)    /* device or pinned allocation failed      */
#define SAIGE_E_IO (-4)       /* file could not be opened or read        */
#define SAIGE_E_DEVICE (-5)   /* HIP runtime error                       */
#define SAIGE_E_NOTCONV (-6)  /* PCG hit maxiter (result still written)  */
#define SAIGE_PENDING 1       /* saige_poll: the request has not finished */

/* GRM arithmetic for saige_build_grm */
#define SAIGE_GRM_FP32 0        /* FP32 Z panels, FP32 MFMA (v_mfma_f32_32x32x2_f32) */
#define SAIGE_GRM_BF16 1        /* 2-bit genotypes standardised to BF16 inside the SYRK,
                                   BF16 MFMA with FP32 accumulation                    */
#define SAIGE_GRM_BF16_PANEL 2  /* same arithmetic from materialised BF16 Z panels     */
#define SAIGE_GRM_BF16_SPLIT 3  /* 2-bit genotypes standardised to two BF16 pieces each,
                                   z = hi + lo (lo the rounded remainder, z - hi - lo
                                   within 2^-16 |z|); hi.hi + hi.lo + lo.hi on the BF16
                                   MFMA with FP32 accumulation in one SYRK pass: an
                                   FP32-accurate S at about three times the BF16 build,
                                   stored rounded to 16 significant bits (within 2^-16
                                   of each value; LOCO ranges of it are stored the same
                                   way), so the block GEMM splits it into two BF16
                                   pieces instead of three with the same bits */

typedef struct saige_ctx saige_ctx;

/* ------------------------------------------------------------ lifecycle */

/* Create a context on HIP device `device`.  Replaces the process-global
 * genoClass (FG.cpp:1188) and gpuSymMatMult's constructor (GPU.hpp:26). */
int saige_create(int device, saige_ctx** out);
/* Free every device buffer (GPU.hpp:27, ~gpuSymMatMult; closeGenoFile_plink,
 * FG.cpp:1191-1209). */
int saige_destroy(saige_ctx* ctx);
/* Message of the last failed call on this thread. */
const char* saige_last_error(void);

/* ------------------------------------------------------------- setters */

/* setminMAFforGRM (FG.cpp:4923) */
int saige_set_min_maf_for_grm(saige_ctx* ctx, float min_maf);
/* setmaxMissingRateforGRM (FG.cpp:4928) */
int saige_set_max_missing_rate_for_grm(saige_ctx* ctx, float max_missing);
/* setminMAC_VarianceRatio (FG.cpp:4970) */
int saige_set_min_mac_variance_ratio(saige_ctx* ctx, float min_mac, float max_mac,
                                     int is_var_ratio_in_geno);
/* The random common-marker pool for variance ratios, g_randMarkerIndforVR
 * (FG.cpp:866-871, drawn there with arma::randi; the caller draws it here). */
int saige_set_vr_random_markers(saige_ctx* ctx, const int32_t* idx_host, int64_t n);

/* ------------------------------------------------------- genotype ingest */

/* setgeno (FG.cpp:2267-2275 -> genoClass::setGenoObj, FG.cpp:739-1024).
 * Counts samples in the .fam and markers in the .bim, reads the .bed in one
 * pass, decodes, QCs, imputes and repacks on the device.
 *   sub_sample_in_geno : n_sub 1-based .fam row indices of the phenotyped
 *                        samples (subSampleInGeno)
 *   indicator          : N flags (indicatorGenoSamplesWithPheno); NULL derives
 *                        them from sub_sample_in_geno
 *   is_diag_of_kin_set_as_one : isDiagofKinSetAsOne */
int saige_setgeno(saige_ctx* ctx, const char* bed_file, const char* bim_file,
                  const char* fam_file, const int32_t* sub_sample_in_geno, int64_t n_sub,
                  const uint8_t* indicator, int is_diag_of_kin_set_as_one);
/* Same from a .bed body (magic bytes stripped) already in host memory,
 * M * ceil(N/4) bytes. */
int saige_setgeno_host(saige_ctx* ctx, const uint8_t* bed_body_host, int64_t N, int64_t M,
                       const int32_t* sub_sample_in_geno, int64_t n_sub,
                       const uint8_t* indicator, int is_diag_of_kin_set_as_one);
/* Same from a .bed body already in device memory of this context's device
 * (not freed, not modified). */
int saige_setgeno_device(saige_ctx* ctx, const void* bed_body_dev, int64_t N, int64_t M,
                         const int32_t* sub_sample_in_geno, int64_t n_sub,
                         const uint8_t* indicator, int is_diag_of_kin_set_as_one);

/* ------------------------------------------------------------- getters */

int64_t saige_get_nnomissing(const saige_ctx* ctx);   /* getNnomissingOut (FG.cpp:1262) */
int64_t saige_get_total_markers(const saige_ctx* ctx);/* gettotalMarker (FG.cpp:1213)   */
/* getMsub_MAFge_minMAFtoConstructGRM (FG.cpp:1271): QC-passing marker count */
int64_t saige_get_mqc(const saige_ctx* ctx);
int64_t saige_get_n_vr_markers(const saige_ctx* ctx);
/* getAlleleFreqVec (FG.cpp:1219): Mqc floats */
int saige_get_allele_freq(saige_ctx* ctx, float* out_host);
/* getMACVec (FG.cpp:1224): Mqc ints */
int saige_get_mac(saige_ctx* ctx, int32_t* out_host);
/* invstdvVec: Mqc floats */
int saige_get_inv_std(saige_ctx* ctx, float* out_host);
/* getQCdMarkerIndex (FG.cpp:1249): M flags */
int saige_get_qc_marker_index(saige_ctx* ctx, uint8_t* out_host);
/* getMACVec_forVarRatio / getIndexVec_forVarRatio (FG.cpp:1230, 1235) */
int saige_get_mac_vr(saige_ctx* ctx, int32_t* out_host);
int saige_get_index_vr(saige_ctx* ctx, int32_t* out_host);
/* Per-marker decode results for all M markers (the values
 * Get_OneSNP_Geno_atBeginning computes, FG.cpp:326-579): allele count after
 * imputation, missing count, MAC, imputation fill, freq, missing rate,
 * 1/sd and the QC / VR flags (bit0 QC, bit1 VR).  Any pointer may be NULL. */
int saige_get_marker_stats(saige_ctx* ctx, int32_t* allele_count, int32_t* num_missing,
                           int32_t* mac, int32_t* fill, float* freq, float* missing_rate,
                           float* inv_std, uint8_t* flags);
/* Get_OneSNP_Geno (FG.cpp:2281): integer genotypes of QC'd marker idx */
int saige_get_one_snp_geno(saige_ctx* ctx, int64_t idx, int32_t* out_host);
/* Get_OneSNP_Geno_forVarRatio (FG.cpp:2291) */
int saige_get_one_snp_geno_vr(saige_ctx* ctx, int64_t idx, int32_t* out_host);
/* Get_OneSNP_StdGeno (FG.cpp:582-662) of QC'd marker idx */
int saige_get_one_snp_std_geno(saige_ctx* ctx, int64_t idx, float* out_host);
/* Get_Diagof_StdGeno (FG.cpp:665-704): sum_j z_ij^2, FP32, markers in order */
int saige_get_diag_std(saige_ctx* ctx, float* out_host);
/* get_DiagofKin (FG.cpp:4358-4375) / get_GRMdiagVec (FG.cpp:4913-4920) */
int saige_get_diag_of_kin(saige_ctx* ctx, float* out_host);

/* ------------------------------------------------------------------ GRM */

/* Build the resident GRM S = Z Z^T (FP32 storage, full N x N) with an MFMA
 * SYRK over marker panels.  Replaces gpuDistributeSNPs + set_matrix
 * (FG.cpp:1893-1950, GPU.cu:30-138; GPU.hpp:29), which upload the FP32 Z
 * instead.  K = S / Mqc. */
int saige_build_grm(saige_ctx* ctx, int precision);
/* The three calls below inspect the resident GRM; the reference keeps no
 * GRM (it multiplies by Z twice per product), so they have no counterpart.
 * Copy K[r0:r0+nr, c0:c0+nc] (= S / Mqc, FP32) to host, row-major. */
int saige_get_grm_block(saige_ctx* ctx, int64_t r0, int64_t c0, int64_t nr, int64_t nc,
                        float* out_host);
/* Seconds of device time of the last saige_build_grm (HIP events). */
double saige_last_grm_seconds(const saige_ctx* ctx);
/* The SAIGE_GRM_* precision the resident GRM was built with, -1 if none. */
int saige_grm_precision(const saige_ctx* ctx);

/* getCrossprodMatAndKin / parallelCrossProd (FG.cpp:1953-1981, 1669-1708):
 * y = K b.  Replaces gpuSymMatMult::sym_sgemv (GPU.hpp:30, GPU.cu:211-252). */
int saige_crossprod_mat_and_kin(saige_ctx* ctx, const float* b_host, float* out_host);
/* k columns at once: Y = K B (column-major N x k). */
int saige_grm_matvec(saige_ctx* ctx, const float* B_host, int64_t k, float* Y_host);

/* getDiagOfSigma (FG.cpp:2322-2365); tau is {tau0, tau1}. */
int saige_get_diag_of_sigma(saige_ctx* ctx, const float* w_host, const float* tau,
                            float* out_host);
/* getCrossprod (FG.cpp:2397-2425): Sigma b. */
int saige_get_crossprod(saige_ctx* ctx, const float* b_host, const float* w_host,
                        const float* tau, float* out_host);

/* ------------------------------------------------------------------ PCG */

/* getPCG1ofSigmaAndVector (FG.cpp:2593-2809): one right-hand side, Jacobi
 * preconditioner, stop when ||r||^2 <= tol (absolute) or at maxiter.  FP32
 * vectors and GRM product, FP64 dot products and convergence test.  Writes
 * the iteration count to *iters (may be NULL).  Returns SAIGE_E_NOTCONV (x
 * still written) if maxiter was reached, where the reference prints a
 * warning (FG.cpp:2794-2797). */
int saige_pcg1(saige_ctx* ctx, const float* w_host, const float* tau, const float* b_host,
               int maxiter, float tol, float* x_host, int32_t* iters);
/* getSigma_X (FG.cpp:3347-3366) and every other "same Sigma, k right-hand
 * sides" solve: block PCG, one GRM GEMM per iteration for all columns; each
 * column keeps its own alpha/beta and stopping test, so column c equals its
 * own single solve up to the GEMM's rounding. */
int saige_pcg_block(saige_ctx* ctx, const float* w_host, const float* tau, const float* B_host,
                    int64_t k, int maxiter, float tol, float* X_host, int32_t* iters);

/* ------------------------------------------------ leave one chromosome out
 * setStartEndIndex (FG.cpp:3067-3082) together with set_Diagof_StdGeno_LOCO
 * (FG.cpp:4934-4958) for that one range: GRM markers [start, end] (indices
 * into the QC'd marker list, inclusive, as the reference's loops use them)
 * are left out.  Builds, on the device, S_loco = S - Z_r Z_r^T (a copy of the
 * resident GRM minus the range's SYRK, same kernel and precision as
 * saige_build_grm) and diag_loco = diag - sum_{range} z^2 (FP32, marker
 * order), with divisor M_loco = Mqc - (end - start + 1) as in
 * parallelCrossProd_LOCO (FG.cpp:1833-1849) and getDiagOfSigma_LOCO
 * (FG.cpp:2368-2392).  chrom is the reference's chromIndex; it is recorded
 * and returned by saige_loco_range but the range alone defines the GRM.
 * Needs a second Npad^2 FP32 buffer (kept until saige_destroy).
 * SAIGE_E_STATE if the block engine has live LOCO slots. */
int saige_set_start_end_index(saige_ctx* ctx, int64_t start, int64_t end, int chrom);
/* The range set last: start, end, chrom, M_loco (-1s if none). */
int saige_loco_range(const saige_ctx* ctx, int64_t* start, int64_t* end, int* chrom,
                     int64_t* m_loco);
/* getCrossprodMatAndKin_LOCO (FG.cpp:1989-2000): K_loco b. */
int saige_crossprod_mat_and_kin_loco(saige_ctx* ctx, const float* b_host, float* out_host);
/* getDiagOfSigma_LOCO (FG.cpp:2368-2392); ignores isDiagofKinSetAsOne as the
 * reference does. */
int saige_get_diag_of_sigma_loco(saige_ctx* ctx, const float* w_host, const float* tau,
                                 float* out_host);
/* getCrossprod_LOCO (FG.cpp:2430-2460): Sigma_loco b. */
int saige_get_crossprod_loco(saige_ctx* ctx, const float* b_host, const float* w_host,
                             const float* tau, float* out_host);
/* getPCG1ofSigmaAndVector_LOCO (FG.cpp:2943-3030): as saige_pcg1 on Sigma_loco. */
int saige_pcg1_loco(saige_ctx* ctx, const float* w_host, const float* tau, const float* b_host,
                    int maxiter, float tol, float* x_host, int32_t* iters);

/* ------------------------------------------- block-PCG slot engine
 * The engine keeps up to `width` (<= 64) solves in flight over shared K.
 * Each slot belongs to a "system" (w, tau, optional sample mask), so many
 * phenotypes, probes and sample subsets advance together; a slot joins or
 * leaves between iterations (continuous batching).  The reference has no
 * counterpart: it solves one right-hand side at a time (FG.cpp:2593). */

/* Allocate the engine for `n_systems` systems; it always holds 64 slots
 * (width, 1..64, is checked and otherwise unused). */
int saige_engine_reserve(saige_ctx* ctx, int width, int n_systems);
/* Set system s: Sigma_s = tau0 diag(1/w) + tau1 K restricted to mask (N
 * flags; NULL = all samples).  Samples outside the mask are held at zero. */
int saige_engine_set_system(saige_ctx* ctx, int s, const float* w_host, const float* tau,
                            const uint8_t* mask_host);
/* Start solving Sigma_s x = b in slot `slot` (b: N floats on host). */
int saige_engine_add(saige_ctx* ctx, int slot, int s, const float* b_host, int maxiter,
                     float tol);
/* Start k solves in one call: column j of B_host (N x k, column-major)
 * goes to slot slots[j] with system systems[j]. */
int saige_engine_add_batch(saige_ctx* ctx, int k, const int32_t* slots, const int32_t* systems,
                           const float* B_host, int maxiter, float tol);
/* Start k GRM products K b in slots: column j of B_host (N x k) in slot
 * slots[j], restricted to system systems[j]'s samples.  A product slot rides
 * along in the next iteration's GEMM and is then done; its result is read
 * with saige_engine_take_batch like a solution. */
int saige_engine_add_products(saige_ctx* ctx, int k, const int32_t* slots, const int32_t* systems,
                              const float* B_host);
/* Copy the solutions of k slots to X_host (N x k, column-major; NULL to
 * discard) and free the slots. */
int saige_engine_take_batch(saige_ctx* ctx, int k, const int32_t* slots, float* X_host);
/* Run `n_iter` iterations on every live slot.  done[slot] = 1 when the slot
 * has stopped; iters[slot] = its iteration count (either may be NULL; each
 * holds 64 entries). */
int saige_engine_step(saige_ctx* ctx, int n_iter, uint8_t* done_host, int32_t* iters_host);
/* Copy the solution of a slot to host and free the slot. */
int saige_engine_take(saige_ctx* ctx, int slot, float* x_host);
/* ------------------------------------------ device-resident trace terms
 * No reference counterparts: the reference downloads every solve and forms
 * GetTrace's t_i = (K u_i)^T P u_i with Armadillo on the host
 * (FG.cpp:3127-3145).  These keep the probe vectors, K u and Sigma^-1 u on
 * the device and return only FP64 dot products. */
/* engine_add_batch / engine_add_products with B already on this device
 * (column-major, leading dimension Nnomissing). */
int saige_engine_add_batch_device(saige_ctx* ctx, int k, const int32_t* slots, const int32_t* systems,
                                  const float* B_dev, int maxiter, float tol);
int saige_engine_add_products_device(saige_ctx* ctx, int k, const int32_t* slots, const int32_t* systems,
                                     const float* B_dev);
/* Take k finished slots: results to X_host (N x k, may be NULL) and/or to
 * columns xcols[j] (NULL: j) of the device buffer X_dev (leading dimension
 * ldx, may be NULL), and the FP64 dots of each result x_j with
 * P1_dev[:, pcols[j]] and P2_dev[:, pcols[j]] (pairwise, either may be NULL;
 * pcols NULL: j; leading dimension ldp) and with the m shared columns of
 * Q_dev (leading dimension ldq) into dots_host[j * nd + d],
 * nd = [P1] + [P2] + m (at most 16), in that order.  Sums run in a fixed
 * order.  Block engine only. */
int saige_engine_take_batch_ex(saige_ctx* ctx, int k, const int32_t* slots, float* X_host, float* X_dev,
                               int64_t ldx, const int32_t* xcols, const float* P1_dev, const float* P2_dev,
                               int64_t ldp, const int32_t* pcols, const float* Q_dev, int m, int64_t ldq,
                               double* dots_host);
/* FP64 dots of the k columns of A_dev (leading dimension lda >= N) with the
 * m columns of Q_dev: out_host[j * m + r]. */
int saige_dots_device(saige_ctx* ctx, const float* A_dev, int64_t lda, int k, const float* Q_dev, int64_t ldq,
                      int m, double* out_host);
/* saige_dots_device without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Its scratch is a fixed per-device array of 8.4 MB: at most
 * 262,144 samples (SAIGE_E_ARG above), k <= 64 columns, 1 <= m <= 16
 * partners.  Every call on one GPU shares that array, so calls that overlap
 * on one GPU race; the results are bitwise those of saige_dots_device. */
int saige_dots_device_noalloc(saige_ctx* ctx, const float* A_dev, int64_t lda, int k, const float* Q_dev,
                              int64_t ldq, int m, double* out_host);
/* Rademacher probes i0 .. i0+k-1 (entries +-1 from a counter-based hash of
 * (seed, i, row); FG.cpp:3133-3137 draws them with R's rbinom) into dst_dev,
 * column-major with leading dimension ld >= N. */
int saige_probes_device(saige_ctx* ctx, float* dst_dev, int64_t ld, uint64_t seed, int64_t i0, int k);
/* The same probes for a fit on a sample subset: entry i of probe i0 + j
 * (the hash of the position i in the subset, as for a fit of n samples) goes
 * to row rows_dev[i] of column j, every other row of the k columns is zero.
 * rows_dev: n ascending indices into the N samples, on this device. */
int saige_probes_device_rows(saige_ctx* ctx, float* dst_dev, int64_t ld, uint64_t seed, int64_t i0, int k,
                             const int64_t* rows_dev, int64_t n);
/* Host-to-device copy into a buffer from saige_alloc_device. */
int saige_copy_host_to_device(saige_ctx* ctx, void* dst_dev, const void* src_host, int64_t bytes);

/* Point the block engine at the LOCO GRM of saige_set_start_end_index
 * (on = 1) or back at the full GRM (on = 0); the Get_Coef_LOCO solves of many
 * fits (FG.R:1240-1244) then run batched.  Call with no live slots, before
 * saige_engine_set_system (the systems' preconditioner uses diag_loco). */
int saige_engine_use_loco(saige_ctx* ctx, int on);
/* ------------------------------------------------------------ autotuner
 * No reference counterpart (the reference's GRM product is one cuBLAS
 * SGEMV per right-hand side, GPU.cu:211-252).  The block-PCG GEMM
 * Y = S P (64 columns) has three shape parameters: rt, the 32-row tiles per
 * wave (1 or 2); pd, the 32-column chunks of S requested ahead of use (1 or
 * 3); nsplit, the column slices of S summed in a fixed order (1, 2, 4 or 8).
 * saige_autotune times every candidate on the resident GRM (reps launches
 * each after one warm-up, on hashed operands) and keeps the fastest; then,
 * for launches over one 32-column half of the block, it tunes rt and pd
 * again at the chosen nsplit (nsplit fixes the summation order, so a
 * column's bits must not depend on whether the other half is live).
 * ms_table (may be NULL) receives 24 per-launch times in ms: the 16 full
 * shapes in the order of saige_autotune_candidates, then the one-half
 * shapes (rt, pd) = (1,1), (2,1), (1,3), (2,3), then the 16-column shapes of
 * the lower-triangle mode in the same order (-1 in other modes).  Call it
 * once before solving; results are reproducible for a fixed shape.
 * SAIGE_E_STATE if the engine has live slots. */
int saige_autotune(saige_ctx* ctx, int reps, double* ms_table);
/* The candidate shapes (rt[i], pd[i], nsplit[i]); returns their count (16). */
int saige_autotune_candidates(int* rt, int* pd, int* nsplit);
/* Set the block-GEMM shape explicitly, for full and one-half launches
 * (engine idle). */
int saige_set_gemm_config(saige_ctx* ctx, int rt, int pd, int nsplit);
/* Current shape, and its tuned per-launch time in ms (0 if not tuned). */
int saige_get_gemm_config(const saige_ctx* ctx, int* rt, int* pd, int* nsplit, double* tuned_ms);
/* Shape of one-half launches (nsplit as the full shape). */
int saige_get_gemm_config_half(const saige_ctx* ctx, int* rt, int* pd, double* tuned_ms);
/* Shape of 16-column launches (lower-triangle mode; rt as the full shape). */
int saige_get_gemm_config_quarter(const saige_ctx* ctx, int* rt, int* pd, double* tuned_ms);
/* Arithmetic of the block GEMM (no reference counterpart; the reference's
 * product is an FP32 SGEMV).  SAIGE_GEMM_FP32: FP32 MFMA 32x32x2.
 * SAIGE_GEMM_BF16X3: every FP32 operand split exactly into three BF16
 * pieces (8 significand bits each), the six piece products of weight
 * >= 2^-16 summed on the BF16 MFMA with FP32 accumulation; the dropped
 * products are below 2^-23 of each |s p| term, the size of one FP32
 * rounding.  Both give FP32-accurate products; their bits differ.  The
 * shape parameters mean the same in both, except that in SAIGE_GEMM_BF16X3
 * rt = 1 runs workgroups of twelve one-tile waves (384 rows share each
 * staged chunk of P) whose registers hold at most two chunks ahead, so pd = 3
 * means two there.  SAIGE_GEMM_BF16X3_TRI: the same six piece products (five
 * on the split GRM, whose values need two pieces), on the 16x16x32 MFMA,
 * read from the lower triangle of the GRM only (half the bytes): each 32x32
 * block left of the diagonal gives both S_IJ P_J and S_IJ^T P_I.  A launch
 * covers 64, 32 or 16 columns, as few as hold live slots (at most 16 or 32
 * occupied slots run as one 16- or 32-column launch wherever they sit), with
 * the same per-column arithmetic (SAIGE_LAUNCH_GEMM_QUARTER counts the
 * 16-column ones, SAIGE_LAUNCH_GEMM_HALF the 32-column ones).  rt selects
 * the workgroup (1: twelve waves, 2: eight), pd 1 or 3 the prefetch depth
 * (one or two chunks ahead); nsplit is unused.  It needs device scratch for
 * its partial products, about 128 N^2 / (32 * waves) bytes (3.3 GB at
 * N = 100k with twelve waves), allocated on first use.  Changing the mode
 * clears the tuned times (re-run saige_autotune).  Engine idle. */
#define SAIGE_GEMM_FP32 0
#define SAIGE_GEMM_BF16X3 1
#define SAIGE_GEMM_BF16X3_TRI 2
int saige_set_gemm_mode(saige_ctx* ctx, int mode);
int saige_get_gemm_mode(const saige_ctx* ctx, int* mode);
/* Single-column GRM products (the PCG of saige_pcg1, saige_get_crossprod,
 * saige_crossprod_mat_and_kin, one-column saige_grm_matvec; no reference
 * counterpart, the reference's is an FP32 SGEMV).  SAIGE_GEMV_TRI (the
 * default) reads only the GRM's lower triangle: each element serves
 * y_i += s_ij x_j and y_j += s_ij x_i, partial sums are added in a fixed
 * order, and it allocates about 2 N^2 / 256 bytes of scratch on first use
 * (80 MB at N = 100k).  SAIGE_GEMV_FULL reads all of S (k_gemv).  Both are
 * FP32 with different summation orders. */
#define SAIGE_GEMV_FULL 0
#define SAIGE_GEMV_TRI 1
int saige_set_gemv_mode(saige_ctx* ctx, int mode);
int saige_get_gemv_mode(const saige_ctx* ctx, int* mode);

/* Free and total device memory (hipMemGetInfo), for admission control. */
int saige_device_memory(saige_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);
/* Device bytes one engine system (w, tau, sample mask) occupies. */
int64_t saige_engine_system_bytes(const saige_ctx* ctx);
/* Seconds of device time spent in the engines' GRM product launches of one
 * kind, and their count (HIP events around each launch). */
#define SAIGE_LAUNCH_GEMV 0       /* width-1 engine (saige_pcg1): k_gemv          */
#define SAIGE_LAUNCH_GEMM_HALF 1  /* block GEMM over 32 columns: one half, or (lower
                                     triangle) at most 32 occupied slots anywhere */
#define SAIGE_LAUNCH_GEMM_FULL 2  /* block GEMM over all 64 columns                */
#define SAIGE_LAUNCH_GEMM_QUARTER 3 /* lower-triangle block GEMM over 16 columns
                                     (at most 16 occupied slots, anywhere)       */
int saige_engine_stats(const saige_ctx* ctx, int kind, double* gemm_seconds, int64_t* gemm_launches);

/* ------------------------------------------------------- synthetic data */

/* Fill a device .bed body (M x ceil(N/4) bytes, PLINK-1 codes) with
 * genotypes g ~ Binomial(2, p_j), p_j ~ U(maf_lo, maf_hi), each call missing
 * with probability miss_rate, from a counter-based generator keyed by seed.
 * If pair_frac > 0, samples [0, 2*npairs) form sibling pairs (2i, 2i+1) with
 * npairs = floor(pair_frac * N / 2): the second copies the first's
 * genotype at each marker with probability 1/2.  Test/benchmark input only. */
int saige_synth_bed_device(saige_ctx* ctx, void* bed_body_dev, int64_t N, int64_t M,
                           float maf_lo, float maf_hi, float miss_rate, float pair_frac,
                           uint64_t seed);
/* Device allocation helper for the above (freed by saige_free_device). */
int saige_alloc_device(saige_ctx* ctx, int64_t bytes, void** out);
int saige_free_device(saige_ctx* ctx, void* p);
/* Page-locked host memory (freed by saige_free_host): host columns staged in
 * it move to and from the device at the link's rate.  No counterpart in the
 * reference, whose vectors stay in host memory; the library's own
 * scheduler stages its column transfers here. */
int saige_alloc_host(saige_ctx* ctx, int64_t bytes, void** out);
int saige_free_host(saige_ctx* ctx, void* p);
/* Device-to-device copy on the context's stream (the known-size copy that
 * calibrates HBM traffic counters in profiles). */
int saige_copy_device(saige_ctx* ctx, void* dst_dev, const void* src_dev, int64_t bytes);
/* Copy rows of a device .bed body to host (for building phenotypes). */
int saige_copy_device_to_host(saige_ctx* ctx, const void* src_dev, int64_t bytes, void* dst_host);

/* ---------------------------------------------- null fits in C++ (host)
 * The R drivers of step 1 in C++, so that an R or C++ caller gets whole
 * null fits -- many of them batched on the block-PCG engine -- without
 * another runtime: glm() start values (FG.R:1119, 1370), Get_Coef
 * (FG.R:2-35), glmmkin.ai_PCG_Rcpp_Binary / _Quantitative (FG.R:79-304,
 * 309-549) with getCoefficients, getAIScore(_q), GetTrace(_q) and
 * fitglmmaiRPCG(_q) (FG.cpp:3113-3160, 3167-3199, 3279-3341, 3409-3662),
 * Covariate_Transform (FG.R:1612-1647) and ScoreTest_NULL_Model
 * (FG.R:579-591).  Host arithmetic is FP64; trace probes are the
 * deterministic Rademacher vectors of saige_probes_device (keyed by
 * probe_seed and the probe index) in place of R's rbinom draws.  Matrices
 * are column-major. */
#define SAIGE_TRAIT_BINARY 0
#define SAIGE_TRAIT_QUANTITATIVE 1
typedef struct saige_fit_options {
    int maxiter;            /* AI-REML and Get_Coef iterations (20)          */
    double tol;             /* tau tolerance (0.02)                           */
    int nrun;               /* probes in the first trace round (30)          */
    double tol_pcg;         /* PCG tolerance (1e-5)                           */
    int maxiter_pcg;        /* PCG iterations (500)                           */
    double trace_cv_cutoff; /* trace CV cutoff (0.0025)                       */
    uint64_t probe_seed;    /* probe hash seed (200)                          */
} saige_fit_options;
typedef struct saige_fit_spec {
    int trait;              /* SAIGE_TRAIT_BINARY or SAIGE_TRAIT_QUANTITATIVE */
    int64_t n;              /* samples in this fit                            */
    const int64_t* subset;  /* n ascending indices into the N phenotyped
                               samples, or NULL for all of them (n == N)      */
    const double* y;        /* n                                              */
    const double* X;        /* n x p, covariate-transformed                   */
    int p;
    const double* offset;   /* n, or NULL                                     */
} saige_fit_spec;
typedef struct saige_fit_result {
    double theta[2];
    double* coefficients;       /* p (caller-allocated; required)     */
    double* linear_predictors;  /* n, or NULL                         */
    double* fitted_values;      /* n, or NULL                         */
    double* working_y;          /* n: Y of the last Get_Coef, or NULL */
    double* cov;                /* p x p, or NULL                     */
    int converged;
    int iterations;
    int status;                 /* SAIGE_OK, or why this fit failed   */
} saige_fit_result;
/* The defaults of fitNULLGLMM (FG.R:1036-1063). */
void saige_fit_options_default(saige_fit_options* opt);
/* Fit n_fits null GLMMs on the resident GRM (engine idle).  Up to
 * max_concurrent fits (<= 0: 16) are admitted at a time, each with its own
 * engine system (weights, tau, sample mask); their solves and GRM products
 * share the engine's 64 slots (continuous batching), and each fit's host
 * work runs in its own thread while the engine steps.  Returns SAIGE_OK when
 * the engine ran; results[i].status says whether fit i succeeded (e.g.
 * SAIGE_E_STATE when its first variance component went to 0, FG.R:449-451). */
int saige_fit_null_batch(saige_ctx* ctx, int n_fits, const saige_fit_spec* specs, const saige_fit_options* opt,
                         int max_concurrent, saige_fit_result* results);
/* ------------------------------------------------- serving null fits
 * Replaces the reference's per-phenotype task farming, one step-1 R
 * process per phenotype (mvp_gwphewas_study/data_analysis/01_gwas_gpu/
 * scripts_for_run/submit_step1_with_nvme_gpu_v2.r:118-120), with one
 * process that keeps the resident GRM busy with a stream of requests
 * (continuous batching).  saige_submit copies a request and returns at once;
 * the server's scheduler thread admits requests while an engine system
 * (max_concurrent of them) and HBM (hbm_reserve_bytes free; 0 = no check)
 * are free; an admitted request's solves and GRM products take engine
 * slots as they free up and join the iterations already under way, so a
 * request joins a mix that is mid-solve; saige_poll hands a finished
 * request's result over.  A fit's arithmetic does not depend on the other
 * columns, so its result is bit-identical alone and in any mix.  While a
 * server runs, its context does no other engine work. */
typedef struct saige_server saige_server;
/* Start a server on the resident GRM with the fit options (NULL: the
 * defaults). */
int saige_server_start(saige_ctx* ctx, const saige_fit_options* opt, int max_concurrent, int64_t hbm_reserve_bytes,
                       saige_server** out);
/* Queue one null fit (the spec's arrays are copied); *ticket names it. */
int saige_submit(saige_server* srv, const saige_fit_spec* spec, int64_t* ticket);
/* The result of request `ticket` once it has finished: SAIGE_OK with
 * *result filled as saige_fit_null_batch fills it (result->status says
 * whether the fit succeeded; the ticket is then released), or SAIGE_PENDING
 * when it has not finished within wait_ms milliseconds (0: do not wait;
 * negative: wait until it has).  times (2, or NULL): seconds from
 * submission to admission and to completion.  Polls of one ticket from
 * several threads are safe: one collects the result, the others get
 * SAIGE_E_ARG. */
int saige_poll(saige_server* srv, int64_t ticket, int wait_ms, saige_fit_result* result, double* times);
/* Finish every queued request, stop the scheduler thread and free the
 * server (results not yet polled are dropped). */
int saige_server_stop(saige_server* srv);
/* Covariate_Transform (FG.R:1612-1647): X (n x p) -> sqrt(n) Q of its QR
 * decomposition (X_out n x p_kept, R_out p_kept x p_kept or NULL), after
 * dropping linearly dependent columns (kept: their indices, or NULL). */
int saige_covariate_transform(const double* X, int64_t n, int p, double* X_out, double* R_out, int32_t* kept,
                              int* p_kept);
/* ScoreTest_NULL_Model (FG.R:579-591) for fitted mu and variance V; any
 * output may be NULL: XV (p x n), XVX, XVX_inv (p x p), XXVX_inv and
 * XVX_inv_XV (n x p), S_a (p). */
int saige_score_test_null(const double* mu, const double* V, const double* y, const double* X, int64_t n, int p,
                          double* XV, double* XVX, double* XVX_inv, double* XXVX_inv, double* XVX_inv_XV,
                          double* S_a);
/* The start values fitNULLGLMM takes from glm() (FG.R:1119, 1370): R's
 * glm.fit IRLS with glm.control's defaults (25 iterations, epsilon 1e-8,
 * relative deviance change) for trait SAIGE_TRAIT_BINARY (logit) or
 * SAIGE_TRAIT_QUANTITATIVE (identity) on X (n x p column-major) with an
 * optional offset (n, or NULL for 0).  Outputs coef (p), eta and mu (n).
 * Host only, no context: callers may run several at once on their threads. */
int saige_glm_start(int trait, int64_t n, const double* y, const double* X, int p, const double* offset,
                    double* coef, double* eta, double* mu);
/* Get_Coef's working response and weights at eta (FG.R:9-12, 24-27), the
 * host step between its getCoefficients calls: mu = linkinv(eta),
 * Y = eta - offset + (y - mu) / mu_eta(eta), W = (mu_eta(eta) /
 * sqrt(variance(mu)))^2, all n long; offset may be NULL (0).  Host only. */
int saige_glm_working(int trait, int64_t n, const double* y, const double* eta, const double* offset, double* mu,
                      double* Y, double* W);
/* getCoefficients (FG.cpp:3167-3199) after its solves, for hosts that batch
 * the solves themselves: sol = Sigma^-1 [Y, X] (n x (1+p), column-major),
 * X (n x p, column-major), w and Y (n).  Outputs cov = inv_sympd(X^T Sigma^-1
 * X) (p x p), alpha = cov (Sigma^-1 X)^T Y (p) and eta = Y - tau0
 * (Sigma^-1 Y - Sigma^-1 X alpha) / w (n).  Host only. */
int saige_coefficients(int64_t n, int p, const double* Y, const double* X, const double* w, double tau0,
                       const double* sol, double* cov, double* alpha, double* eta);

#ifdef __cplusplus
}
#endif
#endif /* SAIGE_HIP_H */
