// Internal declarations shared by the libsaigehip translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/saige_hip.h"

namespace saige {

// Per-marker decode record written by the decode kernel; field meaning as in
// genoClass::Get_OneSNP_Geno_atBeginning (FG.cpp:326-579).
struct MarkerStat {
    int32_t allele_count;
    int32_t num_missing;
    int32_t mac;
    int32_t fill;
    float freq;
    float missing_rate;
    float inv_std;
    uint8_t pass_qc;
    uint8_t pass_vr;
    uint8_t pad[2];
};
static_assert(sizeof(MarkerStat) == 32, "MarkerStat layout");

struct DecodeParams {
    int64_t N;            // samples in the .fam
    int64_t M;            // markers in the .bim
    int64_t Nnm;          // phenotyped samples (Nnomissing)
    int64_t bpm;          // bytes per marker in the .bed, ceil(N/4)
    float min_maf;
    float max_missing;
    int vr_mode;          // 0 none, 1 categorical, 2 single category
    float min_mac_vr;
    float max_mac_vr;
};

// Engine tile width: every block-PCG launch works on EW columns.
constexpr int EW = 64;
// Row padding of all N-length device arrays and of the GRM.
constexpr int64_t NPAD_ALIGN = 256;

// Block-GEMM shape: rt = 32-row tiles per wave (1 or 2), pd = 32-column
// chunks of S requested ahead (1 or 3), nsplit = column slices of S (1, 2, 4
// or 8).  The slice count fixes the summation order, so results are
// reproducible for a fixed configuration.
// Block-GEMM arithmetic: FP32 MFMA (32x32x2), or FP32 operands split
// exactly into three BF16 pieces on the BF16 MFMA (32x32x16), over all of S
// or over its lower triangle (each block used for S_IJ P_J and S_IJ^T P_I)
enum GemmMode { GEMM_FP32 = 0, GEMM_BF16X3 = 1, GEMM_BF16X3_TRI = 2 };

// Work list and partial-product scratch of the lower-triangle product for
// one Npad.  Row set R = rows 32*nwv*R + [0, 32*nwv); item = (R, chunk range
// [ca, cb) of 32 columns, use-1 slot); row set R's use-1 partials are the
// slots rtab[R].x + [0, rtab[R].y).
// the block columns of a 16- or 32-column launch of the lower-triangle
// GEMM: launch column k < width computes block column col[k]; distinct
// columns below EW
struct NarrowCols {
    int width = 16;
    uint8_t col[32] = {};
    static NarrowCols tile(int width, int first) {   // columns first + [0, width)
        NarrowCols c;
        c.width = width;
        for (int k = 0; k < width; k++) c.col[k] = (uint8_t)(first + k);
        return c;
    }
};
// the same map as a kernel argument: byte k of w[k / 4] is col[k]
struct ColMap32 {
    uint32_t w[8];
};

struct TriPlan {
    int64_t Npad = 0;
    int nwv = 0;                 // waves per workgroup (8 or 12)
    int ch = 0;                  // most chunks per item left of the diagonal
    int64_t nitems = 0, nr = 0, y1_slots = 0, y2_slots = 0;
    int4* items = nullptr;       // device [nitems], longest first
    int2* rtab = nullptr;        // device [nr]
    float* y1p = nullptr;        // [y1_slots][32*nwv][W], W = the launch's width (<= EW)
    float* y2p = nullptr;        // [y2_slots][32][W]
};
// builds (or keeps, when Npad/nwv/ch match) the plan; frees on failure
hipError_t tri_plan_build(TriPlan& t, int64_t Npad, int nwv, int ch);
void tri_plan_free(TriPlan& t);

// The lower-triangle GEMV's work list and partials: row sets of 256 rows,
// chunks of 256 columns; items as in TriPlan (at most ch chunks left of the
// diagonal block, or the diagonal block); use-1 partials [slot][256], use-2
// partials [chunk c][row set R > c][256].
struct GemvPlan {
    int64_t Npad = 0;
    int ch = 0;
    int64_t nitems = 0, nr = 0, y1_slots = 0, y2_slots = 0;
    int4* items = nullptr;
    int2* rtab = nullptr;
    float* y1p = nullptr;
    float* y2p = nullptr;
};
hipError_t gemv_plan_build(GemvPlan& t, int64_t Npad, int ch);
void gemv_plan_free(GemvPlan& t);

struct GemmConfig {
    int rt = 1;
    int nsplit = 4;
    int pd = 1;
    int mode = GEMM_FP32;
};

struct Engine {
    int width = 0;        // live slot capacity (<= EW)
    int nsys = 0;
    int loco = 0;         // 1: products with the LOCO GRM S_loco, divisor M_loco
    int kpad = 0;         // 1 (GEMV) or EW (MFMA GEMM)
    // [Npad][kpad] row-major column blocks
    float* X = nullptr; float* R = nullptr; float* Z = nullptr; float* P = nullptr; float* AP = nullptr;
    // per system [nsys][Npad]
    float* winv = nullptr; float* minv = nullptr; uint8_t* mask = nullptr;
    float* tau = nullptr;         // [nsys][2]
    // per slot state.  What the host sets per slot lives in one device block
    // (slot_state, slot_state_bytes(kpad)) mirrored in page-locked host memory
    // (h_state), so a call that changes it uploads it with one copy:
    // d_list, slot_sys, maxiter, tol, live, kind and d_take point into the
    // device block, h_list .. h_take at the same offsets into the mirror
    uint8_t* slot_state = nullptr;
    uint8_t* h_state = nullptr;
    int32_t* slot_sys = nullptr;  // device [kpad]
    double* rz = nullptr;         // z^T r
    double* sumr2 = nullptr;      // r^T r
    double* pap = nullptr;        // p^T Sigma p
    float* alpha = nullptr;
    float* beta = nullptr;
    int32_t* iter = nullptr;
    int32_t* maxiter = nullptr;
    double* tol = nullptr;
    uint8_t* live = nullptr;      // slot is solving
    uint8_t* done = nullptr;      // slot stopped (device)
    uint8_t* kind = nullptr;      // 0 = PCG solve, 1 = GRM product (one GEMM pass)
    uint8_t* h_kind = nullptr;
    double* partial = nullptr;    // [nparts][kpad][2] partial dot products
    int64_t nparts = 0;
    float* stage = nullptr;       // [Npad] staging for one column
    float* stageB = nullptr;      // [kpad][Npad] staging for a block of columns
    float* ypart = nullptr;       // split-K partial products of the block GEMM
    int32_t* d_list = nullptr;    // [2][kpad] slot / system lists of a batch call
    int32_t* d_take = nullptr;    // [3][kpad] slots, destination and partner columns of a take
    int32_t* h_list = nullptr;
    int32_t* h_take = nullptr;
    int32_t* h_maxiter = nullptr;
    double* h_tol = nullptr;
    // pinned host mirrors
    uint8_t* h_done = nullptr;
    int32_t* h_iter = nullptr;
    int32_t* h_slot_sys = nullptr;
    uint8_t* h_live = nullptr;
    // device time of the GRM product launches by kind: [0] GEMV (width 1),
    // [1] GEMM over one 32-column half, [2] GEMM over both halves
    double gemm_seconds[4] = {0.0, 0.0, 0.0, 0.0};   // by SAIGE_LAUNCH_* kind
    int64_t gemm_launches[4] = {0, 0, 0, 0};
};

}  // namespace saige

struct saige_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // settings
    float min_maf = 0.0f;
    float max_missing = 1.0f;
    float min_mac_vr = 20.0f, max_mac_vr = -1.0f;
    int is_var_ratio = 0;
    std::vector<int32_t> vr_rand;
    int diag_one = 0;
    // sizes
    int64_t N = 0, M = 0, Nnm = 0, Npad = 0, Mqc = 0, Mvr = 0, bpm_nm = 0;
    bool loaded = false;
    // decode outputs
    saige::MarkerStat* d_stats = nullptr;        // [M]
    std::vector<saige::MarkerStat> h_stats;      // [M]
    std::vector<int32_t> qc_index, vr_index;     // original marker index of QC'd / VR markers
    uint8_t* d_packed = nullptr;                 // [Mqc][bpm_nm] internal code
    uint8_t* d_packed_vr = nullptr;              // [Mvr][bpm_nm]
    float4* d_lut = nullptr;                     // [Mqc] (L0, L1, L2, 0)
    float4* d_lut_vr = nullptr;
    float* d_diag = nullptr;                     // [Npad] sum z^2
    // GRM S = Z Z^T, [Npad][Npad] FP32
    float* d_grm = nullptr;
    // LOCO (setStartEndIndex, FG.cpp:3067): S minus one range of GRM markers
    float* d_grm_loco = nullptr;                 // [Npad][Npad]
    float* d_diag_loco = nullptr;                // [Npad]
    bool loco_ready = false;
    int64_t loco_start = -1, loco_end = -1;      // inclusive range in the QC'd marker list
    int64_t loco_m = 0;                          // Mqc - (end - start + 1)
    int loco_chrom = -1;
    bool loco_inplace = false;                   // the range is subtracted from d_grm itself
    bool grm_ready = false;
    saige::GemmConfig gemm_cfg;                  // block-GEMM shape (saige_autotune)
    double gemm_tuned_ms = 0.0;                  // its measured time per launch, 0 if untuned
    // shape for launches over one 32-column half: same nsplit (the summation
    // order, so a column's bits do not depend on the other half), own rt/pd
    saige::GemmConfig gemm_cfg_half;
    saige::GemmConfig gemm_cfg_quarter;     // 16-column launches (lower-triangle mode)
    saige::TriPlan tri;                          // lower-triangle product (mode 2), built on first use
    int gemv_mode = SAIGE_GEMV_TRI;              // single-column products: lower triangle or all of S
    saige::GemvPlan triv;                        // the lower-triangle GEMV's plan, built on first use
    double gemm_tuned_ms_half = 0.0;
    double gemm_tuned_ms_quarter = 0.0;
    // FP64 column-dot scratch (saige_engine_take_batch_ex, saige_dots_device)
    double* d_dots_partial = nullptr;
    double* d_noalloc = nullptr;            // this device's address of the allocation-free scratch
    // device blocks parked by the scheduler between runs (engine_park.hpp)
    std::mutex park_mu;
    std::vector<std::pair<int64_t, void*>> parked;
    double* d_dots_out = nullptr;
    int grm_precision = -1;
    double grm_seconds = 0.0;
    // scratch vectors
    float* d_vec = nullptr;                      // [4][Npad]
    // GRM product scratch (saige_grm_matvec), allocated on first use
    float* d_mv_in = nullptr;                    // [EW][Nnm]
    float* d_mv_blk = nullptr;                   // [Npad][EW]
    float* d_mv_out = nullptr;                   // [Npad][EW]
    float* d_mv_ypart = nullptr;                 // split-K scratch
    saige::Engine eng;
    saige::Engine eng1;                          // kpad = 1 engine for saige_pcg1
};

namespace saige {

void set_error(const char* fmt, ...);

#define SAIGE_HIP_TRY(expr)                                                            \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) {                                                        \
            saige::set_error("%s failed at %s:%d: %s", #expr, __FILE__, __LINE__,      \
                             hipGetErrorString(_e));                                   \
            return (_e == hipErrorOutOfMemory) ? SAIGE_E_NOMEM : SAIGE_E_DEVICE;       \
        }                                                                              \
    } while (0)

inline int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// Correctly rounded FP32 division and square root.  The FP32 hardware
// reciprocal/sqrt paths are not exact on gfx950; computing in FP64 and
// rounding once is, because FP64 carries more than 2*24+2 significand bits
// (double rounding is then innocuous for / and sqrt).
__device__ __forceinline__ float div_rn(float a, float b) { return (float)((double)a / (double)b); }
__device__ __forceinline__ float sqrt_rn(float a) { return (float)__builtin_sqrt((double)a); }

// ---- genotype kernels (geno_kernels.hip)
// imask: NULL when every .fam sample is phenotyped, else 16 copies (stride
// ld, a multiple of 16) of the rows' 2-bit code mask, copy r shifted by r
// bytes (decode_mask_copies)
hipError_t launch_decode_stats(hipStream_t s, const uint8_t* bed, const uint8_t* imask, int64_t ld,
                               const uint8_t* randset, DecodeParams p, MarkerStat* stats);
hipError_t launch_repack(hipStream_t s, const uint8_t* bed, int64_t bpm, const int32_t* markers,
                         int64_t nmark, const MarkerStat* stats, const int32_t* subidx,
                         int64_t Nnm, int64_t bpm_nm, uint8_t* out);
hipError_t launch_lut(hipStream_t s, const int32_t* markers, int64_t nmark,
                      const MarkerStat* stats, float4* lut);
hipError_t launch_diag_std(hipStream_t s, const uint8_t* packed, const float4* lut, int64_t Mqc,
                           int64_t Nnm, int64_t bpm_nm, float* diag);
hipError_t launch_get_geno(hipStream_t s, const uint8_t* packed_row, int64_t Nnm, int32_t* out);
hipError_t launch_std_geno(hipStream_t s, const uint8_t* packed_row, const float4* lut_row,
                           int64_t Nnm, float* out);
hipError_t launch_synth_bed(hipStream_t s, uint8_t* bed, int64_t N, int64_t M, float maf_lo,
                            float maf_hi, float miss, int64_t npairs, uint64_t seed);
// Z panel [Npad][ldz] for markers [j0, j0+np) (np <= ldz; columns beyond np
// and rows beyond Nnm are zero).  Element type float or bf16.
hipError_t launch_std_panel_f32(hipStream_t s, const uint8_t* packed, const float4* lut, int64_t j0,
                                int64_t np, int64_t Nnm, int64_t Npad, int64_t bpm_nm, int64_t ldz,
                                float* Zp);
hipError_t launch_std_panel_bf16(hipStream_t s, const uint8_t* packed, const float4* lut, int64_t j0,
                                 int64_t np, int64_t Nnm, int64_t Npad, int64_t bpm_nm, int64_t ldz,
                                 __hip_bfloat16* Zp);

// ---- GRM kernels (grm_kernels.hip)
// S (Npad x Npad, ld = Npad) (+)= Zp Zp^T over the lower-triangle tiles,
// mirrored into the upper triangle.  kdim is a multiple of the kernel's k-chunk.
// tiles: the lower-triangle tile list from syrk_tile_order(Npad / syrk_tile())
hipError_t launch_syrk_f32(hipStream_t s, const float* Zp, int64_t Npad, int64_t ldz, int64_t kdim,
                           const int2* tiles, float* S, int accumulate);
hipError_t launch_syrk_bf16(hipStream_t s, const __hip_bfloat16* Zp, int64_t Npad, int64_t ldz,
                            int64_t kdim, const int2* tiles, float* S, int accumulate);
std::vector<int2> syrk_tile_order(int64_t nt, int64_t g);
// 2-bit-input SYRK: tp = tiled packed genotypes ((Npad/128) x nchunk x 2 KB,
// nchunk = ceil(Mqc/64)), lut64 = nchunk*64 entries; both from launch_pack_tiles
hipError_t launch_pack_tiles(hipStream_t s, const uint8_t* packed, const float4* lut, int64_t Mqc, int64_t Nnm,
                             int64_t Npad, int64_t bpm_nm, uint8_t* tp, uint64_t* lut64);
// markers [j0, j1) only (j1 < 0: all); lut64 must be zero outside the range.
// accumulate: 0 overwrite, 1 add, -1 subtract.
// tile: 128 or 256 (tiles from syrk_tile_order(Npad / tile, ...)).
hipError_t launch_syrk_pk(hipStream_t s, const uint8_t* tp, const uint64_t* lut64, int64_t Mqc, int64_t Npad,
                          const int2* tiles, float* S, int accumulate, int64_t j0, int64_t j1, int tile);
int syrk_pk_tile();
// copy each diagonal tile's lower half onto its upper half (tile edge `tile`)
hipError_t launch_mirror_diag(hipStream_t s, float* S, int64_t Npad, int tile);
// the split GRM in one pass over markers [j0, j1): hi.hi + hi.lo + lo.hi per
// 32-marker chunk from the two tables (256 tiles; accumulate as launch_syrk_pk)
hipError_t launch_syrk_pk_split(hipStream_t s, const uint8_t* tp, const uint64_t* lhi, const uint64_t* llo,
                                int64_t Mqc, int64_t Npad, const int2* tiles, float* S, int accumulate, int64_t j0,
                                int64_t j1);
// piece 0: the values rounded to BF16 (hi); 1: the remainders rounded to BF16 (lo)
hipError_t launch_lut64_range(hipStream_t s, const float4* lut, int64_t Mqc, int64_t j0, int64_t j1, int piece,
                              uint64_t* lut64);
hipError_t launch_sub(hipStream_t s, const float* a, const float* b, int64_t n, float* out);
int syrk_tile();
int syrk_kchunk_f32();
int syrk_kchunk_bf16();

// ---- PCG kernels (pcg_kernels.hip)
struct SigmaArgs {
    const float* winv;     // [nsys][Npad]
    const uint8_t* mask;   // [nsys][Npad]
    const float* tau;      // [nsys][2]
    const int32_t* sys;    // [kpad] system of each column
    const uint8_t* live;   // [kpad] column is being solved (else output zero)
    float inv_m;           // unused; the product is divided by Mqc as float
    float mqc;             // Mqc as float (divisor, FG.cpp:1706)
    int mode;              // 0: out = S p / Mqc ; 1: out = Sigma p
    const uint8_t* kind;   // [kpad] 1 = product slot (out = K p, also written to xout); may be null
    float* xout;           // result block of product slots
};
// out[N][kpad] = epilogue(S * in), plus per-column partial p^T out dots when
// mode == 1 (partial: [nparts][kpad][2], slot 0).  Returns nparts used.
// tri: read only the lower triangle (k_gemv_tri + k_gemv_tri_epi) with this
// plan for Npad; nullptr: k_gemv over all of S.
hipError_t launch_grm_gemv(hipStream_t s, const float* S, int64_t Npad, int64_t Nnm, const float* in,
                           float* out, SigmaArgs a, double* partial, int64_t* nparts,
                           const GemvPlan* tri = nullptr);
bool gemm_config_valid(GemmConfig g);
// ypart: gemm_ypart_floats(Npad) floats of scratch for the split-K slices.
// halves: bit h set = columns [32h, 32h+32) are computed; the epilogue's
// values for a half that is not computed are meaningless (the caller only
// passes halves without live columns).
// tri: the plan for Npad, required in mode GEMM_BF16X3_TRI (else unused).
// s16: every value of S has at most 16 significant bits (the split GRM's
// storage); the lower-triangle mode then splits S into two BF16 pieces
// instead of three, with the same bits.
// narrow (lower-triangle mode): only the 16 or 32 block columns it lists
// are computed (halves is then unused); nullptr: the halves.
hipError_t launch_grm_gemm(hipStream_t s, const float* S, int64_t Npad, int64_t Nnm, const float* in,
                           float* out, SigmaArgs a, double* partial, int64_t* nparts, float* ypart,
                           GemmConfig g, int halves = 3, const TriPlan* tri = nullptr, bool s16 = false,
                           const NarrowCols* narrow = nullptr);
int64_t gemm_ypart_floats(int64_t Npad);
int64_t gemv_max_parts(int64_t Npad);
int64_t gemm_max_parts(int64_t Npad);
int64_t vec_max_parts(int64_t Npad, int kpad);

hipError_t launch_sigma_setup(hipStream_t s, const float* w, const uint8_t* mask_in, const float* diag,
                              int64_t Nnm, int64_t Npad, float tau0, float tau1, float mqc,
                              int diag_one, float* winv, float* minv, uint8_t* mask_out, float* dout);
// column init: X=0, R=b (staged), Z=minv*R, P=Z; partial rz/rr dots
hipError_t launch_col_init(hipStream_t s, const float* b, int slot, int kpad, int64_t Nnm,
                           int64_t Npad, const float* minv_sys, const uint8_t* mask_sys,
                           float* X, float* R, float* Z, float* P, double* partial,
                           int64_t* nparts);
hipError_t launch_col_init_finish(hipStream_t s, int slot, int kpad, const double* partial,
                                  int64_t nparts, double* rz, double* sumr2, const int32_t* maxiter,
                                  const double* tol, int32_t* iter, uint8_t* done, uint8_t* live);
hipError_t launch_alpha(hipStream_t s, int kpad, const double* partial, int64_t nparts,
                        const double* rz, const uint8_t* done, const uint8_t* live, const uint8_t* kind,
                        float* alpha);
hipError_t launch_update(hipStream_t s, int kpad, int64_t Nnm, int64_t Npad, const float* alpha,
                         const uint8_t* done, const uint8_t* live, const uint8_t* kind, const float* minv,
                         const int32_t* sys, float* X, float* R, float* Z, const float* P,
                         const float* AP, double* partial, int64_t* nparts);
hipError_t launch_beta(hipStream_t s, int kpad, const double* partial, int64_t nparts, double* rz,
                       double* sumr2, const int32_t* maxiter, const double* tol, int32_t* iter,
                       uint8_t* done, const uint8_t* live, const uint8_t* kind, float* beta);
hipError_t launch_pupdate(hipStream_t s, int kpad, int64_t Nnm, int64_t Npad, const float* beta,
                          const uint8_t* done, const uint8_t* live, const uint8_t* kind, const float* Z,
                          float* P);
hipError_t launch_cols_init(hipStream_t s, const float* B, int k, const int32_t* slots, const int32_t* sys,
                            int kpad, int64_t Nnm, int64_t Npad, const float* minv, const uint8_t* mask, float* X,
                            float* R, float* Z, float* P, double* partial, int64_t* nparts, int product);
hipError_t launch_cols_init_finish(hipStream_t s, int k, const int32_t* slots, int kpad, const double* partial,
                                   int64_t nparts, double* rz, double* sumr2, const int32_t* maxiter,
                                   const double* tol, int32_t* iter, uint8_t* done, int product);
hipError_t launch_cols_gather(hipStream_t s, const float* src, int k, const int32_t* slots, int kpad, int64_t Nnm,
                              float* dst);
hipError_t launch_gather_col(hipStream_t s, const float* src, int slot, int kpad, int64_t Nnm,
                             float* dst);
hipError_t launch_scatter_cols(hipStream_t s, const float* src, int64_t k, int64_t Nnm,
                               int64_t Npad, int kpad, float* dst);
hipError_t launch_gather_cols(hipStream_t s, const float* src, int64_t k, int64_t Nnm, int kpad,
                              float* dst);
hipError_t launch_fill_zero(hipStream_t s, void* p, int64_t bytes);
hipError_t launch_fill_hash(hipStream_t s, float* p, int64_t n, uint32_t seed);
// Rademacher probes i0 .. i0+k-1 into dst (column-major, ld >= N)
hipError_t launch_rademacher(hipStream_t s, float* dst, int64_t ld, int64_t N, uint64_t seed, int64_t i0, int k);
// probes of a sample subset: entry i of probe i0 + j to row rows[i] of column j
hipError_t launch_rademacher_rows(hipStream_t s, float* dst, int64_t ld, int64_t N, const int64_t* rows, int64_t n,
                                  uint64_t seed, int64_t i0, int k);
// FP64 column dots (see k_col_dots); out[j * nd + d], nd = [P1] + [P2] + m
int col_dots_max();
int64_t col_dots_partial_doubles(int64_t Npad, int k);
// pcols: partner column of each x_j in P1/P2 (nullptr: j)
hipError_t launch_col_dots(hipStream_t s, const float* X, int64_t ldx, int kpad, const int32_t* slots, int k,
                           const float* P1, const float* P2, int64_t ldp, const int32_t* pcols, const float* Q,
                           int m, int64_t ldq, int64_t Nnm, double* partial, double* out);
hipError_t launch_cols_gather_to(hipStream_t s, const float* src, int k, const int32_t* slots, const int32_t* cols,
                                 int kpad, int64_t Nnm, int64_t ld, float* dst);

}  // namespace saige
