// C ABI entry points of libsaigehip (declared in include/saige_hip.h).
//
// Host-side orchestration only: file ingest, compaction of the QC'd marker
// list, panel scheduling of the SYRK and the block-PCG slot engine.  Every
// O(N) or larger computation runs in the kernels of geno_kernels.hip,
// grm_kernels.hip and pcg_kernels.hip.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>

#include "engine_park.hpp"
#include "saige_internal.hpp"

#include <mutex>

using namespace saige;

namespace saige {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

}  // namespace saige

namespace {

#define CHECK_CTX(ctx)                                       \
    do {                                                     \
        if (!(ctx)) {                                        \
            set_error("null context");                       \
            return SAIGE_E_ARG;                              \
        }                                                    \
        hipError_t _se = hipSetDevice((ctx)->device);        \
        if (_se != hipSuccess) {                             \
            set_error("hipSetDevice: %s", hipGetErrorString(_se)); \
            return SAIGE_E_DEVICE;                           \
        }                                                    \
    } while (0)

#define CHECK_LOADED(ctx)                                                      \
    do {                                                                       \
        if (!(ctx)->loaded) {                                                  \
            set_error("no genotypes loaded: call saige_setgeno first");       \
            return SAIGE_E_STATE;                                              \
        }                                                                      \
    } while (0)

#define CHECK_GRM(ctx)                                                         \
    do {                                                                       \
        CHECK_LOADED(ctx);                                                     \
        if (!(ctx)->grm_ready) {                                               \
            set_error("GRM not built: call saige_build_grm first");           \
            return SAIGE_E_STATE;                                              \
        }                                                                      \
    } while (0)

template <typename T>
hipError_t dalloc(T** p, int64_t count) {
    *p = nullptr;
    if (count <= 0) count = 1;
    return hipMalloc((void**)p, sizeof(T) * (size_t)count);
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// the per-slot state block: lists [2][kpad] and slot systems, maxiter [kpad]
// int32, tol [kpad] double, live and kind [kpad] bytes, take lists [3][kpad]
// int32
int64_t slot_state_bytes(int kpad) { return 40 * (int64_t)kpad; }

void set_slot_state_views(Engine& e) {
    const int64_t k = e.kpad;
    uint8_t* d = e.slot_state;
    uint8_t* h = e.h_state;
    e.d_list = reinterpret_cast<int32_t*>(d);
    e.h_list = reinterpret_cast<int32_t*>(h);
    e.slot_sys = reinterpret_cast<int32_t*>(d + 8 * k);
    e.h_slot_sys = reinterpret_cast<int32_t*>(h + 8 * k);
    e.maxiter = reinterpret_cast<int32_t*>(d + 12 * k);
    e.h_maxiter = reinterpret_cast<int32_t*>(h + 12 * k);
    e.tol = reinterpret_cast<double*>(d + 16 * k);
    e.h_tol = reinterpret_cast<double*>(h + 16 * k);
    e.live = d + 24 * k;
    e.h_live = h + 24 * k;
    e.kind = d + 25 * k;
    e.h_kind = h + 25 * k;
    e.d_take = reinterpret_cast<int32_t*>(d + 28 * k);
    e.h_take = reinterpret_cast<int32_t*>(h + 28 * k);
}

// uploads the host mirror of the per-slot state (asynchronous on the
// context's stream: the mirror is left unchanged until the stream is synced)
hipError_t push_slot_state(saige_ctx* c, Engine& e) {
    return hipMemcpyAsync(e.slot_state, e.h_state, slot_state_bytes(e.kpad), hipMemcpyHostToDevice, c->stream);
}

void free_engine(Engine& e) {
    dfree(e.X); dfree(e.R); dfree(e.Z); dfree(e.P); dfree(e.AP);
    dfree(e.winv); dfree(e.minv); dfree(e.mask); dfree(e.tau);
    dfree(e.rz); dfree(e.sumr2); dfree(e.pap); dfree(e.alpha); dfree(e.beta);
    dfree(e.iter); dfree(e.done); dfree(e.slot_state);
    dfree(e.partial); dfree(e.stage); dfree(e.stageB); dfree(e.ypart);
    if (e.h_done) (void)hipHostFree(e.h_done);
    if (e.h_iter) (void)hipHostFree(e.h_iter);
    if (e.h_state) (void)hipHostFree(e.h_state);
    e.h_done = nullptr;
    e.h_iter = nullptr;
    e.h_state = nullptr;
    e.d_list = e.h_list = e.slot_sys = e.h_slot_sys = e.maxiter = e.h_maxiter = e.d_take = e.h_take = nullptr;
    e.tol = e.h_tol = nullptr;
    e.live = e.h_live = e.kind = e.h_kind = nullptr;
    e.width = e.nsys = e.kpad = 0;
}

void free_parked(saige_ctx* c) {
    std::lock_guard<std::mutex> lk(c->park_mu);
    for (auto& b : c->parked) (void)hipFree(b.second);
    c->parked.clear();
}

void free_geno(saige_ctx* c) {
    free_parked(c);
    dfree(c->d_stats); dfree(c->d_packed); dfree(c->d_packed_vr); dfree(c->d_lut); dfree(c->d_lut_vr);
    dfree(c->d_diag); dfree(c->d_grm); dfree(c->d_vec); dfree(c->d_grm_loco); dfree(c->d_diag_loco);
    dfree(c->d_mv_in); dfree(c->d_mv_blk); dfree(c->d_mv_out); dfree(c->d_mv_ypart);
    dfree(c->d_dots_partial); dfree(c->d_dots_out);
    tri_plan_free(c->tri);
    gemv_plan_free(c->triv);
    free_engine(c->eng);
    free_engine(c->eng1);
    c->h_stats.clear();
    c->qc_index.clear();
    c->vr_index.clear();
    c->loaded = false;
    c->grm_ready = false;
    c->loco_ready = false;
    c->loco_inplace = false;
    c->loco_start = c->loco_end = -1;
    c->loco_chrom = -1;
    c->loco_m = 0;
    c->N = c->M = c->Nnm = c->Npad = c->Mqc = c->Mvr = c->bpm_nm = 0;
}

int64_t count_lines(const char* path, int* err) {
    std::ifstream f(path);
    if (!f.is_open()) {
        *err = 1;
        return 0;
    }
    int64_t n = 0;
    std::string line;
    while (std::getline(f, line)) n++;   // FG.cpp:766-771, 781-786
    *err = 0;
    return n;
}

// Allocate and fill the slot engine: kpad columns, nsys systems.
int engine_reserve(saige_ctx* c, Engine& e, int kpad, int nsys) {
    if (e.kpad == kpad && e.nsys >= nsys && e.X) return SAIGE_OK;
    free_engine(e);
    const int64_t Np = c->Npad;
    SAIGE_HIP_TRY(dalloc(&e.X, Np * kpad));
    SAIGE_HIP_TRY(dalloc(&e.R, Np * kpad));
    SAIGE_HIP_TRY(dalloc(&e.Z, Np * kpad));
    SAIGE_HIP_TRY(dalloc(&e.P, Np * kpad));
    SAIGE_HIP_TRY(dalloc(&e.AP, Np * kpad));
    SAIGE_HIP_TRY(dalloc(&e.winv, Np * nsys));
    SAIGE_HIP_TRY(dalloc(&e.minv, Np * nsys));
    SAIGE_HIP_TRY(dalloc(&e.mask, Np * nsys));
    SAIGE_HIP_TRY(dalloc(&e.tau, 2 * nsys));
    SAIGE_HIP_TRY(dalloc(&e.slot_state, slot_state_bytes(kpad)));
    SAIGE_HIP_TRY(hipHostMalloc((void**)&e.h_state, slot_state_bytes(kpad), hipHostMallocDefault));
    SAIGE_HIP_TRY(dalloc(&e.rz, kpad));
    SAIGE_HIP_TRY(dalloc(&e.sumr2, kpad));
    SAIGE_HIP_TRY(dalloc(&e.pap, kpad));
    SAIGE_HIP_TRY(dalloc(&e.alpha, kpad));
    SAIGE_HIP_TRY(dalloc(&e.beta, kpad));
    SAIGE_HIP_TRY(dalloc(&e.iter, kpad));
    SAIGE_HIP_TRY(dalloc(&e.done, kpad));
    SAIGE_HIP_TRY(dalloc(&e.stage, Np));
    SAIGE_HIP_TRY(dalloc(&e.stageB, Np * kpad));
    if (kpad == EW) SAIGE_HIP_TRY(dalloc(&e.ypart, gemm_ypart_floats(Np)));
    int64_t np = std::max<int64_t>(gemv_max_parts(Np), gemm_max_parts(Np) * kpad);
    np = std::max<int64_t>(np, vec_max_parts(Np, kpad) * kpad);
    e.nparts = np;
    SAIGE_HIP_TRY(dalloc(&e.partial, np * 2));
    SAIGE_HIP_TRY(hipHostMalloc((void**)&e.h_done, kpad, hipHostMallocDefault));
    SAIGE_HIP_TRY(hipHostMalloc((void**)&e.h_iter, sizeof(int32_t) * kpad, hipHostMallocDefault));
    e.kpad = kpad;
    e.width = kpad;
    e.nsys = nsys;
    set_slot_state_views(e);
    std::memset(e.h_state, 0, (size_t)slot_state_bytes(kpad));
    // zero all column blocks so padded rows and idle columns read as zero
    hipStream_t s = c->stream;
    SAIGE_HIP_TRY(launch_fill_zero(s, e.X, Np * kpad * 4));
    SAIGE_HIP_TRY(launch_fill_zero(s, e.R, Np * kpad * 4));
    SAIGE_HIP_TRY(launch_fill_zero(s, e.Z, Np * kpad * 4));
    SAIGE_HIP_TRY(launch_fill_zero(s, e.P, Np * kpad * 4));
    SAIGE_HIP_TRY(launch_fill_zero(s, e.AP, Np * kpad * 4));
    SAIGE_HIP_TRY(launch_fill_zero(s, e.winv, Np * nsys * 4));
    SAIGE_HIP_TRY(launch_fill_zero(s, e.minv, Np * nsys * 4));
    SAIGE_HIP_TRY(launch_fill_zero(s, e.mask, round_up(Np * nsys, 4)));
    SAIGE_HIP_TRY(launch_fill_zero(s, e.stage, Np * 4));
    const std::vector<uint8_t> zeros(kpad, 0);
    SAIGE_HIP_TRY(hipMemcpyAsync(e.done, zeros.data(), kpad, hipMemcpyHostToDevice, s));
    SAIGE_HIP_TRY(push_slot_state(c, e));
    SAIGE_HIP_TRY(hipStreamSynchronize(s));
    return SAIGE_OK;
}

// The GRM a product uses: the full one (S/Mqc, diag) or the LOCO one
// (S_loco/M_loco, diag_loco; getDiagOfSigma_LOCO ignores the diag-as-one
// option, FG.cpp:2368-2392).  s16: S's values have at most 16 significant
// bits (the split GRM and its LOCO ranges are stored so), and the block GEMM
// may split them into two BF16 pieces; SAIGE_GEMM_S_PIECES=3 keeps three
// (the same bits, more work).
struct GrmView {
    const float* S;
    const float* diag;
    int64_t m;
    int diag_one;
    bool s16;
};

GrmView grm_view(const saige_ctx* c, int loco) {
    const char* pieces = std::getenv("SAIGE_GEMM_S_PIECES");
    const bool s16 = c->grm_precision == SAIGE_GRM_BF16_SPLIT && !(pieces && std::atoi(pieces) == 3);
    if (loco) return {c->loco_inplace ? c->d_grm : c->d_grm_loco, c->d_diag_loco, c->loco_m, 0, s16};
    return {c->d_grm, c->d_diag, c->Mqc, c->diag_one, s16};
}

int check_loco(const saige_ctx* c) {
    if (!c->loco_ready) {
        set_error("no LOCO range: call saige_set_start_end_index first");
        return SAIGE_E_STATE;
    }
    return SAIGE_OK;
}

int engine_set_system(saige_ctx* c, Engine& e, int s, const float* w_host, const float* tau,
                      const uint8_t* mask_host) {
    if (s < 0 || s >= e.nsys) {
        set_error("system %d out of range [0, %d)", s, e.nsys);
        return SAIGE_E_ARG;
    }
    if (!w_host || !tau) {
        set_error("w and tau are required");
        return SAIGE_E_ARG;
    }
    hipStream_t st = c->stream;
    const int64_t Np = c->Npad, N = c->Nnm;
    float* d_w = c->d_vec;                         // [Npad]
    uint8_t* d_m = reinterpret_cast<uint8_t*>(c->d_vec + Np);
    SAIGE_HIP_TRY(hipMemcpyAsync(d_w, w_host, sizeof(float) * N, hipMemcpyHostToDevice, st));
    if (mask_host) SAIGE_HIP_TRY(hipMemcpyAsync(d_m, mask_host, N, hipMemcpyHostToDevice, st));
    SAIGE_HIP_TRY(hipMemcpyAsync(e.tau + 2 * s, tau, sizeof(float) * 2, hipMemcpyHostToDevice, st));
    const GrmView g = grm_view(c, e.loco);
    SAIGE_HIP_TRY(launch_sigma_setup(st, d_w, mask_host ? d_m : nullptr, g.diag, N, Np, tau[0], tau[1],
                                     (float)g.m, g.diag_one, e.winv + s * Np, e.minv + s * Np,
                                     e.mask + s * Np, nullptr));
    SAIGE_HIP_TRY(hipStreamSynchronize(st));
    return SAIGE_OK;
}

// uploads the per-slot state after the host changed it (live slots and
// their systems)
int engine_sync_host_arrays(saige_ctx* c, Engine& e) {
    SAIGE_HIP_TRY(push_slot_state(c, e));
    SAIGE_HIP_TRY(hipStreamSynchronize(c->stream));
    return SAIGE_OK;
}

// Start k solves at once: column j of B_host (N x k, column-major) in slot
// slots[j] with system systems[j].  One upload, two launches, one sync.
int engine_add_batch(saige_ctx* c, Engine& e, int k, const int32_t* slots, const int32_t* systems,
                     const float* B_host, int maxiter, float tol, int product = 0, bool b_on_device = false) {
    if (product && e.kpad != EW) {
        set_error("GRM-product slots need the block engine");
        return SAIGE_E_STATE;
    }
    if (k < 1 || k > e.kpad || !slots || !systems || !B_host) {
        set_error("invalid batch of %d columns (engine width %d)", k, e.kpad);
        return SAIGE_E_ARG;
    }
    std::vector<uint8_t> seen(e.kpad, 0);
    for (int j = 0; j < k; j++) {
        const int slot = slots[j], sy = systems[j];
        if (slot < 0 || slot >= e.kpad) {
            set_error("slot %d out of range [0, %d)", slot, e.kpad);
            return SAIGE_E_ARG;
        }
        if (sy < 0 || sy >= e.nsys) {
            set_error("system %d out of range [0, %d)", sy, e.nsys);
            return SAIGE_E_ARG;
        }
        if (e.h_live[slot] || seen[slot]) {
            set_error("slot %d is busy", slot);
            return SAIGE_E_STATE;
        }
        seen[slot] = 1;
    }
    for (int j = 0; j < k; j++) {
        e.h_slot_sys[slots[j]] = systems[j];
        e.h_live[slots[j]] = 1;
        e.h_maxiter[slots[j]] = maxiter;
        e.h_tol[slots[j]] = (double)tol;
        e.h_kind[slots[j]] = product ? 1 : 0;
    }
    hipStream_t st = c->stream;
    const int64_t Np = c->Npad, N = c->Nnm;
    std::fill(e.h_list, e.h_list + 2 * e.kpad, 0);
    std::copy(slots, slots + k, e.h_list);
    std::copy(systems, systems + k, e.h_list + e.kpad);
    SAIGE_HIP_TRY(push_slot_state(c, e));                 // lists, live, systems, maxiter, tol, kind
    // right-hand sides: column-major N x k, uploaded, or read in place on the device
    const float* B = B_host;
    if (!b_on_device) {
        SAIGE_HIP_TRY(hipMemcpyAsync(e.stageB, B_host, sizeof(float) * N * k, hipMemcpyHostToDevice, st));
        B = e.stageB;
    }
    int64_t np = 0;
    SAIGE_HIP_TRY(launch_cols_init(st, B, k, e.d_list, e.d_list + e.kpad, e.kpad, N, Np, e.minv, e.mask, e.X,
                                   e.R, e.Z, e.P, e.partial, &np, product));
    SAIGE_HIP_TRY(launch_cols_init_finish(st, k, e.d_list, e.kpad, e.partial, np, e.rz, e.sumr2, e.maxiter, e.tol,
                                          e.iter, e.done, product));
    SAIGE_HIP_TRY(hipStreamSynchronize(st));
    return SAIGE_OK;
}

int engine_add(saige_ctx* c, Engine& e, int slot, int s, const float* b_host, int maxiter, float tol) {
    return engine_add_batch(c, e, 1, &slot, &s, b_host, maxiter, tol);
}

// items per row set left of its diagonal block in the lower-triangle
// product: at most TRI_CHUNKS chunks of 32 columns (3072 columns) per item
constexpr int TRI_CHUNKS = 96;

// the lower-triangle plan for the context's Npad and wave count (rt 1: 12
// waves, rt 2: 8), when that mode is on
int ensure_tri(saige_ctx* c) {
    if (c->gemm_cfg.mode != GEMM_BF16X3_TRI) return SAIGE_OK;
    const hipError_t e = tri_plan_build(c->tri, c->Npad, c->gemm_cfg.rt == 2 ? 8 : 12, TRI_CHUNKS);
    if (e != hipSuccess) {
        saige::set_error("lower-triangle product scratch for Npad=%lld: %s", (long long)c->Npad,
                         hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? SAIGE_E_NOMEM : SAIGE_E_DEVICE;
    }
    return SAIGE_OK;
}

// items of the lower-triangle GEMV: at most TRIV_CHUNKS chunks of 256
// columns left of a row set's diagonal block
constexpr int TRIV_CHUNKS = 24;

// the lower-triangle GEMV's plan when that mode is on (else nullptr); rc != 0
// on an allocation failure
const GemvPlan* gemv_plan(saige_ctx* c, int* rc) {
    *rc = SAIGE_OK;
    if (c->gemv_mode != SAIGE_GEMV_TRI) return nullptr;
    const hipError_t e = gemv_plan_build(c->triv, c->Npad, TRIV_CHUNKS);
    if (e != hipSuccess) {
        saige::set_error("lower-triangle GEMV scratch for Npad=%lld: %s", (long long)c->Npad, hipGetErrorString(e));
        *rc = e == hipErrorOutOfMemory ? SAIGE_E_NOMEM : SAIGE_E_DEVICE;
    }
    return &c->triv;
}

int engine_step(saige_ctx* c, Engine& e, int n_iter, uint8_t* done_out, int32_t* iters_out) {
    hipStream_t st = c->stream;
    const int64_t Np = c->Npad, N = c->Nnm;
    if (e.kpad != 1)
        if (int rc = ensure_tri(c)) return rc;
    int prc = 0;
    const GemvPlan* tv = e.kpad == 1 ? gemv_plan(c, &prc) : nullptr;
    if (prc) return prc;
    SigmaArgs a;
    a.winv = e.winv;
    a.mask = e.mask;
    a.tau = e.tau;
    a.sys = e.slot_sys;
    a.live = e.live;
    const GrmView g = grm_view(c, e.loco);
    a.inv_m = 0.0f;
    a.mqc = (float)g.m;
    a.mode = 1;
    a.kind = e.kind;
    a.xout = e.X;
    SAIGE_HIP_TRY(hipMemcpyAsync(e.h_done, e.done, e.kpad, hipMemcpyDeviceToHost, st));
    SAIGE_HIP_TRY(hipStreamSynchronize(st));
    // 32-column halves of the block that hold an occupied slot; the GEMM skips
    // the others (a done slot that is not yet taken is recomputed unchanged)
    int halves = 0, nlive = 0;
    for (int k = 0; k < e.kpad; k++)
        if (e.h_live[k]) {
            halves |= 1 << (k / 32);
            nlive++;
        }
    if (halves == 0) halves = 1;
    // the lower-triangle product runs at most 16 (32) occupied slots,
    // wherever they sit, as one 16- (32-) column launch: the occupied slots
    // in order, then free ones (computed, not used)
    NarrowCols nc;
    bool narrow = false;
    if (e.kpad == EW && c->gemm_cfg.mode == GEMM_BF16X3_TRI && nlive > 0 && nlive <= 32) {
        nc.width = nlive <= 16 ? 16 : 32;
        int n = 0;
        for (int k = 0; k < e.kpad; k++)
            if (e.h_live[k]) nc.col[n++] = (uint8_t)k;
        for (int k = 0; k < e.kpad && n < nc.width; k++)
            if (!e.h_live[k]) nc.col[n++] = (uint8_t)k;
        narrow = true;
    }
    const bool quarter = narrow && nc.width == 16;
    if (narrow && !quarter) halves = 1;   // for the shape and the statistics below
    for (int it = 0; it < n_iter; it++) {
        bool any = false;
        for (int k = 0; k < e.kpad; k++) any |= (e.h_live[k] && !e.h_done[k]);
        if (!any) break;
        int64_t np = 0;
        SAIGE_HIP_TRY(hipEventRecord(c->ev0, st));
        if (e.kpad == 1)
            SAIGE_HIP_TRY(launch_grm_gemv(st, g.S, Np, N, e.P, e.AP, a, e.partial, &np, tv));
        else
            SAIGE_HIP_TRY(launch_grm_gemm(st, g.S, Np, N, e.P, e.AP, a, e.partial, &np, e.ypart,
                                          quarter       ? c->gemm_cfg_quarter
                                          : halves == 3 ? c->gemm_cfg
                                                        : c->gemm_cfg_half,
                                          halves, &c->tri, g.s16, narrow ? &nc : nullptr));
        SAIGE_HIP_TRY(hipEventRecord(c->ev1, st));
        SAIGE_HIP_TRY(launch_alpha(st, e.kpad, e.partial, np, e.rz, e.done, e.live, e.kind, e.alpha));
        int64_t np2 = 0;
        SAIGE_HIP_TRY(launch_update(st, e.kpad, N, Np, e.alpha, e.done, e.live, e.kind, e.minv, e.slot_sys, e.X, e.R,
                                    e.Z, e.P, e.AP, e.partial, &np2));
        SAIGE_HIP_TRY(launch_beta(st, e.kpad, e.partial, np2, e.rz, e.sumr2, e.maxiter, e.tol, e.iter, e.done,
                                  e.live, e.kind, e.beta));
        SAIGE_HIP_TRY(launch_pupdate(st, e.kpad, N, Np, e.beta, e.done, e.live, e.kind, e.Z, e.P));
        SAIGE_HIP_TRY(hipMemcpyAsync(e.h_done, e.done, e.kpad, hipMemcpyDeviceToHost, st));
        SAIGE_HIP_TRY(hipStreamSynchronize(st));
        float ms = 0.f;
        SAIGE_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        const int kind = e.kpad == 1 ? SAIGE_LAUNCH_GEMV
                         : quarter      ? SAIGE_LAUNCH_GEMM_QUARTER
                         : halves == 3  ? SAIGE_LAUNCH_GEMM_FULL
                                        : SAIGE_LAUNCH_GEMM_HALF;
        e.gemm_seconds[kind] += ms * 1e-3;
        e.gemm_launches[kind]++;
    }
    SAIGE_HIP_TRY(hipMemcpyAsync(e.h_iter, e.iter, 4 * e.kpad, hipMemcpyDeviceToHost, st));
    SAIGE_HIP_TRY(hipMemcpyAsync(e.h_done, e.done, e.kpad, hipMemcpyDeviceToHost, st));
    SAIGE_HIP_TRY(hipStreamSynchronize(st));
    for (int k = 0; k < e.kpad; k++) {
        if (done_out) done_out[k] = e.h_live[k] ? e.h_done[k] : 1;
        if (iters_out) iters_out[k] = e.h_live[k] ? e.h_iter[k] : 0;
    }
    return SAIGE_OK;
}

// Copy the solutions of k slots (X_host: N x k, column-major; may be NULL)
// and free the slots.
int engine_take_batch(saige_ctx* c, Engine& e, int k, const int32_t* slots, float* X_host) {
    if (k < 1 || k > e.kpad || !slots) {
        set_error("invalid take of %d columns", k);
        return SAIGE_E_ARG;
    }
    for (int j = 0; j < k; j++)
        if (slots[j] < 0 || slots[j] >= e.kpad) {
            set_error("slot %d out of range", slots[j]);
            return SAIGE_E_ARG;
        }
    hipStream_t st = c->stream;
    // the slots' list and their freed state in one upload (the gather does
    // not read the live flags)
    std::fill(e.h_list, e.h_list + 2 * e.kpad, 0);
    std::copy(slots, slots + k, e.h_list);
    for (int j = 0; j < k; j++) e.h_live[slots[j]] = 0;
    SAIGE_HIP_TRY(push_slot_state(c, e));
    if (X_host) {
        SAIGE_HIP_TRY(launch_cols_gather(st, e.X, k, e.d_list, e.kpad, c->Nnm, e.stageB));
        SAIGE_HIP_TRY(hipMemcpyAsync(X_host, e.stageB, sizeof(float) * c->Nnm * k, hipMemcpyDeviceToHost, st));
    }
    SAIGE_HIP_TRY(hipStreamSynchronize(st));
    return SAIGE_OK;
}

int ensure_dots_scratch(saige_ctx* c) {
    if (c->d_dots_partial) return SAIGE_OK;
    if (dalloc(&c->d_dots_partial, col_dots_partial_doubles(c->Npad, EW)) ||
        dalloc(&c->d_dots_out, (int64_t)EW * col_dots_max())) {
        dfree(c->d_dots_partial);
        dfree(c->d_dots_out);
        set_error("device allocation failed (dot scratch)");
        return SAIGE_E_NOMEM;
    }
    return SAIGE_OK;
}

// FP64 dots of k columns (engine slots or a column-major device buffer) with
// pairwise and shared partner columns; the k x nd results go to dots_host.
int column_dots(saige_ctx* c, const float* X, int64_t ldx, int kpad, const int32_t* d_slots, int k,
                const float* P1, const float* P2, int64_t ldp, const int32_t* d_pcols, const float* Q, int m,
                int64_t ldq, double* dots_host) {
    const int nd = (P1 ? 1 : 0) + (P2 ? 1 : 0) + m;
    if (nd == 0) return SAIGE_OK;
    if (m < 0 || nd > col_dots_max() || k < 1 || k > EW || !dots_host) {
        set_error("invalid dot request: %d columns x %d partners", k, nd);
        return SAIGE_E_ARG;
    }
    if ((P1 || P2) && ldp < c->Nnm) {
        set_error("pair buffer leading dimension %lld < N", (long long)ldp);
        return SAIGE_E_ARG;
    }
    if (m > 0 && (!Q || ldq < c->Nnm)) {
        set_error("shared partner columns missing or leading dimension < N");
        return SAIGE_E_ARG;
    }
    if (int rc = ensure_dots_scratch(c)) return rc;
    hipStream_t st = c->stream;
    SAIGE_HIP_TRY(launch_col_dots(st, X, ldx, kpad, d_slots, k, P1, P2, ldp, d_pcols, Q, m, ldq, c->Nnm,
                                  c->d_dots_partial, c->d_dots_out));
    SAIGE_HIP_TRY(hipMemcpyAsync(dots_host, c->d_dots_out, sizeof(double) * k * nd, hipMemcpyDeviceToHost, st));
    SAIGE_HIP_TRY(hipStreamSynchronize(st));
    return SAIGE_OK;
}

// engine_take_batch with device-side options: the results may go to host
// (X_host), to a column-major device buffer (X_dev, ld ldx) or nowhere, and
// FP64 dots with partner columns are returned before the slots are freed.
int engine_take_batch_ex(saige_ctx* c, Engine& e, int k, const int32_t* slots, float* X_host, float* X_dev,
                         int64_t ldx, const int32_t* xcols, const float* P1, const float* P2, int64_t ldp,
                         const int32_t* pcols, const float* Q, int m, int64_t ldq, double* dots_host) {
    if (k < 1 || k > e.kpad || e.kpad != EW || !slots) {
        set_error("invalid take of %d columns (block engine only)", k);
        return SAIGE_E_ARG;
    }
    for (int j = 0; j < k; j++) {
        if (slots[j] < 0 || slots[j] >= e.kpad || !e.h_live[slots[j]]) {
            set_error("slot %d is not a live slot", slots[j]);
            return SAIGE_E_ARG;
        }
        if ((xcols && xcols[j] < 0) || (pcols && pcols[j] < 0)) {
            set_error("negative column index");
            return SAIGE_E_ARG;
        }
    }
    if (X_dev && ldx < c->Nnm) {
        set_error("destination leading dimension %lld < N", (long long)ldx);
        return SAIGE_E_ARG;
    }
    if (int rc = ensure_dots_scratch(c)) return rc;
    hipStream_t st = c->stream;
    // the take's lists and the slots' freed state in one upload (the dots
    // and gathers do not read the live flags)
    std::fill(e.h_take, e.h_take + 3 * EW, 0);
    for (int j = 0; j < k; j++) {
        e.h_take[j] = slots[j];
        e.h_take[EW + j] = xcols ? xcols[j] : j;
        e.h_take[2 * EW + j] = pcols ? pcols[j] : j;
        e.h_live[slots[j]] = 0;
    }
    int32_t* d_slots = e.d_take;
    SAIGE_HIP_TRY(push_slot_state(c, e));
    if (int rc = column_dots(c, e.X, 0, e.kpad, d_slots, k, P1, P2, ldp, d_slots + 2 * EW, Q, m, ldq, dots_host))
        return rc;
    if (X_dev) SAIGE_HIP_TRY(launch_cols_gather_to(st, e.X, k, d_slots, d_slots + EW, e.kpad, c->Nnm, ldx, X_dev));
    if (X_host) {
        SAIGE_HIP_TRY(launch_cols_gather(st, e.X, k, d_slots, e.kpad, c->Nnm, e.stageB));
        SAIGE_HIP_TRY(hipMemcpyAsync(X_host, e.stageB, sizeof(float) * c->Nnm * k, hipMemcpyDeviceToHost, st));
    }
    SAIGE_HIP_TRY(hipStreamSynchronize(st));
    return SAIGE_OK;
}

int engine_take(saige_ctx* c, Engine& e, int slot, float* x_host) {
    return engine_take_batch(c, e, 1, &slot, x_host);
}

// Common tail of the three setgeno variants: decode, compact, repack.
int setgeno_device_impl(saige_ctx* c, const uint8_t* d_bed, int64_t N, int64_t M,
                        const int32_t* subidx, int64_t n_sub, const uint8_t* indicator, int diag_one) {
    if (N <= 0 || M <= 0 || n_sub <= 0 || n_sub > N || !subidx) {
        set_error("invalid sizes: N=%lld M=%lld n_sub=%lld", (long long)N, (long long)M, (long long)n_sub);
        return SAIGE_E_ARG;
    }
    if (N > (int64_t)1 << 30 || n_sub * 2 > (int64_t)INT32_MAX) {
        set_error("N=%lld exceeds the supported sample count", (long long)N);
        return SAIGE_E_ARG;
    }
    bool identity = (n_sub == N);
    for (int64_t i = 0; i < n_sub; i++) {
        if (subidx[i] < 1 || subidx[i] > N) {
            set_error("subSampleInGeno[%lld] = %d outside 1..%lld", (long long)i, subidx[i], (long long)N);
            return SAIGE_E_ARG;
        }
        if (subidx[i] != i + 1) identity = false;
    }
    free_geno(c);
    hipStream_t st = c->stream;
    c->N = N;
    c->M = M;
    c->Nnm = n_sub;
    c->Npad = round_up(n_sub, NPAD_ALIGN);
    c->bpm_nm = (n_sub + 3) / 4;
    c->diag_one = diag_one;
    const int64_t bpm = (N + 3) / 4;

    std::vector<uint8_t> ind(N, 0);
    if (indicator) std::memcpy(ind.data(), indicator, N);
    else for (int64_t i = 0; i < n_sub; i++) ind[subidx[i] - 1] = 1;
    std::vector<uint8_t> rset(M, 0);
    for (int32_t v : c->vr_rand)
        if (v >= 0 && v < M) rset[v] = 1;

    // the decode's sample mask: none when every .fam sample is phenotyped,
    // else the rows' 2-bit code mask in 16 copies, copy r shifted by r bytes,
    // so that it lines up with the decode's 16-byte aligned loads of a row
    // whose address is r mod 16
    bool all_in = true;
    for (int64_t i = 0; i < N && all_in; i++) all_in = ind[i] != 0;
    const int64_t ldm = all_in ? 0 : round_up(bpm + 16, 16);
    std::vector<uint8_t> imask;
    if (!all_in) {
        std::vector<uint8_t> mb(bpm, 0);
        for (int64_t i = 0; i < N; i++)
            if (ind[i]) mb[i / 4] |= (uint8_t)(3u << (2 * (i % 4)));
        imask.assign(16 * ldm, 0);
        for (int64_t r = 0; r < 16; r++) std::memcpy(imask.data() + r * ldm + r, mb.data(), bpm);
    }

    uint8_t *d_ind = nullptr, *d_rset = nullptr;
    int32_t *d_sub = nullptr, *d_list = nullptr;
    auto cleanup = [&]() { dfree(d_ind); dfree(d_rset); dfree(d_sub); dfree(d_list); };
    int rc = SAIGE_OK;
    do {
        if (dalloc(&d_ind, all_in ? 16 : 16 * ldm) || dalloc(&d_rset, M) || dalloc(&d_sub, n_sub) ||
            dalloc(&c->d_stats, M)) {
            set_error("device allocation failed (setgeno)");
            rc = SAIGE_E_NOMEM;
            break;
        }
        if ((!all_in && hipMemcpyAsync(d_ind, imask.data(), 16 * ldm, hipMemcpyHostToDevice, st)) ||
            hipMemcpyAsync(d_rset, rset.data(), M, hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(d_sub, subidx, 4 * n_sub, hipMemcpyHostToDevice, st)) {
            set_error("host to device copy failed (setgeno)");
            rc = SAIGE_E_DEVICE;
            break;
        }
        DecodeParams p;
        p.N = N;
        p.M = M;
        p.Nnm = n_sub;
        p.bpm = bpm;
        p.min_maf = c->min_maf;
        p.max_missing = c->max_missing;
        p.vr_mode = c->is_var_ratio ? (c->max_mac_vr != -1.0f ? 1 : 2) : 0;
        p.min_mac_vr = c->min_mac_vr;
        p.max_mac_vr = c->max_mac_vr;
        hipError_t e = launch_decode_stats(st, d_bed, all_in ? nullptr : d_ind, ldm, d_rset, p, c->d_stats);
        c->h_stats.resize(M);
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->h_stats.data(), c->d_stats, sizeof(MarkerStat) * M, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            set_error("decode failed: %s", hipGetErrorString(e));
            rc = SAIGE_E_DEVICE;
            break;
        }
        // compaction in marker order (SNPIdx_new / SNPIdx_vr, FG.cpp:923-943)
        for (int64_t m = 0; m < M; m++) {
            if (c->h_stats[m].pass_qc) c->qc_index.push_back((int32_t)m);
            if (c->h_stats[m].pass_vr) c->vr_index.push_back((int32_t)m);
        }
        c->Mqc = (int64_t)c->qc_index.size();
        c->Mvr = (int64_t)c->vr_index.size();
        const int64_t nl = std::max<int64_t>(c->Mqc, c->Mvr);
        if (dalloc(&d_list, nl) || dalloc(&c->d_packed, c->Mqc * c->bpm_nm) ||
            dalloc(&c->d_packed_vr, c->Mvr * c->bpm_nm) || dalloc(&c->d_lut, c->Mqc) ||
            dalloc(&c->d_lut_vr, c->Mvr) || dalloc(&c->d_diag, c->Npad) || dalloc(&c->d_vec, 4 * c->Npad)) {
            set_error("device allocation failed (packed genotypes)");
            rc = SAIGE_E_NOMEM;
            break;
        }
        e = launch_fill_zero(st, c->d_diag, 4 * c->Npad);
        if (e == hipSuccess && c->Mqc > 0)
            e = hipMemcpyAsync(d_list, c->qc_index.data(), 4 * c->Mqc, hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = launch_repack(st, d_bed, bpm, d_list, c->Mqc, c->d_stats, identity ? nullptr : d_sub, n_sub,
                              c->bpm_nm, c->d_packed);
        if (e == hipSuccess) e = launch_lut(st, d_list, c->Mqc, c->d_stats, c->d_lut);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e == hipSuccess && c->Mvr > 0) {
            e = hipMemcpyAsync(d_list, c->vr_index.data(), 4 * c->Mvr, hipMemcpyHostToDevice, st);
            if (e == hipSuccess)
                e = launch_repack(st, d_bed, bpm, d_list, c->Mvr, c->d_stats, identity ? nullptr : d_sub,
                                  n_sub, c->bpm_nm, c->d_packed_vr);
            if (e == hipSuccess) e = launch_lut(st, d_list, c->Mvr, c->d_stats, c->d_lut_vr);
        }
        // diag(Z Z^T) in FP32 marker order (Get_Diagof_StdGeno, FG.cpp:665-704)
        if (e == hipSuccess && c->Mqc > 0)
            e = launch_diag_std(st, c->d_packed, c->d_lut, c->Mqc, n_sub, c->bpm_nm, c->d_diag);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            set_error("repack failed: %s", hipGetErrorString(e));
            rc = SAIGE_E_DEVICE;
            break;
        }
    } while (0);
    cleanup();
    if (rc != SAIGE_OK) {
        free_geno(c);
        return rc;
    }
    c->loaded = true;
    return SAIGE_OK;
}

int read_bed_body(const char* bed_file, int64_t N, int64_t M, std::vector<uint8_t>& body) {
    std::ifstream f(bed_file, std::ios::binary);
    if (!f.is_open()) {
        set_error("cannot open bed file %s", bed_file);
        return SAIGE_E_IO;
    }
    unsigned char magic[3] = {0, 0, 0};
    f.read(reinterpret_cast<char*>(magic), 3);
    if (!f || magic[0] != 0x6c || magic[1] != 0x1b || magic[2] != 0x01) {
        set_error("%s is not a SNP-major PLINK-1 .bed (magic 6c 1b 01)", bed_file);
        return SAIGE_E_IO;
    }
    const int64_t bytes = M * ((N + 3) / 4);
    body.resize((size_t)bytes);
    // one bulk read instead of a seekg + read per marker (FG.cpp:902-903)
    f.read(reinterpret_cast<char*>(body.data()), bytes);
    if (f.gcount() != bytes) {
        set_error("%s: expected %lld genotype bytes, read %lld", bed_file, (long long)bytes,
                  (long long)f.gcount());
        return SAIGE_E_IO;
    }
    return SAIGE_OK;
}

}  // namespace

extern "C" {

// Adds an in-place LOCO range back to S (no-op otherwise); every entry point
// that uses the full GRM calls it first.
static int restore_full_grm(saige_ctx* c);

const char* saige_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------ allocation-free dots
// The scratch of saige_dots_device_noalloc: FP64 partial sums of k_col_dots
// (one per 256-row block, column and partner) and the finished dots, sized
// for NOALLOC_MAX_N samples, 64 columns and 16 partners (8.4 MB).  Each
// device holds its own copy for the life of the process; host code reaches
// it only through the address hipGetSymbolAddress returns for that device
// (noalloc_scratch_ptr), never through the array's name.
constexpr int64_t NOALLOC_MAX_N = 262144;
constexpr int64_t NOALLOC_OUT = (int64_t)EW * 16;
constexpr int64_t NOALLOC_PARTIAL = (NOALLOC_MAX_N / 256) * NOALLOC_OUT;
static __device__ __align__(256) double scratch[NOALLOC_PARTIAL + NOALLOC_OUT];

static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch_ptr;   // per device, filled once with that device current

// the device address of `scratch` on the current device (device `dev`)
static int noalloc_scratch_ptr(int dev, double** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[(size_t)dev]) {
        // the symbol's address on device `dev`: looked up with it current
        // (the calling thread's device is put back afterwards)
        int prev = -1;
        if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(dev) != hipSuccess) {
            set_error("cannot make device %d current for the scratch lookup", dev);
            return SAIGE_E_DEVICE;
        }
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != dev) (void)hipSetDevice(prev);
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return SAIGE_E_DEVICE;
        }
        g_scratch_ptr[(size_t)dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[(size_t)dev];
    return SAIGE_OK;
}

int saige_create(int device, saige_ctx** out) {
    if (!out) {
        set_error("out is null");
        return SAIGE_E_ARG;
    }
    *out = nullptr;
    int ndev = 0;
    SAIGE_HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        set_error("device %d out of range (%d devices)", device, ndev);
        return SAIGE_E_ARG;
    }
    SAIGE_HIP_TRY(hipSetDevice(device));
    saige_ctx* c = new saige_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        set_error("stream/event creation failed: %s", hipGetErrorString(e));
        delete c;
        return SAIGE_E_DEVICE;
    }
    // the allocation-free entry point's scratch address on this device,
    // looked up once per device while it is current; if the lookup fails,
    // only saige_dots_device_noalloc reports it
    if (noalloc_scratch_ptr(device, &c->d_noalloc) != SAIGE_OK) c->d_noalloc = nullptr;
    *out = c;
    return SAIGE_OK;
}

int saige_destroy(saige_ctx* c) {
    if (!c) return SAIGE_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_geno(c);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SAIGE_OK;
}

int saige_set_min_maf_for_grm(saige_ctx* c, float v) {
    CHECK_CTX(c);
    c->min_maf = v;
    return SAIGE_OK;
}

int saige_set_max_missing_rate_for_grm(saige_ctx* c, float v) {
    CHECK_CTX(c);
    c->max_missing = v;
    return SAIGE_OK;
}

int saige_set_min_mac_variance_ratio(saige_ctx* c, float mn, float mx, int is_vr) {
    CHECK_CTX(c);
    c->min_mac_vr = mn;
    c->max_mac_vr = mx;
    c->is_var_ratio = is_vr ? 1 : 0;
    return SAIGE_OK;
}

int saige_set_vr_random_markers(saige_ctx* c, const int32_t* idx, int64_t n) {
    CHECK_CTX(c);
    if (n < 0 || (n > 0 && !idx)) {
        set_error("invalid random marker list");
        return SAIGE_E_ARG;
    }
    c->vr_rand.assign(idx, idx + n);
    return SAIGE_OK;
}

int saige_setgeno(saige_ctx* c, const char* bed, const char* bim, const char* fam, const int32_t* subidx,
                  int64_t n_sub, const uint8_t* indicator, int diag_one) {
    CHECK_CTX(c);
    if (!bed || !bim || !fam) {
        set_error("bed, bim and fam paths are required");
        return SAIGE_E_ARG;
    }
    int err = 0;
    const int64_t N = count_lines(fam, &err);
    if (err) {
        set_error("Error! fam file not open! (%s)", fam);
        return SAIGE_E_IO;
    }
    const int64_t M = count_lines(bim, &err);
    if (err) {
        set_error("Error! bim file not open! (%s)", bim);
        return SAIGE_E_IO;
    }
    std::vector<uint8_t> body;
    int rc = read_bed_body(bed, N, M, body);
    if (rc) return rc;
    return saige_setgeno_host(c, body.data(), N, M, subidx, n_sub, indicator, diag_one);
}

int saige_setgeno_host(saige_ctx* c, const uint8_t* body, int64_t N, int64_t M, const int32_t* subidx,
                       int64_t n_sub, const uint8_t* indicator, int diag_one) {
    CHECK_CTX(c);
    if (!body || N <= 0 || M <= 0) {
        set_error("invalid bed body");
        return SAIGE_E_ARG;
    }
    uint8_t* d_bed = nullptr;
    const int64_t bytes = M * ((N + 3) / 4);
    SAIGE_HIP_TRY(dalloc(&d_bed, bytes));
    hipError_t e = hipMemcpy(d_bed, body, (size_t)bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        dfree(d_bed);
        set_error("bed upload failed: %s", hipGetErrorString(e));
        return SAIGE_E_DEVICE;
    }
    int rc = setgeno_device_impl(c, d_bed, N, M, subidx, n_sub, indicator, diag_one);
    dfree(d_bed);
    return rc;
}

int saige_setgeno_device(saige_ctx* c, const void* d_bed, int64_t N, int64_t M, const int32_t* subidx,
                         int64_t n_sub, const uint8_t* indicator, int diag_one) {
    CHECK_CTX(c);
    if (!d_bed) {
        set_error("bed_body_dev is null");
        return SAIGE_E_ARG;
    }
    return setgeno_device_impl(c, static_cast<const uint8_t*>(d_bed), N, M, subidx, n_sub, indicator,
                               diag_one);
}

int64_t saige_get_nnomissing(const saige_ctx* c) { return c ? c->Nnm : 0; }
int64_t saige_get_total_markers(const saige_ctx* c) { return c ? c->M : 0; }
int64_t saige_get_mqc(const saige_ctx* c) { return c ? c->Mqc : 0; }
int64_t saige_get_n_vr_markers(const saige_ctx* c) { return c ? c->Mvr : 0; }

int saige_get_marker_stats(saige_ctx* c, int32_t* ac, int32_t* miss, int32_t* mac, int32_t* fill, float* freq,
                           float* mr, float* inv_std, uint8_t* flags) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    for (int64_t m = 0; m < c->M; m++) {
        const MarkerStat& s = c->h_stats[m];
        if (ac) ac[m] = s.allele_count;
        if (miss) miss[m] = s.num_missing;
        if (mac) mac[m] = s.mac;
        if (fill) fill[m] = s.fill;
        if (freq) freq[m] = s.freq;
        if (mr) mr[m] = s.missing_rate;
        if (inv_std) inv_std[m] = s.inv_std;
        if (flags) flags[m] = (uint8_t)(s.pass_qc | (s.pass_vr << 1));
    }
    return SAIGE_OK;
}

int saige_get_allele_freq(saige_ctx* c, float* out) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    for (int64_t j = 0; j < c->Mqc; j++) out[j] = c->h_stats[c->qc_index[j]].freq;
    return SAIGE_OK;
}

int saige_get_mac(saige_ctx* c, int32_t* out) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    for (int64_t j = 0; j < c->Mqc; j++) out[j] = c->h_stats[c->qc_index[j]].mac;
    return SAIGE_OK;
}

int saige_get_inv_std(saige_ctx* c, float* out) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    for (int64_t j = 0; j < c->Mqc; j++) out[j] = c->h_stats[c->qc_index[j]].inv_std;
    return SAIGE_OK;
}

int saige_get_qc_marker_index(saige_ctx* c, uint8_t* out) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    for (int64_t m = 0; m < c->M; m++) out[m] = c->h_stats[m].pass_qc;
    return SAIGE_OK;
}

int saige_get_mac_vr(saige_ctx* c, int32_t* out) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    for (int64_t j = 0; j < c->Mvr; j++) out[j] = c->h_stats[c->vr_index[j]].mac;
    return SAIGE_OK;
}

int saige_get_index_vr(saige_ctx* c, int32_t* out) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    for (int64_t j = 0; j < c->Mvr; j++) out[j] = c->vr_index[j];
    return SAIGE_OK;
}

static int get_geno_impl(saige_ctx* c, const uint8_t* packed, int64_t n, int64_t idx, int32_t* out) {
    if (idx < 0 || idx >= n) {
        set_error("marker index %lld out of range [0, %lld)", (long long)idx, (long long)n);
        return SAIGE_E_ARG;
    }
    int32_t* d = reinterpret_cast<int32_t*>(c->d_vec);
    SAIGE_HIP_TRY(launch_get_geno(c->stream, packed + idx * c->bpm_nm, c->Nnm, d));
    SAIGE_HIP_TRY(hipMemcpyAsync(out, d, 4 * c->Nnm, hipMemcpyDeviceToHost, c->stream));
    SAIGE_HIP_TRY(hipStreamSynchronize(c->stream));
    return SAIGE_OK;
}

int saige_get_one_snp_geno(saige_ctx* c, int64_t idx, int32_t* out) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    return get_geno_impl(c, c->d_packed, c->Mqc, idx, out);
}

int saige_get_one_snp_geno_vr(saige_ctx* c, int64_t idx, int32_t* out) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    return get_geno_impl(c, c->d_packed_vr, c->Mvr, idx, out);
}

int saige_get_one_snp_std_geno(saige_ctx* c, int64_t idx, float* out) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    if (idx < 0 || idx >= c->Mqc) {
        set_error("marker index %lld out of range", (long long)idx);
        return SAIGE_E_ARG;
    }
    SAIGE_HIP_TRY(launch_std_geno(c->stream, c->d_packed + idx * c->bpm_nm, c->d_lut + idx, c->Nnm, c->d_vec));
    SAIGE_HIP_TRY(hipMemcpyAsync(out, c->d_vec, 4 * c->Nnm, hipMemcpyDeviceToHost, c->stream));
    SAIGE_HIP_TRY(hipStreamSynchronize(c->stream));
    return SAIGE_OK;
}

int saige_get_diag_std(saige_ctx* c, float* out) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    SAIGE_HIP_TRY(hipMemcpy(out, c->d_diag, 4 * c->Nnm, hipMemcpyDeviceToHost));
    return SAIGE_OK;
}

int saige_get_diag_of_kin(saige_ctx* c, float* out) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    if (c->diag_one) {
        for (int64_t i = 0; i < c->Nnm; i++) out[i] = 1.0f;   // FG.cpp:4371-4372
        return SAIGE_OK;
    }
    SAIGE_HIP_TRY(hipMemcpy(out, c->d_diag, 4 * c->Nnm, hipMemcpyDeviceToHost));
    const float m = (float)c->Mqc;   // x / MminMAF (FG.cpp:4368-4369)
    for (int64_t i = 0; i < c->Nnm; i++) out[i] = out[i] / m;
    return SAIGE_OK;
}

// The 2-bit-input SYRK over markers [j0, j1) into S: one pass with the BF16
// table (hi), or, for the split GRM (llo given), hi x hi + hi x lo + lo x hi
// in one pass over both tables.  accumulate: 0 store, +1 add, -1 subtract.
static hipError_t syrk_pk_passes(hipStream_t st, const uint8_t* tp, const uint64_t* lhi, const uint64_t* llo,
                                 int64_t Mqc, int64_t Np, const int2* tiles, float* S, int accumulate, int64_t j0,
                                 int64_t j1, int tile) {
    if (!llo) return launch_syrk_pk(st, tp, lhi, Mqc, Np, tiles, S, accumulate, j0, j1, tile);
    hipError_t e = launch_syrk_pk_split(st, tp, lhi, llo, Mqc, Np, tiles, S, accumulate, j0, j1);
    // the cross terms reach a diagonal tile's mirror entries in opposite
    // orders: keep S exactly symmetric
    if (e == hipSuccess) e = launch_mirror_diag(st, S, Np, tile);
    return e;
}

int saige_build_grm(saige_ctx* c, int precision) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    if (precision != SAIGE_GRM_FP32 && precision != SAIGE_GRM_BF16 && precision != SAIGE_GRM_BF16_PANEL &&
        precision != SAIGE_GRM_BF16_SPLIT) {
        set_error("unknown GRM precision %d", precision);
        return SAIGE_E_ARG;
    }
    if (c->Mqc <= 0) {
        set_error("no marker passed QC; the GRM would be empty");
        return SAIGE_E_STATE;
    }
    hipStream_t st = c->stream;
    const int64_t Np = c->Npad;
    if (!c->d_grm) {
        hipError_t e = dalloc(&c->d_grm, Np * Np);
        if (e != hipSuccess) {
            set_error("cannot allocate the %lld x %lld FP32 GRM (%.1f GB)", (long long)Np, (long long)Np,
                      4.0 * Np * Np / 1e9);
            return SAIGE_E_NOMEM;
        }
    }
    c->grm_ready = false;
    c->loco_ready = false;   // S_loco derives from S
    c->loco_inplace = false;
    if (precision == SAIGE_GRM_BF16 || precision == SAIGE_GRM_BF16_SPLIT) {
        // 2-bit operands straight into the MFMA SYRK: no Z panel
        const bool split = precision == SAIGE_GRM_BF16_SPLIT;
        const int64_t nchunk = (c->Mqc + 63) / 64;
        uint8_t* d_tp = nullptr;
        uint64_t *d_l64 = nullptr, *d_llo = nullptr;
        int2* d_tiles = nullptr;
        const int tile = split ? 256 : syrk_pk_tile();   // the split kernel: 256 tiles only
        const std::vector<int2> order = syrk_tile_order(Np / tile, tile == 256 ? 4 : 8);
        auto release = [&]() { dfree(d_tp); dfree(d_l64); dfree(d_llo); dfree(d_tiles); };
        if (dalloc(&d_tp, (Np / syrk_tile()) * nchunk * 2048) || dalloc(&d_l64, nchunk * 64) ||
            (split && dalloc(&d_llo, nchunk * 64)) || dalloc(&d_tiles, (int64_t)order.size())) {
            release();
            set_error("cannot allocate the tiled genotypes for the SYRK");
            return SAIGE_E_NOMEM;
        }
        hipError_t e = hipMemcpyAsync(d_tiles, order.data(), sizeof(int2) * order.size(), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipEventRecord(c->ev0, st);
        if (e == hipSuccess) e = launch_pack_tiles(st, c->d_packed, c->d_lut, c->Mqc, c->Nnm, Np, c->bpm_nm, d_tp, d_l64);
        if (e == hipSuccess && split) e = launch_lut64_range(st, c->d_lut, c->Mqc, 0, c->Mqc, 1, d_llo);
        if (e == hipSuccess)
            e = syrk_pk_passes(st, d_tp, d_l64, d_llo, c->Mqc, Np, d_tiles, c->d_grm, 0, 0, -1, tile);
        if (e == hipSuccess) e = hipEventRecord(c->ev1, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        release();
        if (e != hipSuccess) {
            set_error("GRM build failed: %s", hipGetErrorString(e));
            return SAIGE_E_DEVICE;
        }
        float ms = 0.f;
        SAIGE_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->grm_seconds = ms * 1e-3;
        c->grm_ready = true;
        c->grm_precision = precision;
        return SAIGE_OK;
    }
    const bool bf16 = precision == SAIGE_GRM_BF16_PANEL;
    const int64_t esz = bf16 ? 2 : 4;
    const int64_t kc = bf16 ? syrk_kchunk_bf16() : syrk_kchunk_f32();
    size_t free_b = 0, total_b = 0;
    SAIGE_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    // panel budget: at most 32 GB and half of what is free
    int64_t budget = std::min<int64_t>((int64_t)32 << 30, (int64_t)(free_b / 2));
    int64_t pw = budget / (Np * esz);
    pw = std::max<int64_t>(kc, pw / 64 * 64);
    const int64_t full = round_up(c->Mqc, 64);
    pw = std::min(pw, full);
    // SAIGE_GRM_PANEL_MARKERS caps the panel width (tests use it to force
    // several panels, i.e. the accumulate path of the SYRK)
    if (const char* env = std::getenv("SAIGE_GRM_PANEL_MARKERS")) {
        const int64_t cap = std::max<int64_t>(64, std::atoll(env) / 64 * 64);
        pw = std::min(pw, cap);
    }
    void* d_panel = nullptr;
    SAIGE_HIP_TRY(hipMalloc(&d_panel, (size_t)(Np * pw * esz)));
    const std::vector<int2> order = syrk_tile_order(Np / syrk_tile(), 8);
    int2* d_tiles = nullptr;
    if (dalloc(&d_tiles, (int64_t)order.size()) ||
        hipMemcpy(d_tiles, order.data(), sizeof(int2) * order.size(), hipMemcpyHostToDevice)) {
        (void)hipFree(d_panel);
        dfree(d_tiles);
        set_error("cannot stage the SYRK tile list");
        return SAIGE_E_NOMEM;
    }
    SAIGE_HIP_TRY(hipEventRecord(c->ev0, st));
    hipError_t e = hipSuccess;
    int first = 1;
    for (int64_t j0 = 0; j0 < c->Mqc && e == hipSuccess; j0 += pw) {
        const int64_t np = std::min(pw, c->Mqc - j0);
        const int64_t ldz = round_up(np, 64);
        if (bf16) {
            e = launch_std_panel_bf16(st, c->d_packed, c->d_lut, j0, np, c->Nnm, Np, c->bpm_nm, ldz,
                                      static_cast<__hip_bfloat16*>(d_panel));
            if (e == hipSuccess)
                e = launch_syrk_bf16(st, static_cast<__hip_bfloat16*>(d_panel), Np, ldz, ldz, d_tiles, c->d_grm,
                                     !first);
        } else {
            e = launch_std_panel_f32(st, c->d_packed, c->d_lut, j0, np, c->Nnm, Np, c->bpm_nm, ldz,
                                     static_cast<float*>(d_panel));
            if (e == hipSuccess)
                e = launch_syrk_f32(st, static_cast<float*>(d_panel), Np, ldz, ldz, d_tiles, c->d_grm, !first);
        }
        first = 0;
    }
    if (e == hipSuccess) e = hipEventRecord(c->ev1, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_panel);
    dfree(d_tiles);
    if (e != hipSuccess) {
        set_error("GRM build failed: %s", hipGetErrorString(e));
        return SAIGE_E_DEVICE;
    }
    float ms = 0.f;
    SAIGE_HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->grm_seconds = ms * 1e-3;
    c->grm_ready = true;
    c->grm_precision = precision;
    return SAIGE_OK;
}

double saige_last_grm_seconds(const saige_ctx* c) { return c ? c->grm_seconds : 0.0; }

int saige_grm_precision(const saige_ctx* c) { return c && c->grm_ready ? c->grm_precision : -1; }

int saige_get_grm_block(saige_ctx* c, int64_t r0, int64_t c0, int64_t nr, int64_t nc, float* out) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (int rc = restore_full_grm(c)) return rc;
    if (r0 < 0 || c0 < 0 || nr < 0 || nc < 0 || r0 + nr > c->Nnm || c0 + nc > c->Nnm) {
        set_error("GRM block out of range");
        return SAIGE_E_ARG;
    }
    SAIGE_HIP_TRY(hipMemcpy2D(out, sizeof(float) * nc, c->d_grm + r0 * c->Npad + c0, sizeof(float) * c->Npad,
                              sizeof(float) * nc, nr, hipMemcpyDeviceToHost));
    const float m = (float)c->Mqc;
    for (int64_t i = 0; i < nr * nc; i++) out[i] = out[i] / m;
    return SAIGE_OK;
}

static int grm_matvec(saige_ctx* c, int loco, const float* B, int64_t k, float* Y) {
    if (k <= 0 || !B || !Y) {
        set_error("invalid matvec arguments");
        return SAIGE_E_ARG;
    }
    hipStream_t st = c->stream;
    const int64_t Np = c->Npad, N = c->Nnm;
    const GrmView g = grm_view(c, loco);
    SigmaArgs a{};
    a.mqc = (float)g.m;
    a.mode = 0;
    if (k == 1) {
        int rc = engine_reserve(c, c->eng1, 1, 1);
        if (rc) return rc;
        Engine& e = c->eng1;
        SAIGE_HIP_TRY(hipMemcpyAsync(e.stage, B, 4 * N, hipMemcpyHostToDevice, st));
        int prc = 0;
        const GemvPlan* tv = gemv_plan(c, &prc);
        if (prc) return prc;
        SAIGE_HIP_TRY(launch_grm_gemv(st, g.S, Np, N, e.stage, e.AP, a, nullptr, nullptr, tv));
        SAIGE_HIP_TRY(hipMemcpyAsync(Y, e.AP, 4 * N, hipMemcpyDeviceToHost, st));
        SAIGE_HIP_TRY(hipStreamSynchronize(st));
        return SAIGE_OK;
    }
    if (int rc = ensure_tri(c)) return rc;
    // own scratch (not the engine's), so live engine slots are untouched
    if (!c->d_mv_in) {
        if (dalloc(&c->d_mv_in, N * EW) || dalloc(&c->d_mv_blk, Np * EW) || dalloc(&c->d_mv_out, Np * EW) ||
            dalloc(&c->d_mv_ypart, gemm_ypart_floats(Np))) {
            dfree(c->d_mv_in); dfree(c->d_mv_blk); dfree(c->d_mv_out); dfree(c->d_mv_ypart); dfree(c->d_mv_ypart);
            set_error("device allocation failed (grm_matvec)");
            return SAIGE_E_NOMEM;
        }
    }
    float *d_in = c->d_mv_in, *d_blk = c->d_mv_blk, *d_out = c->d_mv_out;
    hipError_t e = hipSuccess;
    for (int64_t k0 = 0; k0 < k && e == hipSuccess; k0 += EW) {
        const int kb = (int)std::min<int64_t>(EW, k - k0);
        e = hipMemcpyAsync(d_in, B + k0 * N, 4 * N * kb, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = launch_scatter_cols(st, d_in, kb, N, Np, EW, d_blk);
        // the lower-triangle product runs as narrow as the block: 16, 32 or
        // 64 columns
        const bool q = kb <= 16 && c->gemm_cfg.mode == GEMM_BF16X3_TRI;
        const NarrowCols q0 = NarrowCols::tile(16, 0);
        if (e == hipSuccess)
            e = launch_grm_gemm(st, g.S, Np, N, d_blk, d_out, a, nullptr, nullptr, c->d_mv_ypart,
                                q ? c->gemm_cfg_quarter : kb <= 32 ? c->gemm_cfg_half : c->gemm_cfg,
                                kb <= 32 ? 1 : 3, &c->tri, g.s16, q ? &q0 : nullptr);
        if (e == hipSuccess) e = launch_gather_cols(st, d_out, kb, N, EW, d_in);
        if (e == hipSuccess) e = hipMemcpyAsync(Y + k0 * N, d_in, 4 * N * kb, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    if (e != hipSuccess) {
        set_error("grm_matvec failed: %s", hipGetErrorString(e));
        return SAIGE_E_DEVICE;
    }
    return SAIGE_OK;
}

int saige_grm_matvec(saige_ctx* c, const float* B, int64_t k, float* Y) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (int rc = restore_full_grm(c)) return rc;
    return grm_matvec(c, 0, B, k, Y);
}

int saige_crossprod_mat_and_kin(saige_ctx* c, const float* b, float* out) {
    return saige_grm_matvec(c, b, 1, out);
}

static int get_diag_of_sigma(saige_ctx* c, int loco, const float* w, const float* tau, float* out) {
    if (!w || !tau || !out) {
        set_error("invalid arguments");
        return SAIGE_E_ARG;
    }
    hipStream_t st = c->stream;
    const int64_t Np = c->Npad, N = c->Nnm;
    float* d_w = c->d_vec;
    float* d_winv = c->d_vec + Np;
    float* d_minv = c->d_vec + 2 * Np;
    float* d_d = c->d_vec + 3 * Np;
    uint8_t* d_mask = nullptr;
    SAIGE_HIP_TRY(dalloc(&d_mask, Np));
    SAIGE_HIP_TRY(hipMemcpyAsync(d_w, w, 4 * N, hipMemcpyHostToDevice, st));
    const GrmView g = grm_view(c, loco);
    SAIGE_HIP_TRY(launch_sigma_setup(st, d_w, nullptr, g.diag, N, Np, tau[0], tau[1], (float)g.m, g.diag_one,
                                     d_winv, d_minv, d_mask, d_d));
    SAIGE_HIP_TRY(hipMemcpyAsync(out, d_d, 4 * N, hipMemcpyDeviceToHost, st));
    SAIGE_HIP_TRY(hipStreamSynchronize(st));
    dfree(d_mask);
    return SAIGE_OK;
}

int saige_get_diag_of_sigma(saige_ctx* c, const float* w, const float* tau, float* out) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    return get_diag_of_sigma(c, 0, w, tau, out);
}

static int get_crossprod(saige_ctx* c, int loco, const float* b, const float* w, const float* tau, float* out) {
    if (!b || !w || !tau || !out) {
        set_error("invalid arguments");
        return SAIGE_E_ARG;
    }
    int rc = engine_reserve(c, c->eng1, 1, 1);
    if (rc) return rc;
    Engine& e = c->eng1;
    e.loco = loco;
    rc = engine_set_system(c, e, 0, w, tau, nullptr);
    if (rc) return rc;
    hipStream_t st = c->stream;
    const int64_t Np = c->Npad, N = c->Nnm;
    e.h_slot_sys[0] = 0;
    e.h_live[0] = 1;
    rc = engine_sync_host_arrays(c, e);
    if (rc) return rc;
    SAIGE_HIP_TRY(hipMemcpyAsync(e.stage, b, 4 * N, hipMemcpyHostToDevice, st));
    SigmaArgs a;
    a.winv = e.winv;
    a.mask = e.mask;
    a.tau = e.tau;
    a.sys = e.slot_sys;
    a.live = e.live;
    const GrmView g = grm_view(c, loco);
    a.inv_m = 0.f;
    a.mqc = (float)g.m;
    a.mode = 1;
    a.kind = nullptr;
    a.xout = nullptr;
    int prc = 0;
    const GemvPlan* tv = gemv_plan(c, &prc);
    if (prc) return prc;
    SAIGE_HIP_TRY(launch_grm_gemv(st, g.S, Np, N, e.stage, e.AP, a, e.partial, nullptr, tv));
    SAIGE_HIP_TRY(hipMemcpyAsync(out, e.AP, 4 * N, hipMemcpyDeviceToHost, st));
    SAIGE_HIP_TRY(hipStreamSynchronize(st));
    e.h_live[0] = 0;
    return engine_sync_host_arrays(c, e);
}

int saige_get_crossprod(saige_ctx* c, const float* b, const float* w, const float* tau, float* out) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (int rc = restore_full_grm(c)) return rc;
    return get_crossprod(c, 0, b, w, tau, out);
}

static int pcg1(saige_ctx* c, int loco, const float* w, const float* tau, const float* b, int maxiter, float tol, float* x,
         int32_t* iters) {
    if (!w || !tau || !b || !x) {
        set_error("invalid arguments");
        return SAIGE_E_ARG;
    }
    int rc = engine_reserve(c, c->eng1, 1, 1);
    if (rc) return rc;
    Engine& e = c->eng1;
    e.loco = loco;
    if ((rc = engine_set_system(c, e, 0, w, tau, nullptr))) return rc;
    if ((rc = engine_add(c, e, 0, 0, b, maxiter, tol))) return rc;
    uint8_t done = 0;
    int32_t it = 0;
    while (!done) {
        if ((rc = engine_step(c, e, 64, &done, &it))) return rc;
    }
    if ((rc = engine_take(c, e, 0, x))) return rc;
    if (iters) *iters = it;
    if (it >= maxiter && maxiter > 0) {
        set_error("pcg did not converge. You may increase maxiter number.");
        return SAIGE_E_NOTCONV;
    }
    return SAIGE_OK;
}

int saige_pcg1(saige_ctx* c, const float* w, const float* tau, const float* b, int maxiter, float tol, float* x,
               int32_t* iters) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (int rc = restore_full_grm(c)) return rc;
    return pcg1(c, 0, w, tau, b, maxiter, tol, x, iters);
}

int saige_pcg_block(saige_ctx* c, const float* w, const float* tau, const float* B, int64_t k, int maxiter,
                    float tol, float* X, int32_t* iters) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (int rc = restore_full_grm(c)) return rc;
    if (!w || !tau || !B || !X || k <= 0) {
        set_error("invalid arguments");
        return SAIGE_E_ARG;
    }
    int rc = engine_reserve(c, c->eng, EW, std::max(1, c->eng.nsys));
    if (rc) return rc;
    Engine& e = c->eng;
    // getSigma_X solves with the full GRM whatever the engine was pointed at
    struct LocoGuard {
        Engine& e;
        int saved;
        ~LocoGuard() { e.loco = saved; }
    } guard{e, e.loco};
    e.loco = 0;
    for (int k0 = 0; k0 < e.kpad; k0++)
        if (e.h_live[k0]) {
            set_error("engine has live slots; saige_pcg_block needs an idle engine");
            return SAIGE_E_STATE;
        }
    if ((rc = engine_set_system(c, e, 0, w, tau, nullptr))) return rc;
    const int64_t N = c->Nnm;
    bool notconv = false;
    std::vector<uint8_t> done(EW);
    std::vector<int32_t> it(EW);
    for (int64_t k0 = 0; k0 < k; k0 += EW) {
        const int kb = (int)std::min<int64_t>(EW, k - k0);
        std::vector<int32_t> slots(kb), sys0(kb, 0);
        std::iota(slots.begin(), slots.end(), 0);
        if ((rc = engine_add_batch(c, e, kb, slots.data(), sys0.data(), B + k0 * N, maxiter, tol))) return rc;
        for (;;) {
            if ((rc = engine_step(c, e, 64, done.data(), it.data()))) return rc;
            bool all = true;
            for (int j = 0; j < kb; j++) all &= done[j] != 0;
            if (all) break;
        }
        for (int j = 0; j < kb; j++) {
            if (iters) iters[k0 + j] = it[j];
            if (it[j] >= maxiter && maxiter > 0) notconv = true;
        }
        if ((rc = engine_take_batch(c, e, kb, slots.data(), X + k0 * N))) return rc;
    }
    if (notconv) {
        set_error("pcg did not converge for at least one column. You may increase maxiter number.");
        return SAIGE_E_NOTCONV;
    }
    return SAIGE_OK;
}

// ---------------------------------------------------------------- LOCO

namespace {
// S (+/-)= Z_r Z_r^T over GRM markers [start, end], with the kernel and
// operand precision the resident S was built with (accumulate: 1 add, -1
// subtract).  Asynchronous on the context's stream.
hipError_t range_syrk(saige_ctx* c, int64_t start, int64_t end, float* S, int accumulate) {
    hipStream_t st = c->stream;
    const int64_t Np = c->Npad, N = c->Nnm;
    const int64_t cnt = end - start + 1;
    const bool split = c->grm_precision == SAIGE_GRM_BF16_SPLIT;
    const bool pk = c->grm_precision == SAIGE_GRM_BF16 || split;
    const int tile = split ? 256 : (pk ? syrk_pk_tile() : syrk_tile());
    const std::vector<int2> order = syrk_tile_order(Np / tile, tile == 256 ? 4 : 8);
    int2* d_tiles = nullptr;
    uint8_t* d_tp = nullptr;
    uint64_t *d_l64 = nullptr, *d_llo = nullptr;
    void* d_panel = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(st);
        dfree(d_tiles); dfree(d_tp); dfree(d_l64); dfree(d_llo);
        if (d_panel) (void)hipFree(d_panel);
    };
    hipError_t e = dalloc(&d_tiles, (int64_t)order.size());
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_tiles, order.data(), sizeof(int2) * order.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && pk) {
        const int64_t nchunk = (c->Mqc + 63) / 64;
        e = dalloc(&d_tp, (Np / syrk_tile()) * nchunk * 2048);
        if (e == hipSuccess) e = dalloc(&d_l64, nchunk * 64);
        if (e == hipSuccess && split) e = dalloc(&d_llo, nchunk * 64);
        if (e == hipSuccess) e = launch_pack_tiles(st, c->d_packed, c->d_lut, c->Mqc, N, Np, c->bpm_nm, d_tp, d_l64);
        if (e == hipSuccess) e = launch_lut64_range(st, c->d_lut, c->Mqc, start, end + 1, 0, d_l64);
        if (e == hipSuccess && split) e = launch_lut64_range(st, c->d_lut, c->Mqc, start, end + 1, 1, d_llo);
        if (e == hipSuccess)
            e = syrk_pk_passes(st, d_tp, d_l64, d_llo, c->Mqc, Np, d_tiles, S, accumulate, start, end + 1, tile);
    } else if (e == hipSuccess) {
        const bool bf16 = c->grm_precision == SAIGE_GRM_BF16_PANEL;
        const int64_t esz = bf16 ? 2 : 4;
        const int64_t kc = bf16 ? syrk_kchunk_bf16() : syrk_kchunk_f32();
        size_t free_b = 0, total_b = 0;
        e = hipMemGetInfo(&free_b, &total_b);
        int64_t pw = std::min<int64_t>((int64_t)32 << 30, (int64_t)(free_b / 2)) / (Np * esz);
        pw = std::min(std::max<int64_t>(kc, pw / 64 * 64), round_up(cnt, 64));
        if (const char* env = std::getenv("SAIGE_GRM_PANEL_MARKERS"))
            pw = std::min(pw, std::max<int64_t>(64, std::atoll(env) / 64 * 64));
        if (e == hipSuccess) e = hipMalloc(&d_panel, (size_t)(Np * pw * esz));
        for (int64_t j0 = start; j0 <= end && e == hipSuccess; j0 += pw) {
            const int64_t np = std::min(pw, end + 1 - j0);
            const int64_t ldz = round_up(np, 64);
            if (bf16) {
                e = launch_std_panel_bf16(st, c->d_packed, c->d_lut, j0, np, N, Np, c->bpm_nm, ldz,
                                          static_cast<__hip_bfloat16*>(d_panel));
                if (e == hipSuccess)
                    e = launch_syrk_bf16(st, static_cast<__hip_bfloat16*>(d_panel), Np, ldz, ldz, d_tiles, S,
                                         accumulate);
            } else {
                e = launch_std_panel_f32(st, c->d_packed, c->d_lut, j0, np, N, Np, c->bpm_nm, ldz,
                                         static_cast<float*>(d_panel));
                if (e == hipSuccess)
                    e = launch_syrk_f32(st, static_cast<float*>(d_panel), Np, ldz, ldz, d_tiles, S, accumulate);
            }
        }
    }
    release();
    return e;
}
}  // namespace

static int restore_full_grm(saige_ctx* c) {
    if (!c->loco_inplace) return SAIGE_OK;
    if (c->eng.X && c->eng.loco)
        for (int k = 0; k < c->eng.kpad; k++)
            if (c->eng.h_live[k]) {
                set_error("live LOCO slots use the in-place LOCO GRM; take them before using the full GRM");
                return SAIGE_E_STATE;
            }
    hipError_t e = range_syrk(c, c->loco_start, c->loco_end, c->d_grm, 1);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        set_error("restoring the GRM after an in-place LOCO range failed: %s", hipGetErrorString(e));
        return SAIGE_E_DEVICE;
    }
    c->loco_inplace = false;
    c->loco_ready = false;
    return SAIGE_OK;
}

int saige_set_start_end_index(saige_ctx* c, int64_t start, int64_t end, int chrom) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (start < 0 || end < start || end >= c->Mqc) {
        set_error("LOCO range [%lld, %lld] outside the %lld GRM markers", (long long)start, (long long)end,
                  (long long)c->Mqc);
        return SAIGE_E_ARG;
    }
    const int64_t cnt = end - start + 1;
    if (cnt >= c->Mqc) {
        set_error("LOCO range covers every GRM marker; the leave-out GRM would be empty");
        return SAIGE_E_ARG;
    }
    if (c->eng.X) {
        for (int k = 0; k < c->eng.kpad; k++)
            if (c->eng.h_live[k]) {
                set_error("the block engine has live slots; take them before changing the LOCO range");
                return SAIGE_E_STATE;
            }
    }
    hipStream_t st = c->stream;
    const int64_t Np = c->Npad, N = c->Nnm;
    // A second Npad^2 buffer when HBM allows (S stays intact); otherwise the
    // range is subtracted from S itself and added back before the next
    // full-GRM use (SAIGE_LOCO_INPLACE=1 forces this).
    const char* env = std::getenv("SAIGE_LOCO_INPLACE");
    const bool force_inplace = env && std::atoi(env) != 0;
    if (!c->d_diag_loco && dalloc(&c->d_diag_loco, Np)) {
        set_error("device allocation failed (LOCO diagonal)");
        return SAIGE_E_NOMEM;
    }
    bool two_buffers = !force_inplace && c->d_grm_loco != nullptr;
    if (!force_inplace && !two_buffers) {
        size_t free_b = 0, total_b = 0;
        SAIGE_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        if ((int64_t)free_b >= 4 * Np * Np + ((int64_t)4 << 30) && dalloc(&c->d_grm_loco, Np * Np) == hipSuccess)
            two_buffers = true;
    }
    c->loco_ready = false;
    float* d_rdiag = nullptr;
    if (dalloc(&d_rdiag, Np)) {
        set_error("device allocation failed (LOCO)");
        return SAIGE_E_NOMEM;
    }
    hipError_t e = hipSuccess;
    if (two_buffers) {
        if (c->loco_inplace) {   // a range left in S by an earlier in-place call goes back first
            if (int rc = restore_full_grm(c)) {
                dfree(d_rdiag);
                return rc;
            }
        }
        // S_loco = S - Z_r Z_r^T with the kernel and operand precision S was built with
        e = hipMemcpyAsync(c->d_grm_loco, c->d_grm, sizeof(float) * Np * Np, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = range_syrk(c, start, end, c->d_grm_loco, -1);
    } else {
        if (c->loco_inplace) e = range_syrk(c, c->loco_start, c->loco_end, c->d_grm, 1);
        c->loco_inplace = false;
        if (e == hipSuccess) e = range_syrk(c, start, end, c->d_grm, -1);
        if (e == hipSuccess) c->loco_inplace = true;
    }
    // diag_loco = diag - sum over the range of z^2, the range summed from zero in
    // marker order (set_Diagof_StdGeno_LOCO, FG.cpp:4947-4955)
    if (e == hipSuccess) e = launch_fill_zero(st, d_rdiag, 4 * Np);
    if (e == hipSuccess)
        e = launch_diag_std(st, c->d_packed + start * c->bpm_nm, c->d_lut + start, cnt, N, c->bpm_nm, d_rdiag);
    if (e == hipSuccess) e = launch_sub(st, c->d_diag, d_rdiag, Np, c->d_diag_loco);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dfree(d_rdiag);
    if (e != hipSuccess) {
        set_error("LOCO GRM build failed: %s", hipGetErrorString(e));
        return SAIGE_E_DEVICE;
    }
    c->loco_start = start;
    c->loco_end = end;
    c->loco_chrom = chrom;
    c->loco_m = c->Mqc - cnt;
    c->loco_ready = true;
    return SAIGE_OK;
}

int saige_loco_range(const saige_ctx* c, int64_t* start, int64_t* end, int* chrom, int64_t* m_loco) {
    if (!c) {
        set_error("null context");
        return SAIGE_E_ARG;
    }
    const bool r = c->loco_ready;
    if (start) *start = r ? c->loco_start : -1;
    if (end) *end = r ? c->loco_end : -1;
    if (chrom) *chrom = r ? c->loco_chrom : -1;
    if (m_loco) *m_loco = r ? c->loco_m : -1;
    return SAIGE_OK;
}

int saige_crossprod_mat_and_kin_loco(saige_ctx* c, const float* b, float* out) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (int rc = check_loco(c)) return rc;
    return grm_matvec(c, 1, b, 1, out);
}

int saige_get_diag_of_sigma_loco(saige_ctx* c, const float* w, const float* tau, float* out) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (int rc = check_loco(c)) return rc;
    return get_diag_of_sigma(c, 1, w, tau, out);
}

int saige_get_crossprod_loco(saige_ctx* c, const float* b, const float* w, const float* tau, float* out) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (int rc = check_loco(c)) return rc;
    return get_crossprod(c, 1, b, w, tau, out);
}

int saige_pcg1_loco(saige_ctx* c, const float* w, const float* tau, const float* b, int maxiter, float tol,
                    float* x, int32_t* iters) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (int rc = check_loco(c)) return rc;
    return pcg1(c, 1, w, tau, b, maxiter, tol, x, iters);
}

int saige_engine_reserve(saige_ctx* c, int width, int n_systems) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (width < 1 || width > EW || n_systems < 1) {
        set_error("engine width must be in [1, %d] and n_systems >= 1", EW);
        return SAIGE_E_ARG;
    }
    return engine_reserve(c, c->eng, EW, n_systems);
}

int saige_engine_set_system(saige_ctx* c, int s, const float* w, const float* tau, const uint8_t* mask) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (!c->eng.X) {
        set_error("call saige_engine_reserve first");
        return SAIGE_E_STATE;
    }
    if (int rc = c->eng.loco ? check_loco(c) : restore_full_grm(c)) return rc;
    return engine_set_system(c, c->eng, s, w, tau, mask);
}

int saige_engine_add(saige_ctx* c, int slot, int s, const float* b, int maxiter, float tol) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (!c->eng.X || !b) {
        set_error("engine not reserved or b is null");
        return SAIGE_E_STATE;
    }
    return engine_add(c, c->eng, slot, s, b, maxiter, tol);
}

int saige_engine_add_batch(saige_ctx* c, int k, const int32_t* slots, const int32_t* systems, const float* B,
                           int maxiter, float tol) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (!c->eng.X) {
        set_error("engine not reserved");
        return SAIGE_E_STATE;
    }
    return engine_add_batch(c, c->eng, k, slots, systems, B, maxiter, tol);
}

int saige_engine_add_products(saige_ctx* c, int k, const int32_t* slots, const int32_t* systems, const float* B) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (!c->eng.X) {
        set_error("engine not reserved");
        return SAIGE_E_STATE;
    }
    return engine_add_batch(c, c->eng, k, slots, systems, B, 1, 0.0f, 1);
}

int saige_engine_take_batch(saige_ctx* c, int k, const int32_t* slots, float* X) {
    CHECK_CTX(c);
    if (!c->eng.X) {
        set_error("engine not reserved");
        return SAIGE_E_STATE;
    }
    return engine_take_batch(c, c->eng, k, slots, X);
}

int saige_engine_add_batch_device(saige_ctx* c, int k, const int32_t* slots, const int32_t* systems,
                                  const float* B_dev, int maxiter, float tol) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (!c->eng.X) {
        set_error("engine not reserved");
        return SAIGE_E_STATE;
    }
    return engine_add_batch(c, c->eng, k, slots, systems, B_dev, maxiter, tol, 0, true);
}

int saige_engine_add_products_device(saige_ctx* c, int k, const int32_t* slots, const int32_t* systems,
                                     const float* B_dev) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (!c->eng.X) {
        set_error("engine not reserved");
        return SAIGE_E_STATE;
    }
    return engine_add_batch(c, c->eng, k, slots, systems, B_dev, 1, 0.0f, 1, true);
}

int saige_engine_take_batch_ex(saige_ctx* c, int k, const int32_t* slots, float* X_host, float* X_dev,
                               int64_t ldx, const int32_t* xcols, const float* P1_dev, const float* P2_dev,
                               int64_t ldp, const int32_t* pcols, const float* Q_dev, int m, int64_t ldq,
                               double* dots_host) {
    CHECK_CTX(c);
    if (!c->eng.X) {
        set_error("engine not reserved");
        return SAIGE_E_STATE;
    }
    return engine_take_batch_ex(c, c->eng, k, slots, X_host, X_dev, ldx, xcols, P1_dev, P2_dev, ldp, pcols, Q_dev, m,
                                ldq, dots_host);
}

int saige_dots_device(saige_ctx* c, const float* A_dev, int64_t lda, int k, const float* Q_dev, int64_t ldq, int m,
                      double* out_host) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    if (!A_dev || lda < c->Nnm || m < 1) {
        set_error("invalid dot arguments");
        return SAIGE_E_ARG;
    }
    return column_dots(c, A_dev, lda, 0, nullptr, k, nullptr, nullptr, 0, nullptr, Q_dev, m, ldq, out_host);
}

int saige_dots_device_noalloc(saige_ctx* c, const float* A_dev, int64_t lda, int k, const float* Q_dev, int64_t ldq,
                              int m, double* out_host) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    if (!A_dev || lda < c->Nnm || !Q_dev || ldq < c->Nnm || m < 1 || m > 16 || k < 1 || k > EW || !out_host) {
        set_error("invalid dot arguments: %d columns x %d partners (at most %d x 16)", k, m, EW);
        return SAIGE_E_ARG;
    }
    if (c->Nnm > NOALLOC_MAX_N) {
        set_error("saige_dots_device_noalloc: %lld samples exceed its fixed scratch (at most %lld)",
                  (long long)c->Nnm, (long long)NOALLOC_MAX_N);
        return SAIGE_E_ARG;
    }
    if (!c->d_noalloc) {
        double* p = nullptr;                 // the lookup failed at saige_create: its error again
        if (int rc = noalloc_scratch_ptr(c->device, &p)) return rc;
        c->d_noalloc = p;
    }
    hipStream_t st = c->stream;
    // every partial the finishing kernel reads is written first by k_col_dots:
    // nothing to clear
    SAIGE_HIP_TRY(launch_col_dots(st, A_dev, lda, 0, nullptr, k, nullptr, nullptr, 0, nullptr, Q_dev, m, ldq,
                                  c->Nnm, c->d_noalloc, c->d_noalloc + NOALLOC_PARTIAL));
    SAIGE_HIP_TRY(hipMemcpyAsync(out_host, c->d_noalloc + NOALLOC_PARTIAL, sizeof(double) * k * m,
                                 hipMemcpyDeviceToHost, st));
    SAIGE_HIP_TRY(hipStreamSynchronize(st));
    return SAIGE_OK;
}

int saige_probes_device(saige_ctx* c, float* dst_dev, int64_t ld, uint64_t seed, int64_t i0, int k) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    if (!dst_dev || ld < c->Nnm || k < 1 || i0 < 0) {
        set_error("invalid probe request");
        return SAIGE_E_ARG;
    }
    SAIGE_HIP_TRY(launch_rademacher(c->stream, dst_dev, ld, c->Nnm, seed, i0, k));
    SAIGE_HIP_TRY(hipStreamSynchronize(c->stream));
    return SAIGE_OK;
}

int saige_probes_device_rows(saige_ctx* c, float* dst_dev, int64_t ld, uint64_t seed, int64_t i0, int k,
                             const int64_t* rows_dev, int64_t n) {
    CHECK_CTX(c);
    CHECK_LOADED(c);
    if (!dst_dev || ld < c->Nnm || k < 1 || i0 < 0 || !rows_dev || n < 1 || n > c->Nnm) {
        set_error("invalid subset probe request");
        return SAIGE_E_ARG;
    }
    SAIGE_HIP_TRY(launch_fill_zero(c->stream, dst_dev, 4 * ld * k));
    SAIGE_HIP_TRY(launch_rademacher_rows(c->stream, dst_dev, ld, c->Nnm, rows_dev, n, seed, i0, k));
    SAIGE_HIP_TRY(hipStreamSynchronize(c->stream));
    return SAIGE_OK;
}

int saige_copy_host_to_device(saige_ctx* c, void* dst_dev, const void* src_host, int64_t bytes) {
    CHECK_CTX(c);
    if (!dst_dev || !src_host || bytes < 0) {
        set_error("invalid copy");
        return SAIGE_E_ARG;
    }
    SAIGE_HIP_TRY(hipMemcpy(dst_dev, src_host, (size_t)bytes, hipMemcpyHostToDevice));
    return SAIGE_OK;
}

int saige_engine_step(saige_ctx* c, int n_iter, uint8_t* done, int32_t* iters) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (!c->eng.X) {
        set_error("engine not reserved");
        return SAIGE_E_STATE;
    }
    if (int rc = c->eng.loco ? check_loco(c) : restore_full_grm(c)) return rc;
    return engine_step(c, c->eng, n_iter, done, iters);
}

int saige_engine_take(saige_ctx* c, int slot, float* x) {
    CHECK_CTX(c);
    if (!c->eng.X) {
        set_error("engine not reserved");
        return SAIGE_E_STATE;
    }
    return engine_take(c, c->eng, slot, x);
}

int saige_engine_use_loco(saige_ctx* c, int on) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (on && !c->loco_ready) return check_loco(c);
    if (c->eng.X)
        for (int k = 0; k < c->eng.kpad; k++)
            if (c->eng.h_live[k]) {
                set_error("engine has live slots; switch the GRM with an idle engine");
                return SAIGE_E_STATE;
            }
    c->eng.loco = on ? 1 : 0;
    return SAIGE_OK;
}

// ------------------------------------------------------------ autotuner

namespace {
// Block-GEMM shapes tried by saige_autotune, in the order of its table.
// {rt, nsplit, pd}
constexpr GemmConfig kGemmCandidates[] = {
    {1, 1, 1}, {1, 2, 1}, {1, 4, 1}, {1, 8, 1}, {2, 1, 1}, {2, 2, 1}, {2, 4, 1}, {2, 8, 1},
    {1, 1, 3}, {1, 2, 3}, {1, 4, 3}, {1, 8, 3}, {2, 1, 3}, {2, 2, 3}, {2, 4, 3}, {2, 8, 3}};
constexpr int kNumGemmCandidates = (int)(sizeof(kGemmCandidates) / sizeof(kGemmCandidates[0]));
// {rt, -, pd} for one-half launches (nsplit taken from the full shape)
constexpr GemmConfig kHalfCandidates[] = {{1, 0, 1}, {2, 0, 1}, {1, 0, 3}, {2, 0, 3}};
constexpr int kNumHalfCandidates = (int)(sizeof(kHalfCandidates) / sizeof(kHalfCandidates[0]));

int engine_idle(const saige_ctx* c, const char* what) {
    if (c->eng.X)
        for (int k = 0; k < c->eng.kpad; k++)
            if (c->eng.h_live[k]) {
                set_error("engine has live slots; %s needs an idle engine", what);
                return SAIGE_E_STATE;
            }
    return SAIGE_OK;
}

int ensure_mv_scratch(saige_ctx* c) {
    if (c->d_mv_in) return SAIGE_OK;
    const int64_t Np = c->Npad, N = c->Nnm;
    if (dalloc(&c->d_mv_in, N * EW) || dalloc(&c->d_mv_blk, Np * EW) || dalloc(&c->d_mv_out, Np * EW) ||
        dalloc(&c->d_mv_ypart, gemm_ypart_floats(Np))) {
        dfree(c->d_mv_in); dfree(c->d_mv_blk); dfree(c->d_mv_out); dfree(c->d_mv_ypart);
        set_error("device allocation failed (block product scratch)");
        return SAIGE_E_NOMEM;
    }
    return SAIGE_OK;
}
}  // namespace

int saige_autotune(saige_ctx* c, int reps, double* ms_table) {
    CHECK_CTX(c);
    CHECK_GRM(c);
    if (int rc = engine_idle(c, "saige_autotune")) return rc;
    if (int rc = ensure_mv_scratch(c)) return rc;
    reps = std::max(1, std::min(reps, 100));
    hipStream_t st = c->stream;
    const int64_t Np = c->Npad, N = c->Nnm;
    // operands with realistic bit activity: an all-zero block runs the MFMAs
    // at lower power and higher clocks than real data
    SAIGE_HIP_TRY(launch_fill_hash(st, c->d_mv_blk, Np * EW, 0x5A1E6u));
    SigmaArgs a{};
    a.mqc = (float)c->Mqc;
    a.mode = 0;
    // time one shape: one warm-up launch, then reps launches between events
    const bool s16 = grm_view(c, 0).s16;
    const NarrowCols q0 = NarrowCols::tile(16, 0);
    auto time_shape = [&](GemmConfig g, int halves, double* out_ms, bool q = false) -> hipError_t {
        const NarrowCols* quarter = q ? &q0 : nullptr;
        hipError_t e = launch_grm_gemm(st, c->d_grm, Np, N, c->d_mv_blk, c->d_mv_out, a, nullptr, nullptr,
                                       c->d_mv_ypart, g, halves, &c->tri, s16, quarter);
        if (e == hipSuccess) e = hipEventRecord(c->ev0, st);
        for (int r = 0; r < reps && e == hipSuccess; r++)
            e = launch_grm_gemm(st, c->d_grm, Np, N, c->d_mv_blk, c->d_mv_out, a, nullptr, nullptr, c->d_mv_ypart,
                                g, halves, &c->tri, s16, quarter);
        if (e == hipSuccess) e = hipEventRecord(c->ev1, st);
        if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
        *out_ms = (double)ms / reps;
        return e;
    };
    // shapes are tuned within the context's arithmetic mode.  The
    // lower-triangle mode has no column slices (only the nsplit 8 entries are
    // timed; the others read -1) and its summation order depends on rt (the
    // wave count), so one-half launches keep the full shape's rt there.
    const int mode = c->gemm_cfg.mode;
    const bool tri = mode == GEMM_BF16X3_TRI;
    GemmConfig best = c->gemm_cfg;
    double best_ms = -1.0;
    double tfull[kNumGemmCandidates];
    for (int i = 0; i < kNumGemmCandidates; i++) {
        GemmConfig g = kGemmCandidates[i];
        g.mode = mode;
        tfull[i] = -1.0;
        if (ms_table) ms_table[i] = -1.0;
        if (tri && g.nsplit != 8) continue;
        c->gemm_cfg = g;
        if (int rc = ensure_tri(c)) return rc;
        double per = 0.0;
        SAIGE_HIP_TRY(time_shape(g, 3, &per));
        tfull[i] = per;
        if (ms_table) ms_table[i] = per;
        if (best_ms < 0.0 || per < best_ms) {
            best = g;
            best_ms = per;
        }
    }
    c->gemm_cfg = best;
    if (int rc = ensure_tri(c)) return rc;
    // one-half launches: rt and pd only, the slice count stays the full shape's
    GemmConfig best_h = best, best_q = best;
    double best_h_ms = -1.0, best_q_ms = -1.0;
    if (ms_table)
        for (int i = 0; i < kNumHalfCandidates; i++) ms_table[kNumGemmCandidates + kNumHalfCandidates + i] = -1.0;
    if (tri) {
        // rt fixes the summation order, so the 64-, 32- and 16-column
        // launches share it: per rt the fastest pd of each, and the rt of the
        // least full + half/2 + 16-column time (a fit round runs about as
        // many 16-column launches as full ones, fewer one-half ones)
        double score = -1.0;
        for (int rt = 1; rt <= 2; rt++) {
            GemmConfig bf{}, bh{}, bq{};
            double mf = -1.0, mh = -1.0, mq = -1.0;
            for (int i = 0; i < kNumGemmCandidates; i++)
                if (kGemmCandidates[i].rt == rt && tfull[i] >= 0.0) {
                    const double t = tfull[i];
                    if (mf < 0.0 || t < mf) {
                        bf = kGemmCandidates[i];
                        mf = t;
                    }
                }
            c->gemm_cfg = GemmConfig{rt, 8, 1, mode};
            if (int rc = ensure_tri(c)) return rc;
            for (int i = 0; i < kNumHalfCandidates; i++) {
                if (kHalfCandidates[i].rt != rt) continue;
                const GemmConfig g{rt, 8, kHalfCandidates[i].pd, mode};
                double per = 0.0;
                SAIGE_HIP_TRY(time_shape(g, 1, &per));
                if (ms_table) ms_table[kNumGemmCandidates + i] = per;
                if (mh < 0.0 || per < mh) {
                    bh = g;
                    mh = per;
                }
                SAIGE_HIP_TRY(time_shape(g, 1, &per, true));
                if (ms_table) ms_table[kNumGemmCandidates + kNumHalfCandidates + i] = per;
                if (mq < 0.0 || per < mq) {
                    bq = g;
                    mq = per;
                }
            }
            if (mf >= 0.0 && (score < 0.0 || mf + 0.5 * mh + mq < score)) {
                score = mf + 0.5 * mh + mq;
                best = GemmConfig{rt, 8, bf.pd, mode};
                best_ms = mf;
                best_h = bh;
                best_h_ms = mh;
                best_q = bq;
                best_q_ms = mq;
            }
        }
        c->gemm_cfg = best;
        if (int rc = ensure_tri(c)) return rc;
    }
    for (int i = 0; i < kNumHalfCandidates && !tri; i++) {
        GemmConfig g{kHalfCandidates[i].rt, best.nsplit, kHalfCandidates[i].pd, mode};
        if (ms_table) ms_table[kNumGemmCandidates + i] = -1.0;
        double per = 0.0;
        SAIGE_HIP_TRY(time_shape(g, 1, &per));
        if (ms_table) ms_table[kNumGemmCandidates + i] = per;
        if (best_h_ms < 0.0 || per < best_h_ms) {
            best_h = g;
            best_h_ms = per;
        }
    }
    c->gemm_cfg = best;
    c->gemm_tuned_ms = best_ms;
    c->gemm_cfg_half = best_h;
    c->gemm_tuned_ms_half = best_h_ms;
    c->gemm_cfg_quarter = tri ? best_q : best_h;
    c->gemm_tuned_ms_quarter = tri ? best_q_ms : best_h_ms;
    return SAIGE_OK;
}

int saige_autotune_candidates(int* rt, int* pd, int* nsplit) {
    for (int i = 0; i < kNumGemmCandidates; i++) {
        if (rt) rt[i] = kGemmCandidates[i].rt;
        if (pd) pd[i] = kGemmCandidates[i].pd;
        if (nsplit) nsplit[i] = kGemmCandidates[i].nsplit;
    }
    return kNumGemmCandidates;
}

int saige_set_gemm_config(saige_ctx* c, int rt, int pd, int nsplit) {
    CHECK_CTX(c);
    const GemmConfig g{rt, nsplit, pd, c->gemm_cfg.mode};
    if (!gemm_config_valid(g)) {
        set_error("invalid block-GEMM shape rt=%d pd=%d nsplit=%d (rt 1|2, pd 1|3, nsplit 1|2|4|8)", rt, pd,
                  nsplit);
        return SAIGE_E_ARG;
    }
    if (int rc = engine_idle(c, "changing the block-GEMM shape")) return rc;
    c->gemm_cfg = g;
    c->gemm_tuned_ms = 0.0;
    c->gemm_cfg_half = g;
    c->gemm_cfg_quarter = g;
    c->gemm_tuned_ms_half = 0.0;
    c->gemm_tuned_ms_quarter = 0.0;
    return SAIGE_OK;
}

int saige_set_gemm_mode(saige_ctx* c, int mode) {
    CHECK_CTX(c);
    if (mode != SAIGE_GEMM_FP32 && mode != SAIGE_GEMM_BF16X3 && mode != SAIGE_GEMM_BF16X3_TRI) {
        set_error("unknown block-GEMM mode %d (0 = FP32 MFMA, 1 = three-piece BF16 split, 2 = the split over "
                  "the lower triangle)", mode);
        return SAIGE_E_ARG;
    }
    if (int rc = engine_idle(c, "changing the block-GEMM mode")) return rc;
    if (c->gemm_cfg.mode != mode) {
        // timings of the other mode do not carry over
        c->gemm_cfg.mode = mode;
        c->gemm_cfg_half.mode = mode;
        c->gemm_cfg_quarter.mode = mode;
        c->gemm_tuned_ms = 0.0;
        c->gemm_tuned_ms_half = 0.0;
        c->gemm_tuned_ms_quarter = 0.0;
    }
    return SAIGE_OK;
}

int saige_set_gemv_mode(saige_ctx* c, int mode) {
    CHECK_CTX(c);
    if (mode != SAIGE_GEMV_FULL && mode != SAIGE_GEMV_TRI) {
        set_error("unknown GEMV mode %d (0 = all of S, 1 = the lower triangle)", mode);
        return SAIGE_E_ARG;
    }
    if (int rc = engine_idle(c, "changing the GEMV mode")) return rc;
    c->gemv_mode = mode;
    return SAIGE_OK;
}

int saige_get_gemv_mode(const saige_ctx* c, int* mode) {
    if (!c) {
        set_error("null context");
        return SAIGE_E_ARG;
    }
    if (mode) *mode = c->gemv_mode;
    return SAIGE_OK;
}

int saige_get_gemm_mode(const saige_ctx* c, int* mode) {
    if (!c) {
        set_error("null context");
        return SAIGE_E_ARG;
    }
    if (mode) *mode = c->gemm_cfg.mode;
    return SAIGE_OK;
}

int saige_get_gemm_config_half(const saige_ctx* c, int* rt, int* pd, double* tuned_ms) {
    if (!c) {
        set_error("null context");
        return SAIGE_E_ARG;
    }
    if (rt) *rt = c->gemm_cfg_half.rt;
    if (pd) *pd = c->gemm_cfg_half.pd;
    if (tuned_ms) *tuned_ms = c->gemm_tuned_ms_half;
    return SAIGE_OK;
}

int saige_get_gemm_config_quarter(const saige_ctx* c, int* rt, int* pd, double* tuned_ms) {
    if (!c) {
        set_error("null context");
        return SAIGE_E_ARG;
    }
    if (rt) *rt = c->gemm_cfg_quarter.rt;
    if (pd) *pd = c->gemm_cfg_quarter.pd;
    if (tuned_ms) *tuned_ms = c->gemm_tuned_ms_quarter;
    return SAIGE_OK;
}

int saige_get_gemm_config(const saige_ctx* c, int* rt, int* pd, int* nsplit, double* tuned_ms) {
    if (!c) {
        set_error("null context");
        return SAIGE_E_ARG;
    }
    if (rt) *rt = c->gemm_cfg.rt;
    if (pd) *pd = c->gemm_cfg.pd;
    if (nsplit) *nsplit = c->gemm_cfg.nsplit;
    if (tuned_ms) *tuned_ms = c->gemm_tuned_ms;
    return SAIGE_OK;
}

int saige_engine_stats(const saige_ctx* c, int kind, double* secs, int64_t* launches) {
    if (!c) {
        set_error("null context");
        return SAIGE_E_ARG;
    }
    if (kind < 0 || kind > 3) {
        set_error("unknown launch kind %d", kind);
        return SAIGE_E_ARG;
    }
    if (secs) *secs = c->eng.gemm_seconds[kind] + c->eng1.gemm_seconds[kind];
    if (launches) *launches = c->eng.gemm_launches[kind] + c->eng1.gemm_launches[kind];
    return SAIGE_OK;
}

int saige_device_memory(saige_ctx* c, int64_t* free_bytes, int64_t* total_bytes) {
    CHECK_CTX(c);
    size_t f = 0, t = 0;
    SAIGE_HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    return SAIGE_OK;
}

int64_t saige_engine_system_bytes(const saige_ctx* c) {
    // winv + minv (FP32) and the sample mask of one system
    return c ? c->Npad * (2 * 4 + 1) + 8 : 0;
}

int saige_synth_bed_device(saige_ctx* c, void* d_bed, int64_t N, int64_t M, float maf_lo, float maf_hi,
                           float miss, float pair_frac, uint64_t seed) {
    CHECK_CTX(c);
    if (!d_bed || N <= 0 || M <= 0) {
        set_error("invalid synthetic bed arguments");
        return SAIGE_E_ARG;
    }
    const int64_t npairs = (int64_t)std::floor((double)pair_frac * (double)N / 2.0);
    SAIGE_HIP_TRY(launch_synth_bed(c->stream, static_cast<uint8_t*>(d_bed), N, M, maf_lo, maf_hi, miss, npairs,
                                   seed));
    SAIGE_HIP_TRY(hipStreamSynchronize(c->stream));
    return SAIGE_OK;
}

int saige_alloc_device(saige_ctx* c, int64_t bytes, void** out) {
    CHECK_CTX(c);
    if (!out || bytes <= 0) {
        set_error("invalid allocation request");
        return SAIGE_E_ARG;
    }
    SAIGE_HIP_TRY(hipMalloc(out, (size_t)bytes));
    return SAIGE_OK;
}

int saige_free_device(saige_ctx* c, void* p) {
    CHECK_CTX(c);
    if (p) SAIGE_HIP_TRY(hipFree(p));
    return SAIGE_OK;
}

extern "C++" {
namespace saige {
void* ctx_unpark(saige_ctx* c, int64_t bytes) {
    std::lock_guard<std::mutex> lk(c->park_mu);
    for (size_t i = 0; i < c->parked.size(); i++)
        if (c->parked[i].first == bytes) {
            void* p = c->parked[i].second;
            c->parked.erase(c->parked.begin() + (ptrdiff_t)i);
            return p;
        }
    return nullptr;
}

void ctx_park(saige_ctx* c, void* p, int64_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(c->park_mu);
    c->parked.emplace_back(bytes, p);
}
}  // namespace saige
}  // extern "C++"

int saige_alloc_host(saige_ctx* c, int64_t bytes, void** out) {
    CHECK_CTX(c);
    if (!out || bytes <= 0) {
        set_error("invalid allocation request");
        return SAIGE_E_ARG;
    }
    SAIGE_HIP_TRY(hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault));
    return SAIGE_OK;
}

int saige_free_host(saige_ctx* c, void* p) {
    CHECK_CTX(c);
    if (p) SAIGE_HIP_TRY(hipHostFree(p));
    return SAIGE_OK;
}

int saige_copy_device(saige_ctx* c, void* dst, const void* src, int64_t bytes) {
    CHECK_CTX(c);
    if (!src || !dst || bytes < 0) {
        set_error("invalid copy");
        return SAIGE_E_ARG;
    }
    SAIGE_HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, c->stream));
    SAIGE_HIP_TRY(hipStreamSynchronize(c->stream));
    return SAIGE_OK;
}

int saige_copy_device_to_host(saige_ctx* c, const void* src, int64_t bytes, void* dst) {
    CHECK_CTX(c);
    if (!src || !dst || bytes < 0) {
        set_error("invalid copy");
        return SAIGE_E_ARG;
    }
    SAIGE_HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
    return SAIGE_OK;
}

}  // extern "C"
