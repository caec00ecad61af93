// Test harness only: a CPU stand-in for the library's block-PCG slot engine,
// so that the C++ scheduler and request server (saige-gpu_amd/csrc/
// fit_engine.cpp) and the drivers they run (fit_core.cpp) can be exercised,
// also under the host sanitizers, without a GPU.  It implements the C-ABI
// calls fit_engine.cpp makes (include/saige_hip.h) over a dense FP64 K held
// in host memory:
//   - systems (w, tau, sample mask) and 64 slots; a solve slot runs one
//     Jacobi-preconditioned CG iteration per step in FP64 with the
//     reference's stopping rule (r^T r <= tol, at most maxiter iterations,
//     FG.cpp:2696-2792); a product slot is done after one step;
//   - "device" buffers are host memory; probes are saige::fit::probe_vector;
//   - dots are FP64 sums in index order.
// Every slot's arithmetic depends on its own column and system only, as in
// the real engine, so a fit's result is the same alone and in any mix.
// Hooks (extern "C", mock_*) set K, fail a chosen step, and set the free
// memory the admission check sees.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine_park.hpp"
#include "fit_core.hpp"
#include "saige_hip.h"

namespace saige {
static thread_local std::string g_err;
void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}
}  // namespace saige

namespace {
struct System {
    std::vector<double> w;
    double tau[2] = {1.0, 0.0};
    std::vector<uint8_t> mask;
};
struct Slot {
    bool live = false, done = false, product = false;
    int sys = 0, iters = 0, maxiter = 0;
    double tol = 0.0;
    std::vector<double> x, r, z, p, b;
};
}  // namespace

struct saige_ctx {
    int64_t N = 0;
    std::vector<double> K;                 // N x N, column-major
    std::vector<System> sys;
    std::vector<Slot> slot;
    int64_t steps = 0;
    int64_t fail_at_step = -1;             // mock_fail_step: this step returns SAIGE_E_DEVICE
    std::atomic<int64_t> free_bytes{int64_t(1) << 40};
    std::atomic<int64_t> live_allocs{0};  // device buffers neither freed nor parked
    std::vector<std::pair<int64_t, void*>> parked;
};

namespace {
using saige::set_error;

constexpr int kWidth = 64;

void kmul(const saige_ctx* c, const std::vector<double>& v, std::vector<double>& out) {
    const int64_t N = c->N;
    out.assign((size_t)N, 0.0);
    for (int64_t j = 0; j < N; j++) {
        const double vj = v[(size_t)j];
        if (vj == 0.0) continue;
        const double* col = c->K.data() + j * N;
        for (int64_t i = 0; i < N; i++) out[(size_t)i] += col[i] * vj;
    }
}

// Sigma v on the system's samples (zero outside its mask)
void sigma_mul(const saige_ctx* c, const System& s, const std::vector<double>& v, std::vector<double>& out) {
    kmul(c, v, out);
    for (int64_t i = 0; i < c->N; i++)
        out[(size_t)i] = s.mask[(size_t)i] ? s.tau[0] / s.w[(size_t)i] * v[(size_t)i] + s.tau[1] * out[(size_t)i]
                                           : 0.0;
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
}

void precond(const saige_ctx* c, const System& s, const std::vector<double>& r, std::vector<double>& z) {
    z.assign((size_t)c->N, 0.0);
    for (int64_t i = 0; i < c->N; i++)
        if (s.mask[(size_t)i]) z[(size_t)i] = r[(size_t)i] / (s.tau[0] / s.w[(size_t)i] + s.tau[1] * c->K[(size_t)(i * c->N + i)]);
}

int check_slots(const saige_ctx* c, int k, const int32_t* slots, const int32_t* systems) {
    if (c->slot.empty()) {
        set_error("engine not reserved");
        return SAIGE_E_STATE;
    }
    for (int j = 0; j < k; j++) {
        if (slots[j] < 0 || slots[j] >= kWidth || c->slot[(size_t)slots[j]].live) {
            set_error("slot %d out of range or busy", slots[j]);
            return SAIGE_E_ARG;
        }
        if (systems && (systems[j] < 0 || systems[j] >= (int)c->sys.size())) {
            set_error("system %d out of range", systems[j]);
            return SAIGE_E_ARG;
        }
    }
    return SAIGE_OK;
}

void start(saige_ctx* c, int slot, int s, const float* b, bool product, int maxiter, float tol) {
    Slot& t = c->slot[(size_t)slot];
    const System& sy = c->sys[(size_t)s];
    t = Slot{};
    t.live = true;
    t.product = product;
    t.sys = s;
    t.maxiter = maxiter;
    t.tol = tol;
    t.b.assign((size_t)c->N, 0.0);
    for (int64_t i = 0; i < c->N; i++) t.b[(size_t)i] = sy.mask[(size_t)i] ? (double)b[i] : 0.0;
    if (product) return;
    t.x.assign((size_t)c->N, 0.0);
    t.r = t.b;
    precond(c, sy, t.r, t.z);
    t.p = t.z;
    if (dot(t.r, t.r) <= t.tol) t.done = true;
}

// a finished slot's result: the solution, or K b for a product slot
void take_col(const saige_ctx* c, const Slot& t, double* out) {
    for (int64_t i = 0; i < c->N; i++) out[i] = t.x[(size_t)i];
}
}  // namespace

extern "C" {

// ------------------------------------------------------------------ hooks
saige_ctx* mock_create(const double* K, int64_t N) {
    auto* c = new saige_ctx();
    c->N = N;
    c->K.assign(K, K + N * N);
    return c;
}
void mock_destroy(saige_ctx* c) {
    for (auto& b : c->parked) std::free(b.second);
    delete c;
}
void mock_fail_step(saige_ctx* c, int64_t step) { c->fail_at_step = step; }
void mock_set_free_bytes(saige_ctx* c, int64_t b) { c->free_bytes = b; }
int64_t mock_steps(const saige_ctx* c) { return c->steps; }
int64_t mock_live_allocs(const saige_ctx* c) { return c->live_allocs; }
int64_t mock_parked(const saige_ctx* c) { return (int64_t)c->parked.size(); }

// ------------------------------------------------------- the C-ABI subset
const char* saige_last_error(void) { return saige::g_err.c_str(); }
int64_t saige_get_nnomissing(const saige_ctx* c) { return c ? c->N : 0; }

int saige_device_memory(saige_ctx* c, int64_t* free_b, int64_t* total_b) {
    if (!c) return SAIGE_E_ARG;
    if (free_b) *free_b = c->free_bytes;
    if (total_b) *total_b = int64_t(1) << 40;
    return SAIGE_OK;
}

int saige_alloc_device(saige_ctx* c, int64_t bytes, void** out) {
    if (!c || !out || bytes <= 0) return SAIGE_E_ARG;
    *out = std::calloc((size_t)bytes, 1);
    if (!*out) return SAIGE_E_NOMEM;
    c->live_allocs++;
    return SAIGE_OK;
}
int saige_free_device(saige_ctx* c, void* p) {
    if (!c) return SAIGE_E_ARG;
    if (p) {
        std::free(p);
        c->live_allocs--;
    }
    return SAIGE_OK;
}
// parked blocks: kept on the context, not counted as live
extern "C++" {
namespace saige {
void* ctx_unpark(saige_ctx* c, int64_t bytes) {
    for (size_t i = 0; i < c->parked.size(); i++)
        if (c->parked[i].first == bytes) {
            void* p = c->parked[i].second;
            c->parked.erase(c->parked.begin() + (ptrdiff_t)i);
            c->live_allocs++;
            return p;
        }
    return nullptr;
}
void ctx_park(saige_ctx* c, void* p, int64_t bytes) {
    if (!p) return;
    c->parked.emplace_back(bytes, p);
    c->live_allocs--;
}
}  // namespace saige
}  // extern "C++"

// host blocks: plain memory, not counted among the device buffers (the
// scheduler keeps its blocks for the process)
int saige_alloc_host(saige_ctx* c, int64_t bytes, void** out) {
    if (!c || !out || bytes <= 0) return SAIGE_E_ARG;
    *out = std::calloc((size_t)bytes, 1);
    return *out ? SAIGE_OK : SAIGE_E_NOMEM;
}
int saige_free_host(saige_ctx* c, void* p) {
    if (!c) return SAIGE_E_ARG;
    std::free(p);
    return SAIGE_OK;
}
int saige_copy_host_to_device(saige_ctx* c, void* dst, const void* src, int64_t bytes) {
    if (!c || !dst || !src || bytes < 0) return SAIGE_E_ARG;
    std::memcpy(dst, src, (size_t)bytes);
    return SAIGE_OK;
}

int saige_engine_reserve(saige_ctx* c, int width, int n_systems) {
    if (!c || width < 1 || width > kWidth || n_systems < 1) return SAIGE_E_ARG;
    for (const Slot& t : c->slot)
        if (t.live) {
            set_error("engine has live slots");
            return SAIGE_E_STATE;
        }
    c->slot.assign(kWidth, Slot{});
    c->sys.assign((size_t)n_systems, System{});
    for (System& s : c->sys) {
        s.w.assign((size_t)c->N, 1.0);
        s.mask.assign((size_t)c->N, 1);
    }
    return SAIGE_OK;
}

int saige_engine_set_system(saige_ctx* c, int s, const float* w, const float* tau, const uint8_t* mask) {
    if (!c || s < 0 || s >= (int)c->sys.size() || !w || !tau) return SAIGE_E_ARG;
    System& sy = c->sys[(size_t)s];
    for (const Slot& t : c->slot)
        if (t.live && t.sys == s) {
            set_error("system %d has live slots", s);
            return SAIGE_E_STATE;
        }
    for (int64_t i = 0; i < c->N; i++) {
        sy.w[(size_t)i] = w[i];
        sy.mask[(size_t)i] = mask ? mask[i] : 1;
    }
    sy.tau[0] = tau[0];
    sy.tau[1] = tau[1];
    return SAIGE_OK;
}

int saige_engine_add_batch(saige_ctx* c, int k, const int32_t* slots, const int32_t* systems, const float* B,
                           int maxiter, float tol) {
    if (!c || k < 1 || !slots || !systems || !B) return SAIGE_E_ARG;
    if (int rc = check_slots(c, k, slots, systems)) return rc;
    for (int j = 0; j < k; j++) start(c, slots[j], systems[j], B + (int64_t)j * c->N, false, maxiter, tol);
    return SAIGE_OK;
}
int saige_engine_add_products(saige_ctx* c, int k, const int32_t* slots, const int32_t* systems, const float* B) {
    if (!c || k < 1 || !slots || !systems || !B) return SAIGE_E_ARG;
    if (int rc = check_slots(c, k, slots, systems)) return rc;
    for (int j = 0; j < k; j++) start(c, slots[j], systems[j], B + (int64_t)j * c->N, true, 0, 0.f);
    return SAIGE_OK;
}
int saige_engine_add_batch_device(saige_ctx* c, int k, const int32_t* slots, const int32_t* systems,
                                  const float* B, int maxiter, float tol) {
    return saige_engine_add_batch(c, k, slots, systems, B, maxiter, tol);
}
int saige_engine_add_products_device(saige_ctx* c, int k, const int32_t* slots, const int32_t* systems,
                                     const float* B) {
    return saige_engine_add_products(c, k, slots, systems, B);
}

int saige_engine_step(saige_ctx* c, int n_iter, uint8_t* done, int32_t* iters) {
    if (!c || n_iter < 1) return SAIGE_E_ARG;
    if (c->slot.empty()) {
        set_error("engine not reserved");
        return SAIGE_E_STATE;
    }
    if (c->fail_at_step >= 0 && c->steps == c->fail_at_step) {
        c->steps++;
        set_error("mock engine: injected device failure at step %lld", (long long)c->fail_at_step);
        return SAIGE_E_DEVICE;
    }
    c->steps++;
    std::vector<double> Ap;
    for (int it = 0; it < n_iter; it++) {
        for (Slot& t : c->slot) {
            if (!t.live || t.done) continue;
            const System& sy = c->sys[(size_t)t.sys];
            if (t.product) {
                kmul(c, t.b, t.x);
                for (int64_t i = 0; i < c->N; i++)
                    if (!sy.mask[(size_t)i]) t.x[(size_t)i] = 0.0;
                t.done = true;
                continue;
            }
            sigma_mul(c, sy, t.p, Ap);
            const double rz = dot(t.r, t.z);
            const double alpha = rz / dot(t.p, Ap);
            for (int64_t i = 0; i < c->N; i++) {
                t.x[(size_t)i] += alpha * t.p[(size_t)i];
                t.r[(size_t)i] -= alpha * Ap[(size_t)i];
            }
            t.iters++;
            if (dot(t.r, t.r) <= t.tol || t.iters >= t.maxiter) {
                t.done = true;
                continue;
            }
            precond(c, sy, t.r, t.z);
            const double beta = dot(t.r, t.z) / rz;
            for (int64_t i = 0; i < c->N; i++) t.p[(size_t)i] = t.z[(size_t)i] + beta * t.p[(size_t)i];
        }
    }
    for (int s = 0; s < kWidth; s++) {
        const Slot& t = c->slot[(size_t)s];
        if (done) done[s] = t.live ? (uint8_t)t.done : 1;
        if (iters) iters[s] = t.live ? t.iters : 0;
    }
    return SAIGE_OK;
}

int saige_engine_take_batch(saige_ctx* c, int k, const int32_t* slots, float* X) {
    if (!c || k < 1 || !slots) return SAIGE_E_ARG;
    std::vector<double> col((size_t)c->N);
    for (int j = 0; j < k; j++) {
        if (slots[j] < 0 || slots[j] >= kWidth) return SAIGE_E_ARG;
        Slot& t = c->slot[(size_t)slots[j]];
        if (X) {
            take_col(c, t, col.data());
            for (int64_t i = 0; i < c->N; i++) X[(int64_t)j * c->N + i] = (float)col[(size_t)i];
        }
        t.live = false;
    }
    return SAIGE_OK;
}

int saige_engine_take_batch_ex(saige_ctx* c, int k, const int32_t* slots, float* X_host, float* X_dev, int64_t ldx,
                               const int32_t* xcols, const float* P1, const float* P2, int64_t ldp,
                               const int32_t* pcols, const float* Q, int m, int64_t ldq, double* dots) {
    if (!c || k < 1 || !slots) return SAIGE_E_ARG;
    const int nd = (P1 ? 1 : 0) + (P2 ? 1 : 0) + (Q ? m : 0);
    if (nd > 16 || (nd > 0 && !dots)) return SAIGE_E_ARG;
    std::vector<double> col((size_t)c->N);
    for (int j = 0; j < k; j++) {
        if (slots[j] < 0 || slots[j] >= kWidth) return SAIGE_E_ARG;
        Slot& t = c->slot[(size_t)slots[j]];
        take_col(c, t, col.data());
        // the result as the device holds it: FP32
        for (double& v : col) v = (double)(float)v;
        if (X_host)
            for (int64_t i = 0; i < c->N; i++) X_host[(int64_t)j * c->N + i] = (float)col[(size_t)i];
        if (X_dev) {
            const int64_t xc = xcols ? xcols[j] : j;
            for (int64_t i = 0; i < c->N; i++) X_dev[xc * ldx + i] = (float)col[(size_t)i];
        }
        int d = 0;
        const int64_t pc = pcols ? pcols[j] : j;
        for (const float* P : {P1, P2}) {
            if (!P) continue;
            double s = 0.0;
            for (int64_t i = 0; i < c->N; i++) s += col[(size_t)i] * (double)P[pc * ldp + i];
            dots[j * nd + d++] = s;
        }
        for (int q = 0; Q && q < m; q++) {
            double s = 0.0;
            for (int64_t i = 0; i < c->N; i++) s += col[(size_t)i] * (double)Q[q * ldq + i];
            dots[j * nd + d++] = s;
        }
        t.live = false;
    }
    return SAIGE_OK;
}

int saige_dots_device(saige_ctx* c, const float* A, int64_t lda, int k, const float* Q, int64_t ldq, int m,
                      double* out) {
    if (!c || !A || !Q || !out || k < 1 || m < 1) return SAIGE_E_ARG;
    for (int j = 0; j < k; j++)
        for (int q = 0; q < m; q++) {
            double s = 0.0;
            for (int64_t i = 0; i < c->N; i++) s += (double)A[j * lda + i] * (double)Q[q * ldq + i];
            out[j * m + q] = s;
        }
    return SAIGE_OK;
}

int saige_probes_device_rows(saige_ctx* c, float* dst, int64_t ld, uint64_t seed, int64_t i0, int k,
                             const int64_t* rows, int64_t n) {
    if (!c || !dst || k < 1 || !rows || n < 1 || n > c->N) return SAIGE_E_ARG;
    std::vector<double> v((size_t)n);
    for (int j = 0; j < k; j++) {
        std::fill(dst + j * ld, dst + j * ld + c->N, 0.0f);
        saige::fit::probe_vector(seed, i0 + j, n, v.data());
        for (int64_t i = 0; i < n; i++)
            if (rows[i] >= 0 && rows[i] < c->N) dst[j * ld + rows[i]] = (float)v[(size_t)i];
    }
    return SAIGE_OK;
}

int saige_probes_device(saige_ctx* c, float* dst, int64_t ld, uint64_t seed, int64_t i0, int k) {
    if (!c || !dst || k < 1) return SAIGE_E_ARG;
    std::vector<double> v((size_t)c->N);
    for (int j = 0; j < k; j++) {
        saige::fit::probe_vector(seed, i0 + j, c->N, v.data());
        for (int64_t i = 0; i < c->N; i++) dst[j * ld + i] = (float)v[(size_t)i];
    }
    return SAIGE_OK;
}

}  // extern "C"
