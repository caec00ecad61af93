// Batched null fits in C++: fit_core's drivers run one thread per fit, and
// their solves and GRM products are packed into the block-PCG slot engine by
// a scheduler on the calling thread (continuous batching: a request's
// columns take free slots as they appear, join the next iteration and leave
// when they converge; the fit resumes as soon as its request is done, while
// the engine keeps stepping the other fits' columns).  Only the scheduler
// thread calls into the library's engine; fit threads do host arithmetic.
//
// No reference counterpart: the reference fits one phenotype per R process
// through fitNULLGLMM (FG.R:1036-1298), one right-hand side at a time.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/saige_hip.h"
#include "engine_park.hpp"
#include "fit_core.hpp"

namespace saige {
void set_error(const char* fmt, ...);   // the library's last-error message (saige_capi.hip)
}

namespace {

using saige::fit::Matrix;

struct Scheduler;

// A trace round run on the device (Solver::trace_round): the probes U, K U
// and Sigma^-1 X stay in the fit's system buffers; phase 1 takes the
// products K [PY, U] (K u into AU with h = (Sigma^-1 X)^T K u), phase 2 the
// solves Sigma^-1 [U, K PY, (PY)] (the U solves only as the dots d and e).
struct TraceState {
    const Matrix* SX = nullptr;
    const std::vector<double>* PY = nullptr;
    bool quantitative = false;
    int64_t start = 0, k = 0, p = 0;
    uint64_t seed = 0;
    saige::fit::TraceTerms* out = nullptr;
    int phase = 1;
    bool setup = false;           // probes, Sigma^-1 X and g are on the device
    Matrix host1, res1;           // phase 1 host columns (PY) and their products
    Matrix host2, res2;           // phase 2 host columns (K PY, PY) and their solves
};

struct Request {
    bool solve = true;            // else a GRM product
    const Matrix* B = nullptr;    // n x k over the fit's samples
    Matrix* out = nullptr;
    std::vector<double> w;        // solves: the system's weights (n)
    double tau[2] = {0.0, 0.0};
    int64_t remaining = 0;        // columns not yet done
    bool done = false;
    std::string error;
    // product_then_solve: when the products are done, the scheduler turns the
    // request into the solve of C = gather_columns(B, KB, src) into chain_out
    bool chain = false;
    std::vector<int> src;
    Matrix C;
    Matrix* chain_out = nullptr;
    TraceState* trace = nullptr;  // a device trace round
};

struct FitTask : saige::fit::Solver {
    Scheduler* sch = nullptr;
    const saige_fit_spec* spec = nullptr;
    int sys = -1;                 // engine system while admitted
    bool sys_set = false;         // the system holds `set_w`/`set_tau`
    std::vector<double> set_w;
    double set_tau[2] = {0.0, 0.0};
    std::thread th;
    saige::fit::Result result;
    std::string error;
    bool finished = false;
    bool released = false;        // its system went back to the pool (the thread has ended)
    // serving: the request's own copy of its inputs, and when it was
    // submitted, admitted and finished (steady clock)
    saige_fit_spec own_spec{};
    std::vector<double> own_y, own_X, own_offset;
    std::vector<int64_t> own_subset;
    std::chrono::steady_clock::time_point t_submit, t_admit, t_done;
    int64_t posts = 0;            // requests posted so far
    std::condition_variable cv;   // the fit waits here for its request
    std::chrono::steady_clock::time_point t_resumed{};   // its thread last woke (diagnostics)

    void solve(const std::vector<double>& w, const double tau[2], const Matrix& B, Matrix& X) override;
    void product(const Matrix& B, Matrix& KB) override;
    void product_then_solve(const Matrix& B, Matrix& KB, const std::vector<double>& w, const double tau[2],
                            const std::vector<int>& src, Matrix& X) override;
    void trace_round(const std::vector<double>& w, const double tau[2], const Matrix& Sigma_iX,
                     const std::vector<double>* PY, bool quantitative, int64_t start, int64_t end, uint64_t seed,
                     saige::fit::TraceTerms& out) override;
    void submit(Request& r);
};

// What becomes of a column when its slot is done.
enum Action : int8_t {
    kHost = 0,    // download into req->out column j
    kStash = 1,   // trace phase 1: K u_j into AU column j, h_j dots
    kPair = 2,    // trace phase 2: Sigma^-1 u_j as the dots d_j (and e_j)
};

struct Column {
    FitTask* fit;
    Request* req;
    int64_t j;
    int dev = -1;                 // source: column `dev` of the system's U buffer (else req->B column j)
    Action act = kHost;
};

// Device buffers of a system's trace rounds, column-major with leading
// dimension N: probes U and K U (kTraceProbes columns each), Sigma^-1 X
// (kTraceCovariates).
constexpr int kTraceProbes = 64;
constexpr int kTraceCovariates = 16;   // the shared dot partners of one take
struct TraceBuffers {
    float* U = nullptr;
    float* AU = nullptr;
    float* SX = nullptr;
    int64_t* rows = nullptr;        // the sample subset of rows_fit (N entries at most)
    const FitTask* rows_fit = nullptr;
};

struct Scheduler {
    saige_ctx* ctx;
    int64_t N;
    saige::fit::Options opt;
    int width = 64;
    bool device_trace = true;             // trace rounds on the device (else the host round)
    std::mutex mu;
    std::condition_variable cv;           // the scheduler waits here for requests / finished fits
    std::deque<std::pair<FitTask*, Request*>> inbox;
    std::string fatal;                    // engine failure: every fit stops
    std::deque<FitTask*> queue;           // submitted, not yet admitted
    bool stopping = false;                // no more submissions: run() returns when all are done
    int64_t hbm_reserve = 0;              // admit only while this much HBM is free
    std::condition_variable done_cv;      // a fit finished (saige_poll waits here)
    double t_fit_host = 0;                // fits' own host seconds between requests (diagnostics)

    void post(FitTask* f, Request* r) {
        using clk = std::chrono::steady_clock;
        const auto now = clk::now();
        std::unique_lock<std::mutex> lk(mu);
        if (f->t_resumed.time_since_epoch().count() != 0)
            t_fit_host += std::chrono::duration<double>(now - f->t_resumed).count();
        if (!fatal.empty()) {
            r->error = fatal;
            r->done = true;
            return;
        }
        inbox.emplace_back(f, r);
        f->posts++;
        cv.notify_all();
        f->cv.wait(lk, [&] { return r->done; });
        f->t_resumed = clk::now();
    }
};

void FitTask::submit(Request& r) {
    sch->post(this, &r);
    if (!r.error.empty()) throw std::runtime_error(r.error);
}

void FitTask::solve(const std::vector<double>& w, const double tau[2], const Matrix& B, Matrix& X) {
    Request r;
    r.solve = true;
    r.B = &B;
    r.out = &X;
    r.w = w;
    r.tau[0] = tau[0];
    r.tau[1] = tau[1];
    X = Matrix(B.rows, B.cols);
    submit(r);
}

void FitTask::product(const Matrix& B, Matrix& KB) {
    Request r;
    r.solve = false;
    r.B = &B;
    r.out = &KB;
    KB = Matrix(B.rows, B.cols);
    submit(r);
}

void FitTask::product_then_solve(const Matrix& B, Matrix& KB, const std::vector<double>& w, const double tau[2],
                                 const std::vector<int>& src, Matrix& X) {
    Request r;
    r.solve = false;
    r.B = &B;
    r.out = &KB;
    r.w = w;
    r.tau[0] = tau[0];
    r.tau[1] = tau[1];
    r.chain = true;
    r.src = src;
    r.chain_out = &X;
    KB = Matrix(B.rows, B.cols);
    X = Matrix(B.rows, (int64_t)src.size());
    submit(r);
}

void FitTask::trace_round(const std::vector<double>& w, const double tau[2], const Matrix& Sigma_iX,
                          const std::vector<double>* PY, bool quantitative, int64_t start, int64_t end,
                          uint64_t seed, saige::fit::TraceTerms& out) {
    const int64_t k = end - start;
    if (!sch->device_trace || k > kTraceProbes || Sigma_iX.cols > kTraceCovariates) {
        saige::fit::Solver::trace_round(w, tau, Sigma_iX, PY, quantitative, start, end, seed, out);
        return;
    }
    TraceState t;
    t.SX = &Sigma_iX;
    t.PY = PY;
    t.quantitative = quantitative;
    t.start = start;
    t.k = k;
    t.p = Sigma_iX.cols;
    t.seed = seed;
    t.out = &out;
    out.d.assign((size_t)k, 0.0);
    out.e.assign((size_t)k, 0.0);
    out.h = Matrix(k, t.p);
    out.g = Matrix(k, t.p);
    if (PY) {
        t.host1 = Matrix(Sigma_iX.rows, 1);
        std::copy(PY->begin(), PY->end(), t.host1.col(0));
        t.res1 = Matrix(Sigma_iX.rows, 1);
    }
    Request r;
    r.solve = false;
    r.B = &t.host1;
    r.out = &t.res1;
    r.w = w;
    r.tau[0] = tau[0];
    r.tau[1] = tau[1];
    r.trace = &t;
    submit(r);
}

int check(int rc) {
    if (rc != SAIGE_OK) throw std::runtime_error(saige_last_error());
    return rc;
}

// the fit's samples in the N phenotyped samples
int64_t sample_of(const saige_fit_spec* s, int64_t i) { return s->subset ? s->subset[i] : i; }

// Page-locked host blocks from the library, kept for the process once made
// (pinning tens of MB takes milliseconds, about what a fit round's transfers
// gain from it): a scheduler takes the blocks it needs and gives them back
// when it ends.
class HostBlocks {
public:
    static HostBlocks& get() {
        static HostBlocks* b = new HostBlocks();   // never destroyed, like the blocks
        return *b;
    }
    // a block of at least `bytes`, or nullptr if none can be made
    void* take(saige_ctx* ctx, int64_t bytes) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (size_t i = 0; i < free_.size(); i++)
                if (free_[i].first >= bytes) {
                    void* p = free_[i].second;
                    used_.emplace_back(free_[i]);
                    free_.erase(free_.begin() + (ptrdiff_t)i);
                    return p;
                }
        }
        void* p = nullptr;
        if (saige_alloc_host(ctx, bytes, &p) != SAIGE_OK) return nullptr;
        std::lock_guard<std::mutex> lk(mu_);
        used_.emplace_back(bytes, p);
        return p;
    }
    void give(void* p) {
        std::lock_guard<std::mutex> lk(mu_);
        for (size_t i = 0; i < used_.size(); i++)
            if (used_[i].second == p) {
                free_.push_back(used_[i]);
                used_.erase(used_.begin() + (ptrdiff_t)i);
                return;
            }
    }

private:
    std::mutex mu_;
    std::vector<std::pair<int64_t, void*>> free_, used_;
};

// n values in a page-locked block (zeroed plain memory if none can be made):
// host columns staged here move to and from the device at the link's rate
// instead of through the runtime's bounce buffers
template <class T>
class HostArray {
public:
    HostArray(saige_ctx* ctx, int64_t n) {
        const int64_t m = std::max<int64_t>(n, 1);
        p_ = static_cast<T*>(HostBlocks::get().take(ctx, (int64_t)sizeof(T) * m));
        pinned_ = p_ != nullptr;
        if (!pinned_) p_ = new T[(size_t)m]();
    }
    ~HostArray() {
        if (pinned_)
            HostBlocks::get().give(p_);
        else
            delete[] p_;
    }
    HostArray(const HostArray&) = delete;
    HostArray& operator=(const HostArray&) = delete;
    T* data() { return p_; }
    T& operator[](size_t i) { return p_[i]; }

private:
    T* p_ = nullptr;
    bool pinned_ = false;
};

// The scheduler loop: admits the queued fits while an engine system and
// HBM are free, steps the engine over their columns, and returns once
// `stopping` is set and every submitted fit has finished.
void run_scheduler(Scheduler& sch, int n_systems) {
    const int64_t N = sch.N;
    check(saige_engine_reserve(sch.ctx, sch.width, n_systems));
    std::deque<int> free_sys;
    for (int s = 0; s < n_systems; s++) free_sys.push_back(s);
    std::vector<int> free_slots;                      // lowest slot first (min-heap)
    for (int s = sch.width - 1; s >= 0; s--) free_slots.push_back(s);
    std::vector<FitTask*> active;                     // admitted, system not yet released
    std::deque<Column> pending;
    std::vector<Column> owner((size_t)sch.width, Column{nullptr, nullptr, 0});
    std::vector<uint8_t> done((size_t)sch.width);
    std::vector<int32_t> iters((size_t)sch.width);
    size_t live = 0;
    // staging for host columns (at most one per slot, or the trace round's
    // covariate columns), a system's weights and its sample mask
    HostArray<float> colbuf(sch.ctx, N * std::max<int64_t>(sch.width, kTraceCovariates));
    HostArray<float> wfull(sch.ctx, N);
    HostArray<uint8_t> mask(sch.ctx, N);
    // the trace rounds' device buffers, per system, taken on first use from
    // those the context keeps (or made), and handed back to it when the batch
    // ends
    const int64_t probe_bytes = 4 * N * kTraceProbes, cov_bytes = 4 * N * kTraceCovariates;
    struct Buffers {
        saige_ctx* ctx;
        std::vector<TraceBuffers> b;
        int64_t probe_bytes, cov_bytes, rows_bytes;
        ~Buffers() {
            for (auto& t : b) {
                saige::ctx_park(ctx, t.U, probe_bytes);
                saige::ctx_park(ctx, t.AU, probe_bytes);
                saige::ctx_park(ctx, t.SX, cov_bytes);
                saige::ctx_park(ctx, t.rows, rows_bytes);
            }
        }
    } tbufs{sch.ctx, std::vector<TraceBuffers>((size_t)n_systems), probe_bytes, cov_bytes, 8 * N};
    auto device_block = [&](int64_t bytes) {
        void* p = saige::ctx_unpark(sch.ctx, bytes);
        if (!p) check(saige_alloc_device(sch.ctx, bytes, &p));
        return p;
    };
    // fits whose request the last step completed, with their post count then
    std::vector<std::pair<FitTask*, int64_t>> resumed;
    // wall seconds of the loop's parts, printed to stderr at the end when the
    // environment sets SAIGE_ENGINE_TIMING (diagnostics only)
    const bool timing = std::getenv("SAIGE_ENGINE_TIMING") != nullptr;
    double t_idle = 0, t_fill = 0, t_step = 0, t_wait_posts = 0, t_take = 0, t_setup = 0;
    double t_sys = 0, t_upload = 0, t_devadd = 0, t_take_dots = 0, t_take_host = 0;
    int64_t n_steps = 0;
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };

    auto set_system = [&](FitTask* f, const double* w, const double tau[2]) {
        const auto t0 = clk::now();
        const saige_fit_spec* s = f->spec;
        std::fill(mask.data(), mask.data() + N, (uint8_t)(s->subset ? 0 : 1));
        std::fill(wfull.data(), wfull.data() + N, 1.0f);
        for (int64_t i = 0; i < s->n; i++) {
            const int64_t r = sample_of(s, i);
            mask[(size_t)r] = 1;
            wfull[(size_t)r] = w ? (float)w[i] : 1.0f;
        }
        const float t[2] = {(float)tau[0], (float)tau[1]};
        check(saige_engine_set_system(sch.ctx, f->sys, wfull.data(), t, mask.data()));
        t_sys += secs(t0);
    };

    // a trace round's device side before its first column runs: the probes
    // into U (generated in place, or uploaded for a sample subset), Sigma^-1 X
    // into SX, and g = (Sigma^-1 X)^T u
    auto trace_setup = [&](FitTask* f, TraceState& t) {
        const auto t0 = clk::now();
        TraceBuffers& b = tbufs.b[(size_t)f->sys];
        if (!b.U) {
            b.U = (float*)device_block(probe_bytes);
            b.AU = (float*)device_block(probe_bytes);
            b.SX = (float*)device_block(cov_bytes);
        }
        const saige_fit_spec* sp = f->spec;
        const int64_t n = sp->n, k = t.k, p = t.p;
        if (!sp->subset) {
            check(saige_probes_device(sch.ctx, b.U, N, t.seed, t.start, (int)k));
        } else {
            // the subset's rows go to the device once per fit on this system;
            // its probes are then made there like the full cohort's
            if (b.rows_fit != f) {
                if (!b.rows) b.rows = (int64_t*)device_block(8 * N);
                check(saige_copy_host_to_device(sch.ctx, b.rows, sp->subset, 8 * n));
                b.rows_fit = f;
            }
            check(saige_probes_device_rows(sch.ctx, b.U, N, t.seed, t.start, (int)k, b.rows, n));
        }
        for (int64_t c = 0; c < p; c++) {
            const double* x = t.SX->col(c);
            float* d = colbuf.data() + c * N;
            if (sp->subset) std::fill(d, d + N, 0.0f);
            for (int64_t i = 0; i < n; i++) d[sample_of(sp, i)] = (float)x[i];
        }
        check(saige_copy_host_to_device(sch.ctx, b.SX, colbuf.data(), 4 * N * p));
        std::vector<double> g((size_t)(k * p));
        check(saige_dots_device(sch.ctx, b.U, N, (int)k, b.SX, N, (int)p, g.data()));
        for (int64_t j = 0; j < k; j++)
            for (int64_t c = 0; c < p; c++) t.out->g(j, c) = g[(size_t)(j * p + c)];
        t.setup = true;
        t_setup += secs(t0);
    };

    // admission by free HBM: a fit's trace buffers and host-visible state
    // come out of what is free on the device
    auto hbm_ok = [&] {
        if (sch.hbm_reserve <= 0) return true;
        int64_t free_b = 0, total_b = 0;
        return saige_device_memory(sch.ctx, &free_b, &total_b) == SAIGE_OK && free_b >= sch.hbm_reserve;
    };

    try {
    for (;;) {
        // admit fits while a system (and HBM) is free
        std::vector<FitTask*> admit;
        {
            std::lock_guard<std::mutex> lk(sch.mu);
            if (sch.stopping && sch.queue.empty() && active.empty()) break;
            while (!sch.queue.empty() && free_sys.size() > admit.size()) {
                admit.push_back(sch.queue.front());
                sch.queue.pop_front();
            }
        }
        bool hbm_short = false;
        if (!admit.empty() && !hbm_ok()) {
            std::lock_guard<std::mutex> lk(sch.mu);   // back to the head of the queue, in order
            for (auto it = admit.rbegin(); it != admit.rend(); ++it) sch.queue.push_front(*it);
            admit.clear();
            hbm_short = true;
        }
        for (FitTask* f : admit) {
            f->sys = free_sys.front();
            free_sys.pop_front();
            f->sys_set = false;
            f->t_admit = clk::now();
            active.push_back(f);
            resumed.emplace_back(f, 0);               // its first request is awaited like a resumed fit's
            f->th = std::thread([f, &sch] {
                f->t_resumed = std::chrono::steady_clock::now();
                try {
                    const saige_fit_spec* s = f->spec;
                    Matrix X(s->n, s->p);
                    std::memcpy(X.a.data(), s->X, sizeof(double) * (size_t)(s->n * s->p));
                    std::vector<double> y(s->y, s->y + s->n);
                    std::vector<double> off;
                    if (s->offset) off.assign(s->offset, s->offset + s->n);
                    f->result = saige::fit::fit_null(
                        *f, s->trait == SAIGE_TRAIT_QUANTITATIVE ? saige::fit::Trait::Quantitative
                                                                 : saige::fit::Trait::Binary,
                        y, X, off, sch.opt);
                } catch (const std::exception& e) {
                    f->error = e.what();
                }
                std::lock_guard<std::mutex> lk(sch.mu);
                f->t_done = clk::now();
                f->finished = true;
                sch.cv.notify_all();
                sch.done_cv.notify_all();
            });
        }
        // collect new requests; release finished fits' systems
        {
            std::unique_lock<std::mutex> lk(sch.mu);
            const auto t0 = clk::now();
            if (live == 0 && pending.empty()) {
                // nothing runs: wait until every fit the last step resumed (or
                // that was just admitted) has posted its next request or
                // finished, so that the fits' next phases start in the same
                // iteration and share its launches (fits that drift apart run
                // more of them); with none such, for any request, finished fit
                // or arrival that can be admitted (with HBM short, arrivals
                // are looked at again every 10 ms)
                auto ready = [&] {
                    if (!resumed.empty()) {
                        for (auto& fp : resumed)
                            if (!fp.first->finished && fp.first->posts <= fp.second) return false;
                        return true;
                    }
                    if (!sch.inbox.empty()) return true;
                    for (FitTask* f : active)
                        if (f->finished) return true;
                    if (!hbm_short && !sch.queue.empty() && !free_sys.empty()) return true;   // an arrival
                    return sch.stopping && sch.queue.empty() && active.empty();
                };
                if (hbm_short)
                    sch.cv.wait_for(lk, std::chrono::milliseconds(10), ready);
                else
                    sch.cv.wait(lk, ready);
                resumed.clear();
            }
            t_idle += secs(t0);
            while (!sch.inbox.empty()) {
                auto [f, r] = sch.inbox.front();
                sch.inbox.pop_front();
                r->remaining = r->B->cols;
                for (int64_t j = 0; j < r->B->cols; j++) pending.push_back(Column{f, r, j});
                if (r->trace) {                        // phase 1: K [PY, U]; K u stays on the device
                    r->remaining += r->trace->k;
                    for (int64_t j = 0; j < r->trace->k; j++) pending.push_back(Column{f, r, j, (int)j, kStash});
                }
                if (r->remaining == 0) {
                    r->done = true;
                    f->cv.notify_all();
                }
            }
            for (size_t i = 0; i < active.size();) {
                FitTask* f = active[i];
                if (!f->finished) {
                    i++;
                    continue;
                }
                free_sys.push_back(f->sys);
                tbufs.b[(size_t)f->sys].rows_fit = nullptr;   // a later fit may reuse this address
                f->sys = -1;
                active.erase(active.begin() + (ptrdiff_t)i);
                // a released fit may be freed at once by saige_poll: nothing
                // here may point at it any more
                resumed.erase(std::remove_if(resumed.begin(), resumed.end(),
                                             [f](const std::pair<FitTask*, int64_t>& fp) { return fp.first == f; }),
                              resumed.end());
                f->released = true;
                sch.done_cv.notify_all();
            }
        }
        // fill free slots: one batch of host solves, one of host products, and
        // runs of consecutive device columns
        const auto t_fill0 = clk::now();
        std::vector<int32_t> s_slots, s_sys, p_slots, p_sys;
        std::vector<Column> s_cols, p_cols;
        struct DevRun {
            bool solve;
            FitTask* fit;
            int col0;
            std::vector<int32_t> slots, systems;
        };
        std::vector<DevRun> runs;
        const Request* checked = nullptr;             // the request whose system was last checked
        while (!pending.empty() && !free_slots.empty()) {
            Column c = pending.front();
            pending.pop_front();
            FitTask* f = c.fit;
            if (c.req->trace && !c.req->trace->setup) trace_setup(f, *c.req->trace);
            if (c.req != checked) {
                if (c.req->solve || c.req->chain) {
                    // a fit has one request in flight, so its system is free to change
                    const bool same = f->sys_set && f->set_tau[0] == c.req->tau[0] &&
                                      f->set_tau[1] == c.req->tau[1] && f->set_w == c.req->w;
                    if (!same) {
                        set_system(f, c.req->w.data(), c.req->tau);
                        f->set_w = c.req->w;
                        f->set_tau[0] = c.req->tau[0];
                        f->set_tau[1] = c.req->tau[1];
                        f->sys_set = true;
                    }
                } else if (!f->sys_set) {
                    const double t[2] = {1.0, 0.0};
                    set_system(f, nullptr, t);          // products only need the sample mask
                    f->sys_set = true;
                    f->set_w.clear();
                }
                checked = c.req;
            }
            std::pop_heap(free_slots.begin(), free_slots.end(), std::greater<int>());
            const int slot = free_slots.back();
            free_slots.pop_back();
            owner[(size_t)slot] = c;
            live++;
            if (c.dev >= 0) {
                DevRun* run = runs.empty() ? nullptr : &runs.back();
                if (!run || run->solve != c.req->solve || run->fit != f ||
                    run->col0 + (int)run->slots.size() != c.dev) {
                    runs.push_back(DevRun{c.req->solve, f, c.dev, {}, {}});
                    run = &runs.back();
                }
                run->slots.push_back(slot);
                run->systems.push_back(f->sys);
                continue;
            }
            (c.req->solve ? s_slots : p_slots).push_back(slot);
            (c.req->solve ? s_sys : p_sys).push_back(f->sys);
            (c.req->solve ? s_cols : p_cols).push_back(c);
        }
        auto upload = [&](const std::vector<Column>& cs) {
            for (size_t k = 0; k < cs.size(); k++) {
                const saige_fit_spec* s = cs[k].fit->spec;
                const double* b = cs[k].req->B->col(cs[k].j);
                float* d = colbuf.data() + (int64_t)k * N;
                if (s->subset) std::fill(d, d + N, 0.0f);
                for (int64_t i = 0; i < s->n; i++) d[sample_of(s, i)] = (float)b[i];
            }
        };
        const auto t_up0 = clk::now();
        if (!s_cols.empty()) {
            upload(s_cols);
            check(saige_engine_add_batch(sch.ctx, (int)s_slots.size(), s_slots.data(), s_sys.data(), colbuf.data(),
                                         sch.opt.maxiter_pcg, (float)sch.opt.tol_pcg));
        }
        if (!p_cols.empty()) {
            upload(p_cols);
            check(saige_engine_add_products(sch.ctx, (int)p_slots.size(), p_slots.data(), p_sys.data(),
                                            colbuf.data()));
        }
        t_upload += secs(t_up0);
        const auto t_dev0 = clk::now();
        for (const DevRun& r : runs) {
            const float* B = tbufs.b[(size_t)r.fit->sys].U + (int64_t)r.col0 * N;
            if (r.solve)
                check(saige_engine_add_batch_device(sch.ctx, (int)r.slots.size(), r.slots.data(), r.systems.data(), B,
                                                    sch.opt.maxiter_pcg, (float)sch.opt.tol_pcg));
            else
                check(saige_engine_add_products_device(sch.ctx, (int)r.slots.size(), r.slots.data(),
                                                       r.systems.data(), B));
        }
        t_devadd += secs(t_dev0);
        t_fill += secs(t_fill0);
        if (live == 0) continue;
        // the step runs on a helper thread while the fits that the previous
        // step resumed compute their next request, so that its columns join
        // the following iteration; nothing else calls the library meanwhile
        int step_rc = SAIGE_OK;
        std::string step_err;
        std::thread stepper([&] {
            step_rc = saige_engine_step(sch.ctx, 1, done.data(), iters.data());
            if (step_rc != SAIGE_OK) step_err = saige_last_error();   // the message is per thread
        });
        const auto t_step0 = clk::now();
        if (!resumed.empty()) {
            std::unique_lock<std::mutex> lk(sch.mu);
            sch.cv.wait(lk, [&] {
                for (auto& fp : resumed)
                    if (!fp.first->finished && fp.first->posts <= fp.second) return false;
                return true;
            });
            resumed.clear();
        }
        t_wait_posts += secs(t_step0);
        stepper.join();
        t_step += secs(t_step0);
        n_steps++;
        if (step_rc != SAIGE_OK) throw std::runtime_error(step_err);
        const auto t_take0 = clk::now();
        std::vector<int32_t> fin, host;
        for (int s = 0; s < sch.width; s++)
            if (owner[(size_t)s].req && done[(size_t)s]) {
                fin.push_back(s);
                if (owner[(size_t)s].act == kHost) host.push_back(s);
            }
        if (fin.empty()) continue;
        // device columns of a trace round: K u to AU with the h dots, or the
        // Sigma^-1 u dots d (and e); one take per fit and kind
        const auto t_td0 = clk::now();
        {
            std::vector<int32_t> sl, cl;
            std::vector<double> dots;
            std::vector<char> seen(fin.size(), 0);
            for (size_t a = 0; a < fin.size(); a++) {
                const Column& c0 = owner[(size_t)fin[a]];
                if (c0.act == kHost || seen[a]) continue;
                sl.clear();
                cl.clear();
                for (size_t b = a; b < fin.size(); b++) {
                    const Column& c = owner[(size_t)fin[b]];
                    if (!seen[b] && c.act == c0.act && c.fit == c0.fit) {
                        seen[b] = 1;
                        sl.push_back(fin[b]);
                        cl.push_back(c.dev);
                    }
                }
                TraceState& t = *c0.req->trace;
                const TraceBuffers& tb = tbufs.b[(size_t)c0.fit->sys];
                const int kk = (int)sl.size();
                if (c0.act == kStash) {
                    dots.resize((size_t)(kk * t.p));
                    check(saige_engine_take_batch_ex(sch.ctx, kk, sl.data(), nullptr, tb.AU, N, cl.data(), nullptr,
                                                     nullptr, N, nullptr, tb.SX, (int)t.p, N, dots.data()));
                    for (int q = 0; q < kk; q++)
                        for (int64_t c = 0; c < t.p; c++) t.out->h(cl[(size_t)q], c) = dots[(size_t)(q * t.p + c)];
                } else {
                    const int nd = t.quantitative ? 2 : 1;
                    dots.resize((size_t)(kk * nd));
                    check(saige_engine_take_batch_ex(sch.ctx, kk, sl.data(), nullptr, nullptr, N, nullptr, tb.AU,
                                                     t.quantitative ? tb.U : nullptr, N, cl.data(), nullptr, 0, N,
                                                     dots.data()));
                    for (int q = 0; q < kk; q++) {
                        t.out->d[(size_t)cl[(size_t)q]] = dots[(size_t)(q * nd)];
                        if (t.quantitative) t.out->e[(size_t)cl[(size_t)q]] = dots[(size_t)(q * nd + 1)];
                    }
                }
            }
        }
        t_take_dots += secs(t_td0);
        const auto t_th0 = clk::now();
        if (!host.empty()) check(saige_engine_take_batch(sch.ctx, (int)host.size(), host.data(), colbuf.data()));
        t_take_host += secs(t_th0);
        std::lock_guard<std::mutex> lk(sch.mu);
        size_t hk = 0;
        for (const int32_t slot : fin) {
            Column c = owner[(size_t)slot];
            owner[(size_t)slot] = Column{nullptr, nullptr, 0};
            free_slots.push_back(slot);
            std::push_heap(free_slots.begin(), free_slots.end(), std::greater<int>());
            live--;
            if (c.act == kHost) {
                const saige_fit_spec* s = c.fit->spec;
                const float* x = colbuf.data() + (int64_t)(hk++) * N;
                double* o = c.req->out->col(c.j);
                for (int64_t i = 0; i < s->n; i++) o[i] = (double)x[sample_of(s, i)];
            }
            if (--c.req->remaining == 0) {
                Request* r = c.req;
                if (r->chain && !r->solve) {
                    // the products are done: queue the solves now, the fit sleeps on
                    r->C = saige::fit::gather_columns(*r->B, *r->out, r->src);
                    r->solve = true;
                    r->B = &r->C;
                    r->out = r->chain_out;
                    r->remaining = r->C.cols;
                    for (int64_t j = 0; j < r->C.cols; j++) pending.push_back(Column{c.fit, r, j});
                } else if (r->trace && r->trace->phase == 1) {
                    // phase 2: Sigma^-1 [U, K PY, (PY)], the U solves as dots only
                    TraceState& t = *r->trace;
                    t.phase = 2;
                    const int64_t n = c.fit->spec->n;
                    if (t.PY) {
                        t.host2 = Matrix(n, t.quantitative ? 2 : 1);
                        std::copy(t.res1.col(0), t.res1.col(0) + n, t.host2.col(0));
                        if (t.quantitative) std::copy(t.PY->begin(), t.PY->end(), t.host2.col(1));
                        t.res2 = Matrix(n, t.host2.cols);
                    }
                    r->solve = true;
                    r->B = &t.host2;
                    r->out = &t.res2;
                    r->remaining = t.k + t.host2.cols;
                    for (int64_t j = 0; j < t.k; j++) pending.push_back(Column{c.fit, r, j, (int)j, kPair});
                    for (int64_t j = 0; j < t.host2.cols; j++) pending.push_back(Column{c.fit, r, j});
                } else if (r->trace) {
                    TraceState& t = *r->trace;
                    if (t.PY) {
                        const int64_t n = c.fit->spec->n;
                        t.out->APY.assign(t.res1.col(0), t.res1.col(0) + n);
                        t.out->SAPY.assign(t.res2.col(0), t.res2.col(0) + n);
                        if (t.quantitative) t.out->SPY.assign(t.res2.col(1), t.res2.col(1) + n);
                    }
                }
                if (r->remaining == 0) {
                    r->done = true;
                    resumed.emplace_back(c.fit, c.fit->posts);
                    c.fit->cv.notify_all();
                }
            }
        }
        t_take += secs(t_take0);
    }
    if (timing)
        std::fprintf(stderr,
                     "saige engine: %lld steps; idle (waiting for requests) %.4f s, fill %.4f s (of which trace "
                     "setup %.4f s, systems %.4f s, host-column uploads %.4f s, device-column adds %.4f s), step "
                     "%.4f s (of which waiting for resumed fits %.4f s), take %.4f s (of which dot takes %.4f s, "
                     "host-column downloads %.4f s); fits' host work between requests %.4f s in all\n",
                     (long long)n_steps, t_idle, t_fill, t_setup, t_sys, t_upload, t_devadd, t_step, t_wait_posts,
                     t_take, t_take_dots, t_take_host, sch.t_fit_host);
    } catch (const std::exception& e) {
        // the engine failed: fail every request in the inbox, waiting for
        // slots or in a slot, so that every admitted fit's thread ends; later
        // requests fail at once (Scheduler::post)
        std::lock_guard<std::mutex> lk(sch.mu);
        sch.fatal = e.what();
        auto fail = [&](FitTask* f, Request* r) {
            if (r && !r->done) {
                r->error = sch.fatal;
                r->done = true;
                f->cv.notify_all();
            }
        };
        for (auto& fr : sch.inbox) fail(fr.first, fr.second);
        sch.inbox.clear();
        for (auto& c : pending) fail(c.fit, c.req);
        for (auto& c : owner) fail(c.fit, c.req);
        throw;
    }
}

}  // namespace

extern "C" {

void saige_fit_options_default(saige_fit_options* o) {
    if (!o) return;
    const saige::fit::Options d;
    o->maxiter = d.maxiter;
    o->tol = d.tol;
    o->nrun = d.nrun;
    o->tol_pcg = d.tol_pcg;
    o->maxiter_pcg = d.maxiter_pcg;
    o->trace_cv_cutoff = d.trace_cv_cutoff;
    o->probe_seed = d.probe_seed;
}

}  // extern "C"

namespace {

// What R's glm() and binomial() refuse before any arithmetic: non-finite
// responses, covariates or offsets, and binomial responses outside [0, 1]
// ("y values must be 0 <= y <= 1").  The IRLS loops assume finite values
// (their array exp/log are built for finite arguments).
bool valid_glm_inputs(int trait, int64_t n, const double* y, const double* X, int p, const double* offset,
                      const char* who) {
    for (int64_t i = 0; i < n; i++) {
        if (!std::isfinite(y[i]) || (offset && !std::isfinite(offset[i]))) {
            saige::set_error("%s: y or offset of sample %lld is not finite", who, (long long)i);
            return false;
        }
        if (trait == SAIGE_TRAIT_BINARY && (y[i] < 0.0 || y[i] > 1.0)) {
            saige::set_error("%s: binomial y of sample %lld is %g, outside [0, 1]", who, (long long)i, y[i]);
            return false;
        }
    }
    if (X)
        for (int64_t i = 0; i < n * (int64_t)p; i++)
            if (!std::isfinite(X[i])) {
                saige::set_error("%s: covariate %lld of sample %lld is not finite", who, (long long)(i / n),
                                 (long long)(i % n));
                return false;
            }
    return true;
}

// A finished fit's outputs into the caller's result (its buffers sized by
// the spec); false when the fit failed (status says why).
bool fill_result(const FitTask& f, const saige_fit_spec& s, int engine_rc, saige_fit_result& out) {
    const bool ran = f.finished && f.error.empty();
    out.status = ran ? SAIGE_OK : (engine_rc != SAIGE_OK ? SAIGE_E_DEVICE : SAIGE_E_STATE);
    if (!ran) return false;
    const auto& r = f.result;
    out.theta[0] = r.theta[0];
    out.theta[1] = r.theta[1];
    out.converged = r.converged ? 1 : 0;
    out.iterations = r.iterations;
    std::memcpy(out.coefficients, r.coefficients.data(), sizeof(double) * (size_t)s.p);
    if (out.linear_predictors)
        std::memcpy(out.linear_predictors, r.linear_predictors.data(), sizeof(double) * (size_t)s.n);
    if (out.fitted_values) std::memcpy(out.fitted_values, r.fitted_values.data(), sizeof(double) * (size_t)s.n);
    if (out.working_y) std::memcpy(out.working_y, r.working_y.data(), sizeof(double) * (size_t)s.n);
    if (out.cov) std::memcpy(out.cov, r.cov.a.data(), sizeof(double) * (size_t)(s.p * s.p));
    return true;
}

// The checks saige_fit_null_batch and saige_submit make on one request.
int check_spec(const saige_fit_spec& s, int64_t N, const char* who) {
    if (s.n <= 0 || s.p <= 0 || !s.y || !s.X || (!s.subset && s.n != N) || s.n > N ||
        (s.trait != SAIGE_TRAIT_BINARY && s.trait != SAIGE_TRAIT_QUANTITATIVE)) {
        saige::set_error("%s: invalid fit specification", who);
        return SAIGE_E_ARG;
    }
    if (s.subset)
        for (int64_t k = 0; k < s.n; k++)
            if (s.subset[k] < 0 || s.subset[k] >= N || (k > 0 && s.subset[k] <= s.subset[k - 1])) {
                saige::set_error("%s: the sample subset must be ascending indices below %lld", who, (long long)N);
                return SAIGE_E_ARG;
            }
    if (!valid_glm_inputs(s.trait, s.n, s.y, s.X, s.p, s.offset, who)) return SAIGE_E_ARG;
    return SAIGE_OK;
}

}  // namespace

// A request server on one context: the scheduler loop on a thread of its
// own, fed by saige_submit; tasks stay until saige_poll hands their result
// over.
struct saige_server {
    Scheduler sch;
    int n_systems = 16;
    std::thread th;
    std::map<int64_t, std::unique_ptr<FitTask>> tasks;   // under sch.mu
    int64_t next_ticket = 1;
    bool ended = false;          // the scheduler thread has returned
    int rc = SAIGE_OK;
    std::string fatal;
};

extern "C" {

int saige_fit_null_batch(saige_ctx* ctx, int n_fits, const saige_fit_spec* specs, const saige_fit_options* o,
                         int max_concurrent, saige_fit_result* results) {
    if (!ctx || n_fits < 0 || (n_fits > 0 && (!specs || !results))) return SAIGE_E_ARG;
    const int64_t N = saige_get_nnomissing(ctx);
    for (int i = 0; i < n_fits; i++) {
        if (!results[i].coefficients) return SAIGE_E_ARG;
        if (int rc = check_spec(specs[i], N, "saige_fit_null_batch")) return rc;
    }
    Scheduler sch;
    sch.ctx = ctx;
    sch.N = N;
    if (o) {
        sch.opt.maxiter = o->maxiter;
        sch.opt.tol = o->tol;
        sch.opt.nrun = o->nrun;
        sch.opt.tol_pcg = o->tol_pcg;
        sch.opt.maxiter_pcg = o->maxiter_pcg;
        sch.opt.trace_cv_cutoff = o->trace_cv_cutoff;
        sch.opt.probe_seed = o->probe_seed;
    }
    std::vector<std::unique_ptr<FitTask>> fits;
    for (int i = 0; i < n_fits; i++) {
        fits.push_back(std::make_unique<FitTask>());
        fits.back()->sch = &sch;
        fits.back()->spec = &specs[i];
    }
    int n_systems = std::max(1, std::min(n_fits, max_concurrent > 0 ? max_concurrent : 16));
    int rc = SAIGE_OK;
    std::string fatal;
    try {
        {
            std::lock_guard<std::mutex> lk(sch.mu);
            for (auto& f : fits) {
                f->t_submit = std::chrono::steady_clock::now();
                sch.queue.push_back(f.get());
            }
            sch.stopping = true;
        }
        run_scheduler(sch, n_systems);
    } catch (const std::exception& e) {
        rc = SAIGE_E_DEVICE;
        fatal = e.what();
    }
    for (auto& f : fits)
        if (f->th.joinable()) f->th.join();
    std::string first_error;
    for (int i = 0; i < n_fits; i++) {
        const FitTask& f = *fits[(size_t)i];
        if (!fill_result(f, specs[i], rc, results[i]) && first_error.empty())
            first_error = f.error.empty() ? fatal : f.error;
    }
    if (!first_error.empty()) saige::set_error("%s", first_error.c_str());
    return rc;
}

int saige_server_start(saige_ctx* ctx, const saige_fit_options* o, int max_concurrent, int64_t hbm_reserve_bytes,
                       saige_server** out) {
    if (!ctx || !out || hbm_reserve_bytes < 0) return SAIGE_E_ARG;
    const int64_t N = saige_get_nnomissing(ctx);
    if (N <= 0) {
        saige::set_error("saige_server_start: no genotypes loaded");
        return SAIGE_E_STATE;
    }
    auto srv = std::make_unique<saige_server>();
    srv->sch.ctx = ctx;
    srv->sch.N = N;
    if (o) {
        srv->sch.opt.maxiter = o->maxiter;
        srv->sch.opt.tol = o->tol;
        srv->sch.opt.nrun = o->nrun;
        srv->sch.opt.tol_pcg = o->tol_pcg;
        srv->sch.opt.maxiter_pcg = o->maxiter_pcg;
        srv->sch.opt.trace_cv_cutoff = o->trace_cv_cutoff;
        srv->sch.opt.probe_seed = o->probe_seed;
    }
    srv->sch.hbm_reserve = hbm_reserve_bytes;
    srv->n_systems = std::max(1, max_concurrent > 0 ? max_concurrent : 16);
    saige_server* sp = srv.get();
    sp->th = std::thread([sp] {
        int rc = SAIGE_OK;
        std::string err;
        try {
            run_scheduler(sp->sch, sp->n_systems);
        } catch (const std::exception& e) {
            rc = SAIGE_E_DEVICE;
            err = e.what();
        }
        std::lock_guard<std::mutex> lk(sp->sch.mu);
        sp->rc = rc;
        sp->fatal = err;
        sp->ended = true;
        sp->sch.done_cv.notify_all();
    });
    *out = srv.release();
    return SAIGE_OK;
}

int saige_submit(saige_server* srv, const saige_fit_spec* spec, int64_t* ticket) {
    if (!srv || !spec || !ticket) return SAIGE_E_ARG;
    if (int rc = check_spec(*spec, srv->sch.N, "saige_submit")) return rc;
    auto f = std::make_unique<FitTask>();
    const saige_fit_spec& s = *spec;
    f->own_y.assign(s.y, s.y + s.n);
    f->own_X.assign(s.X, s.X + s.n * (int64_t)s.p);
    if (s.offset) f->own_offset.assign(s.offset, s.offset + s.n);
    if (s.subset) f->own_subset.assign(s.subset, s.subset + s.n);
    f->own_spec = s;
    f->own_spec.y = f->own_y.data();
    f->own_spec.X = f->own_X.data();
    f->own_spec.offset = s.offset ? f->own_offset.data() : nullptr;
    f->own_spec.subset = s.subset ? f->own_subset.data() : nullptr;
    f->spec = &f->own_spec;
    f->sch = &srv->sch;
    std::lock_guard<std::mutex> lk(srv->sch.mu);
    if (srv->ended || srv->sch.stopping) {
        saige::set_error("saige_submit: the server has stopped%s%s", srv->fatal.empty() ? "" : ": ",
                         srv->fatal.c_str());
        return SAIGE_E_STATE;
    }
    f->t_submit = std::chrono::steady_clock::now();
    srv->sch.queue.push_back(f.get());
    *ticket = srv->next_ticket++;
    srv->tasks.emplace(*ticket, std::move(f));
    srv->sch.cv.notify_all();
    return SAIGE_OK;
}

int saige_poll(saige_server* srv, int64_t ticket, int wait_ms, saige_fit_result* result, double* times) {
    if (!srv || !result || !result->coefficients) return SAIGE_E_ARG;
    std::unique_lock<std::mutex> lk(srv->sch.mu);
    if (srv->tasks.find(ticket) == srv->tasks.end()) {
        saige::set_error("saige_poll: unknown ticket %lld", (long long)ticket);
        return SAIGE_E_ARG;
    }
    // the task is looked up afresh after every wake: another thread's poll
    // of the same ticket may have collected (and freed) it meanwhile
    auto ready = [&] {
        const auto jt = srv->tasks.find(ticket);
        if (jt == srv->tasks.end()) return true;
        const FitTask* g = jt->second.get();
        return g->released || (srv->ended && (g->finished || !g->th.joinable()));
    };
    if (wait_ms < 0)
        srv->sch.done_cv.wait(lk, ready);
    else if (wait_ms > 0)
        srv->sch.done_cv.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
    if (!ready()) return SAIGE_PENDING;
    auto it = srv->tasks.find(ticket);
    if (it == srv->tasks.end()) {
        saige::set_error("saige_poll: ticket %lld was collected by another poll", (long long)ticket);
        return SAIGE_E_ARG;
    }
    std::unique_ptr<FitTask> task = std::move(it->second);
    srv->tasks.erase(it);
    const int engine_rc = srv->rc;
    const std::string fatal = srv->fatal;
    lk.unlock();
    if (task->th.joinable()) task->th.join();
    if (times) {
        using sec = std::chrono::duration<double>;
        const bool admitted = task->t_admit.time_since_epoch().count() != 0;
        times[0] = admitted ? sec(task->t_admit - task->t_submit).count() : -1.0;
        times[1] = task->finished ? sec(task->t_done - task->t_submit).count() : -1.0;
    }
    if (!fill_result(*task, task->own_spec, engine_rc, *result))
        saige::set_error("%s", task->error.empty() ? (fatal.empty() ? "not run" : fatal.c_str()) : task->error.c_str());
    return SAIGE_OK;
}

int saige_server_stop(saige_server* srv) {
    if (!srv) return SAIGE_E_ARG;
    {
        std::lock_guard<std::mutex> lk(srv->sch.mu);
        srv->sch.stopping = true;
        srv->sch.cv.notify_all();
    }
    if (srv->th.joinable()) srv->th.join();
    for (auto& kv : srv->tasks)
        if (kv.second->th.joinable()) kv.second->th.join();
    const int rc = srv->rc;
    if (rc != SAIGE_OK) saige::set_error("%s", srv->fatal.c_str());
    delete srv;
    return rc;
}

int saige_covariate_transform(const double* X, int64_t n, int p, double* X_out, double* R_out, int32_t* kept,
                              int* p_kept) {
    if (!X || n <= 0 || p <= 0 || !X_out || !p_kept) return SAIGE_E_ARG;
    saige::fit::Matrix Xm(n, p);
    std::memcpy(Xm.a.data(), X, sizeof(double) * (size_t)(n * p));
    const auto t = saige::fit::covariate_transform(Xm);
    *p_kept = (int)t.kept.size();
    std::memcpy(X_out, t.X.a.data(), sizeof(double) * t.X.a.size());
    if (R_out) std::memcpy(R_out, t.R.a.data(), sizeof(double) * t.R.a.size());
    if (kept)
        for (size_t i = 0; i < t.kept.size(); i++) kept[i] = t.kept[i];
    return SAIGE_OK;
}

int saige_score_test_null(const double* mu, const double* V, const double* y, const double* X, int64_t n, int p,
                          double* XV, double* XVX, double* XVX_inv, double* XXVX_inv, double* XVX_inv_XV,
                          double* S_a) {
    if (!mu || !V || !y || !X || n <= 0 || p <= 0) return SAIGE_E_ARG;
    try {
        saige::fit::Matrix Xm(n, p);
        std::memcpy(Xm.a.data(), X, sizeof(double) * (size_t)(n * p));
        const auto o = saige::fit::score_test_null(std::vector<double>(mu, mu + n), std::vector<double>(V, V + n),
                                                   std::vector<double>(y, y + n), Xm);
        if (XV) std::memcpy(XV, o.XV.a.data(), sizeof(double) * o.XV.a.size());
        if (XVX) std::memcpy(XVX, o.XVX.a.data(), sizeof(double) * o.XVX.a.size());
        if (XVX_inv) std::memcpy(XVX_inv, o.XVX_inv.a.data(), sizeof(double) * o.XVX_inv.a.size());
        if (XXVX_inv) std::memcpy(XXVX_inv, o.XXVX_inv.a.data(), sizeof(double) * o.XXVX_inv.a.size());
        if (XVX_inv_XV) std::memcpy(XVX_inv_XV, o.XVX_inv_XV.a.data(), sizeof(double) * o.XVX_inv_XV.a.size());
        if (S_a) std::memcpy(S_a, o.S_a.data(), sizeof(double) * o.S_a.size());
    } catch (const std::exception&) {
        return SAIGE_E_ARG;
    }
    return SAIGE_OK;
}

int saige_glm_start(int trait, int64_t n, const double* y, const double* X, int p, const double* offset,
                    double* coef, double* eta, double* mu) {
    if (!y || !X || n <= 0 || p <= 0 || !coef || !eta || !mu ||
        (trait != SAIGE_TRAIT_BINARY && trait != SAIGE_TRAIT_QUANTITATIVE))
        return SAIGE_E_ARG;
    if (!valid_glm_inputs(trait, n, y, X, p, offset, "saige_glm_start")) return SAIGE_E_ARG;
    try {
        saige::fit::Matrix Xm(n, p);
        std::memcpy(Xm.a.data(), X, sizeof(double) * (size_t)(n * p));
        const std::vector<double> off = offset ? std::vector<double>(offset, offset + n) : std::vector<double>();
        const auto g = saige::fit::glm_start(trait == SAIGE_TRAIT_BINARY ? saige::fit::Trait::Binary
                                                                         : saige::fit::Trait::Quantitative,
                                             std::vector<double>(y, y + n), Xm, off);
        std::memcpy(coef, g.coef.data(), sizeof(double) * (size_t)p);
        std::memcpy(eta, g.eta.data(), sizeof(double) * (size_t)n);
        std::memcpy(mu, g.mu.data(), sizeof(double) * (size_t)n);
    } catch (const std::exception&) {
        return SAIGE_E_ARG;
    }
    return SAIGE_OK;
}

int saige_glm_working(int trait, int64_t n, const double* y, const double* eta, const double* offset, double* mu,
                      double* Y, double* W) {
    if (!y || !eta || n <= 0 || !mu || !Y || !W || (trait != SAIGE_TRAIT_BINARY && trait != SAIGE_TRAIT_QUANTITATIVE))
        return SAIGE_E_ARG;
    if (!valid_glm_inputs(trait, n, y, nullptr, 0, offset, "saige_glm_working")) return SAIGE_E_ARG;
    for (int64_t i = 0; i < n; i++)
        if (!std::isfinite(eta[i])) {
            saige::set_error("saige_glm_working: eta of sample %lld is not finite", (long long)i);
            return SAIGE_E_ARG;
        }
    saige::fit::glm_working(trait == SAIGE_TRAIT_BINARY ? saige::fit::Trait::Binary : saige::fit::Trait::Quantitative,
                            n, y, eta, offset, mu, Y, W);
    return SAIGE_OK;
}

int saige_coefficients(int64_t n, int p, const double* Y, const double* X, const double* w, double tau0,
                       const double* sol, double* cov, double* alpha, double* eta) {
    if (!Y || !X || !w || !sol || n <= 0 || p <= 0 || !cov || !alpha || !eta) return SAIGE_E_ARG;
    saige::fit::coefficients_from_solution(n, p, Y, X, w, tau0, sol, cov, alpha, eta);
    return SAIGE_OK;
}

}  // extern "C"
