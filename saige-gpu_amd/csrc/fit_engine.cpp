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
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/saige_hip.h"
#include "fit_core.hpp"

namespace saige {
void set_error(const char* fmt, ...);   // the library's last-error message (saige_capi.hip)
}

namespace {

using saige::fit::Matrix;

struct Scheduler;

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
    int64_t posts = 0;            // requests posted so far
    std::condition_variable cv;   // the fit waits here for its request

    void solve(const std::vector<double>& w, const double tau[2], const Matrix& B, Matrix& X) override;
    void product(const Matrix& B, Matrix& KB) override;
    void product_then_solve(const Matrix& B, Matrix& KB, const std::vector<double>& w, const double tau[2],
                            const std::vector<int>& src, Matrix& X) override;
    void submit(Request& r);
};

struct Column {
    FitTask* fit;
    Request* req;
    int64_t j;
};

struct Scheduler {
    saige_ctx* ctx;
    int64_t N;
    saige::fit::Options opt;
    int width = 64;
    std::mutex mu;
    std::condition_variable cv;           // the scheduler waits here for requests / finished fits
    std::deque<std::pair<FitTask*, Request*>> inbox;
    std::string fatal;                    // engine failure: every fit stops

    void post(FitTask* f, Request* r) {
        std::unique_lock<std::mutex> lk(mu);
        if (!fatal.empty()) {
            r->error = fatal;
            r->done = true;
            return;
        }
        inbox.emplace_back(f, r);
        f->posts++;
        cv.notify_all();
        f->cv.wait(lk, [&] { return r->done; });
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

int check(int rc) {
    if (rc != SAIGE_OK) throw std::runtime_error(saige_last_error());
    return rc;
}

// the fit's samples in the N phenotyped samples
int64_t sample_of(const saige_fit_spec* s, int64_t i) { return s->subset ? s->subset[i] : i; }

void run_batch(Scheduler& sch, std::vector<std::unique_ptr<FitTask>>& fits, int n_systems) {
    const int64_t N = sch.N;
    check(saige_engine_reserve(sch.ctx, sch.width, n_systems));
    std::deque<int> free_sys;
    for (int s = 0; s < n_systems; s++) free_sys.push_back(s);
    std::vector<int> free_slots;                      // lowest slot first (min-heap)
    for (int s = sch.width - 1; s >= 0; s--) free_slots.push_back(s);
    std::deque<FitTask*> waiting;
    for (auto& f : fits) waiting.push_back(f.get());
    std::deque<Column> pending;
    std::vector<Column> owner((size_t)sch.width, Column{nullptr, nullptr, 0});
    std::vector<uint8_t> done((size_t)sch.width);
    std::vector<int32_t> iters((size_t)sch.width);
    size_t n_finished = 0, live = 0;
    std::vector<float> colbuf;
    std::vector<uint8_t> mask((size_t)N);
    std::vector<float> wfull((size_t)N);
    // fits whose request the last step completed, with their post count then
    std::vector<std::pair<FitTask*, int64_t>> resumed;
    // wall seconds of the loop's parts, printed to stderr at the end when the
    // environment sets SAIGE_ENGINE_TIMING (diagnostics only)
    const bool timing = std::getenv("SAIGE_ENGINE_TIMING") != nullptr;
    double t_idle = 0, t_fill = 0, t_step = 0, t_wait_posts = 0, t_take = 0;
    int64_t n_steps = 0;
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };

    auto set_system = [&](FitTask* f, const double* w, const double tau[2]) {
        const saige_fit_spec* s = f->spec;
        std::fill(mask.begin(), mask.end(), (uint8_t)(s->subset ? 0 : 1));
        std::fill(wfull.begin(), wfull.end(), 1.0f);
        for (int64_t i = 0; i < s->n; i++) {
            const int64_t r = sample_of(s, i);
            mask[(size_t)r] = 1;
            wfull[(size_t)r] = w ? (float)w[i] : 1.0f;
        }
        const float t[2] = {(float)tau[0], (float)tau[1]};
        check(saige_engine_set_system(sch.ctx, f->sys, wfull.data(), t, mask.data()));
    };

    try {
    while (n_finished < fits.size()) {
        // admit fits while a system is free
        while (!waiting.empty() && !free_sys.empty()) {
            FitTask* f = waiting.front();
            waiting.pop_front();
            f->sys = free_sys.front();
            free_sys.pop_front();
            resumed.emplace_back(f, 0);               // its first request is awaited like a resumed fit's
            f->th = std::thread([f, &sch] {
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
                f->finished = true;
                sch.cv.notify_all();
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
                // more of them); with none such, for any request or finished fit
                sch.cv.wait(lk, [&] {
                    if (!resumed.empty()) {
                        for (auto& fp : resumed)
                            if (!fp.first->finished && fp.first->posts <= fp.second) return false;
                        return true;
                    }
                    if (!sch.inbox.empty()) return true;
                    for (auto& f : fits)
                        if (f->finished && f->sys >= 0) return true;
                    return false;
                });
                resumed.clear();
            }
            t_idle += secs(t0);
            while (!sch.inbox.empty()) {
                auto [f, r] = sch.inbox.front();
                sch.inbox.pop_front();
                r->remaining = r->B->cols;
                for (int64_t j = 0; j < r->B->cols; j++) pending.push_back(Column{f, r, j});
                if (r->B->cols == 0) {
                    r->done = true;
                    f->cv.notify_all();
                }
            }
            for (auto& f : fits)
                if (f->finished && f->sys >= 0) {
                    free_sys.push_back(f->sys);
                    f->sys = -1;
                    n_finished++;
                }
        }
        // fill free slots: one batch of solves and one of products
        const auto t_fill0 = clk::now();
        std::vector<int32_t> s_slots, s_sys, p_slots, p_sys;
        std::vector<Column> s_cols, p_cols;
        std::deque<Column> keep;
        while (!pending.empty() && !free_slots.empty()) {
            Column c = pending.front();
            pending.pop_front();
            FitTask* f = c.fit;
            if (c.req->solve || c.req->chain) {
                // a fit has one request in flight, so its system is free to change
                const bool same = f->sys_set && f->set_tau[0] == c.req->tau[0] && f->set_tau[1] == c.req->tau[1] &&
                                  f->set_w == c.req->w;
                if (!same) {
                    set_system(f, c.req->w.data(), c.req->tau);
                    f->set_w = c.req->w;
                    f->set_tau[0] = c.req->tau[0];
                    f->set_tau[1] = c.req->tau[1];
                    f->sys_set = true;
                }
            } else if (!f->sys_set) {
                const double t[2] = {1.0, 0.0};
                set_system(f, nullptr, t);              // products only need the sample mask
                f->sys_set = true;
                f->set_w.clear();
            }
            std::pop_heap(free_slots.begin(), free_slots.end(), std::greater<int>());
            const int slot = free_slots.back();
            free_slots.pop_back();
            owner[(size_t)slot] = c;
            (c.req->solve ? s_slots : p_slots).push_back(slot);
            (c.req->solve ? s_sys : p_sys).push_back(f->sys);
            (c.req->solve ? s_cols : p_cols).push_back(c);
            live++;
        }
        auto upload = [&](const std::vector<Column>& cs) {
            colbuf.assign((size_t)(N * (int64_t)cs.size()), 0.0f);
            for (size_t k = 0; k < cs.size(); k++) {
                const saige_fit_spec* s = cs[k].fit->spec;
                const double* b = cs[k].req->B->col(cs[k].j);
                float* d = colbuf.data() + (int64_t)k * N;
                for (int64_t i = 0; i < s->n; i++) d[sample_of(s, i)] = (float)b[i];
            }
        };
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
        std::vector<int32_t> fin;
        for (int s = 0; s < sch.width; s++)
            if (owner[(size_t)s].req && done[(size_t)s]) fin.push_back(s);
        if (fin.empty()) continue;
        colbuf.assign((size_t)(N * (int64_t)fin.size()), 0.0f);
        check(saige_engine_take_batch(sch.ctx, (int)fin.size(), fin.data(), colbuf.data()));
        std::lock_guard<std::mutex> lk(sch.mu);
        for (size_t k = 0; k < fin.size(); k++) {
            const int slot = fin[k];
            Column c = owner[(size_t)slot];
            owner[(size_t)slot] = Column{nullptr, nullptr, 0};
            free_slots.push_back(slot);
            std::push_heap(free_slots.begin(), free_slots.end(), std::greater<int>());
            live--;
            const saige_fit_spec* s = c.fit->spec;
            const float* x = colbuf.data() + (int64_t)k * N;
            double* o = c.req->out->col(c.j);
            for (int64_t i = 0; i < s->n; i++) o[i] = (double)x[sample_of(s, i)];
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
                     "saige engine: %lld steps; idle (waiting for requests) %.4f s, fill %.4f s, step %.4f s "
                     "(of which waiting for resumed fits %.4f s), take %.4f s\n",
                     (long long)n_steps, t_idle, t_fill, t_step, t_wait_posts, t_take);
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

}  // namespace

extern "C" {

int saige_fit_null_batch(saige_ctx* ctx, int n_fits, const saige_fit_spec* specs, const saige_fit_options* o,
                         int max_concurrent, saige_fit_result* results) {
    if (!ctx || n_fits < 0 || (n_fits > 0 && (!specs || !results))) return SAIGE_E_ARG;
    const int64_t N = saige_get_nnomissing(ctx);
    for (int i = 0; i < n_fits; i++) {
        const saige_fit_spec& s = specs[i];
        if (s.n <= 0 || s.p <= 0 || !s.y || !s.X || (!s.subset && s.n != N) || s.n > N ||
            (s.trait != SAIGE_TRAIT_BINARY && s.trait != SAIGE_TRAIT_QUANTITATIVE) || !results[i].coefficients)
            return SAIGE_E_ARG;
        if (!valid_glm_inputs(s.trait, s.n, s.y, s.X, s.p, s.offset, "saige_fit_null_batch")) return SAIGE_E_ARG;
        if (s.subset)
            for (int64_t k = 0; k < s.n; k++)
                if (s.subset[k] < 0 || s.subset[k] >= N || (k > 0 && s.subset[k] <= s.subset[k - 1]))
                    return SAIGE_E_ARG;
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
        run_batch(sch, fits, n_systems);
    } catch (const std::exception& e) {
        rc = SAIGE_E_DEVICE;
        fatal = e.what();
    }
    for (auto& f : fits)
        if (f->th.joinable()) f->th.join();
    std::string first_error;
    for (int i = 0; i < n_fits; i++) {
        saige_fit_result& out = results[i];
        const saige_fit_spec& s = specs[i];
        const FitTask& f = *fits[(size_t)i];
        const bool ran = f.finished && f.error.empty();
        out.status = ran ? SAIGE_OK : (rc != SAIGE_OK ? SAIGE_E_DEVICE : SAIGE_E_STATE);
        if (!ran) {
            if (first_error.empty()) first_error = f.error.empty() ? fatal : f.error;
            continue;
        }
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
    }
    if (!first_error.empty()) saige::set_error("%s", first_error.c_str());
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
