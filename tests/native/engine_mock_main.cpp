// Test harness only: runs the C++ scheduler and request server
// (saige-gpu_amd/csrc/fit_engine.cpp) over the CPU engine stand-in
// (engine_mock.cpp) from several threads, so that the host sanitizers
// (AddressSanitizer + UndefinedBehaviorSanitizer, or ThreadSanitizer) watch
// the whole path: a batch, each fit alone (bitwise equal), the server with
// two submitting and two polling threads and one ticket polled by two
// threads at once, and an injected engine failure in both.  Prints
// "all scenarios passed" and returns 0, or names the first mismatch.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "saige_hip.h"

extern "C" {
saige_ctx* mock_create(const double* K, int64_t N);
void mock_destroy(saige_ctx* c);
void mock_fail_step(saige_ctx* c, int64_t step);
int64_t mock_steps(const saige_ctx* c);
int64_t mock_live_allocs(const saige_ctx* c);
}

namespace {
constexpr int64_t N = 160, M = 400;

struct Job {
    int trait;
    std::vector<double> y, X;   // X: n x 3, column-major
    std::vector<int64_t> subset;
    int64_t n;
};
struct Out {
    double theta[2];
    std::vector<double> coef, fitted;
    int iterations, status;
};

bool same(const Out& a, const Out& b) {
    return std::memcmp(a.theta, b.theta, sizeof a.theta) == 0 && a.coef == b.coef && a.fitted == b.fitted &&
           a.iterations == b.iterations && a.status == b.status;
}

saige_fit_spec spec_of(const Job& j) {
    saige_fit_spec s{};
    s.trait = j.trait;
    s.n = j.n;
    s.subset = j.subset.empty() ? nullptr : j.subset.data();
    s.y = j.y.data();
    s.X = j.X.data();
    s.p = 3;
    s.offset = nullptr;
    return s;
}

void wire(Out& o, saige_fit_result& r, int64_t n) {
    o.coef.assign(3, 0.0);
    o.fitted.assign((size_t)n, 0.0);
    r = saige_fit_result{};
    r.coefficients = o.coef.data();
    r.fitted_values = o.fitted.data();
}

void unwire(Out& o, const saige_fit_result& r) {
    o.theta[0] = r.theta[0];
    o.theta[1] = r.theta[1];
    o.iterations = r.iterations;
    o.status = r.status;
}

int batch(saige_ctx* c, const std::vector<Job>& jobs, const saige_fit_options& opt, std::vector<Out>& out) {
    std::vector<saige_fit_spec> specs;
    std::vector<saige_fit_result> res(jobs.size());
    out.assign(jobs.size(), Out{});
    for (size_t i = 0; i < jobs.size(); i++) {
        specs.push_back(spec_of(jobs[i]));
        wire(out[i], res[i], jobs[i].n);
    }
    const int rc = saige_fit_null_batch(c, (int)jobs.size(), specs.data(), &opt, 16, res.data());
    for (size_t i = 0; i < jobs.size(); i++) unwire(out[i], res[i]);
    return rc;
}

// the server: two threads submit (alternate jobs), two threads poll them;
// the first request's ticket is polled by both pollers at once
int serve(saige_ctx* c, const std::vector<Job>& jobs, const saige_fit_options& opt, std::vector<Out>& out,
          int* stop_rc) {
    saige_server* srv = nullptr;
    if (saige_server_start(c, &opt, 3, 0, &srv) != SAIGE_OK) return 1;
    // a submission after an engine failure is refused (the server has
    // stopped): that request then counts as failed and is not polled
    std::vector<int64_t> ticket(jobs.size(), 0);
    std::vector<int> src(jobs.size(), -100);
    std::vector<std::thread> th;
    for (int t = 0; t < 2; t++)
        th.emplace_back([&, t] {
            for (size_t i = (size_t)t; i < jobs.size(); i += 2) {
                std::this_thread::sleep_for(std::chrono::milliseconds(2 * (int)i));
                const saige_fit_spec s = spec_of(jobs[i]);
                int64_t tk = 0;
                src[i] = saige_submit(srv, &s, &tk);
                ticket[i] = tk;
            }
        });
    for (auto& t : th) t.join();
    th.clear();
    out.assign(jobs.size(), Out{});
    std::vector<saige_fit_result> res(jobs.size());
    std::vector<int> prc(jobs.size(), -100);
    for (size_t i = 0; i < jobs.size(); i++) wire(out[i], res[i], jobs[i].n);
    // a second result buffer for the doubly polled ticket
    Out extra;
    saige_fit_result extra_r{};
    wire(extra, extra_r, jobs[0].n);
    int extra_rc = -100;
    for (int t = 0; t < 2; t++)
        th.emplace_back([&, t] {
            for (size_t i = (size_t)t; i < jobs.size(); i += 2)
                if (src[i] == SAIGE_OK) prc[i] = saige_poll(srv, ticket[i], -1, &res[i], nullptr);
            if (t == 1 && src[0] == SAIGE_OK) extra_rc = saige_poll(srv, ticket[0], -1, &extra_r, nullptr);
        });
    for (auto& t : th) t.join();
    *stop_rc = saige_server_stop(srv);
    for (size_t i = 0; i < jobs.size(); i++) {
        unwire(out[i], res[i]);
        if (src[i] != SAIGE_OK) {
            if (src[i] != SAIGE_E_STATE) {
                std::printf("submission %zu returned %d\n", i, src[i]);
                return 1;
            }
            out[i].status = src[i];
        }
    }
    if (src[0] == SAIGE_OK) {
        // ticket 0: exactly one of the two polls collected it
        const bool first = prc[0] == SAIGE_OK, second = extra_rc == SAIGE_OK;
        if (first == second || (prc[0] != SAIGE_OK && prc[0] != SAIGE_E_ARG) ||
            (extra_rc != SAIGE_OK && extra_rc != SAIGE_E_ARG)) {
            std::printf("doubly polled ticket: return codes %d and %d\n", prc[0], extra_rc);
            return 1;
        }
        if (second) {
            unwire(extra, extra_r);
            out[0] = extra;
        }
    }
    for (size_t i = 1; i < jobs.size(); i++)
        if (src[i] == SAIGE_OK && prc[i] != SAIGE_OK) {
            std::printf("poll of request %zu returned %d\n", i, prc[i]);
            return 1;
        }
    return 0;
}
}  // namespace

int main() {
    std::mt19937_64 g(11);
    std::normal_distribution<double> nd;
    std::uniform_real_distribution<double> ud(0.05, 0.5);
    // K = Z Z^T / M from standardised Binomial(2, f) genotypes
    std::vector<double> Z((size_t)(N * M));
    for (int64_t j = 0; j < M; j++) {
        const double f = ud(g);
        std::binomial_distribution<int> bd(2, f);
        for (int64_t i = 0; i < N; i++) Z[(size_t)(j * N + i)] = (bd(g) - 2 * f) / std::sqrt(2 * f * (1 - f));
    }
    std::vector<double> K((size_t)(N * N), 0.0);
    for (int64_t a = 0; a < N; a++)
        for (int64_t b = 0; b <= a; b++) {
            double s = 0.0;
            for (int64_t j = 0; j < M; j++) s += Z[(size_t)(j * N + a)] * Z[(size_t)(j * N + b)];
            K[(size_t)(a * N + b)] = K[(size_t)(b * N + a)] = s / M;
        }
    // phenotypes: a genetic part from 100 markers, two covariates
    std::vector<double> gen((size_t)N, 0.0);
    for (int64_t j = 0; j < 100; j++) {
        const double b = nd(g) * 0.14;
        for (int64_t i = 0; i < N; i++) gen[(size_t)i] += b * Z[(size_t)(j * N + i)];
    }
    auto make_X = [&](const std::vector<int64_t>& rows) {
        std::vector<double> X(rows.size() * 3);
        std::mt19937_64 gx(3);
        std::vector<double> c1((size_t)N), c2((size_t)N);
        for (int64_t i = 0; i < N; i++) {
            c1[(size_t)i] = nd(gx);
            c2[(size_t)i] = nd(gx) > 0 ? 1.0 : 0.0;
        }
        const size_t n = rows.size();
        for (size_t r = 0; r < n; r++) {
            X[r] = 1.0;
            X[n + r] = c1[(size_t)rows[r]];
            X[2 * n + r] = c2[(size_t)rows[r]];
        }
        return X;
    };
    std::vector<int64_t> all((size_t)N);
    for (int64_t i = 0; i < N; i++) all[(size_t)i] = i;
    std::vector<int64_t> sub;
    for (int64_t i = 0; i < N; i += 4) sub.push_back(i), sub.push_back(i + 1), sub.push_back(i + 3);
    std::vector<Job> jobs;
    for (int k = 0; k < 5; k++) {
        Job j;
        j.trait = k == 2 ? SAIGE_TRAIT_QUANTITATIVE : SAIGE_TRAIT_BINARY;
        const std::vector<int64_t>& rows = k == 4 ? sub : all;
        j.n = (int64_t)rows.size();
        if (k == 4) j.subset = sub;
        j.X = make_X(rows);
        std::uniform_real_distribution<double> u(0, 1);
        for (int64_t r : rows) {
            const double eta = -1.0 + 0.3 * k + 1.2 * gen[(size_t)r];
            j.y.push_back(j.trait == SAIGE_TRAIT_BINARY ? (u(g) < 1 / (1 + std::exp(-eta)) ? 1.0 : 0.0)
                                                        : eta + nd(g));
        }
        jobs.push_back(std::move(j));
    }
    saige_fit_options opt;
    saige_fit_options_default(&opt);
    opt.nrun = 8;
    opt.tol_pcg = 1e-6;

    saige_ctx* c = mock_create(K.data(), N);
    std::vector<Out> mixed, alone, served;
    if (batch(c, jobs, opt, mixed) != SAIGE_OK) {
        std::printf("batch failed: %s\n", saige_last_error());
        return 1;
    }
    for (size_t i = 0; i < jobs.size(); i++) {
        if (mixed[i].status != SAIGE_OK) {
            std::printf("fit %zu status %d\n", i, mixed[i].status);
            return 1;
        }
        if (batch(c, {jobs[i]}, opt, alone) != SAIGE_OK || !same(alone[0], mixed[i])) {
            std::printf("fit %zu differs alone and in the batch\n", i);
            return 1;
        }
    }
    int stop_rc = 0;
    if (serve(c, jobs, opt, served, &stop_rc) != 0 || stop_rc != SAIGE_OK) return 1;
    for (size_t i = 0; i < jobs.size(); i++)
        if (!same(served[i], mixed[i])) {
            std::printf("served fit %zu differs from the batch\n", i);
            return 1;
        }
    // an engine failure: every fit ends with an error, nothing hangs
    mock_fail_step(c, mock_steps(c) + 4);
    std::vector<Out> failed;
    if (batch(c, jobs, opt, failed) != SAIGE_E_DEVICE) {
        std::printf("batch with a failing engine did not report it\n");
        return 1;
    }
    for (const Out& o : failed)
        if (o.status == SAIGE_OK) {
            std::printf("a fit succeeded over a failed engine\n");
            return 1;
        }
    mock_fail_step(c, mock_steps(c) + 6);
    if (serve(c, jobs, opt, served, &stop_rc) != 0 || stop_rc != SAIGE_E_DEVICE) {
        std::printf("server over a failing engine: stop returned %d\n", stop_rc);
        return 1;
    }
    for (const Out& o : served)
        if (o.status == SAIGE_OK) {
            std::printf("a served fit succeeded over a failed engine\n");
            return 1;
        }
    if (mock_live_allocs(c) != 0) {
        std::printf("%lld device buffers neither freed nor back on the context\n", (long long)mock_live_allocs(c));
        return 1;
    }
    mock_destroy(c);
    std::printf("all scenarios passed\n");
    return 0;
}
