// Null-GLMM fitting logic (host, FP64); see fit_core.hpp.
#include "fit_core.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <tuple>

namespace saige {
namespace fit {

namespace {

// ------------------------------------------------------------- host pieces

// The O(n) loops of a fit (the working response, the coefficients after a
// solve, glm()'s IRLS) run as kPieces fixed pieces of the samples when n is
// at least kPieceMin: the pieces go to a small pool of host threads and the
// calling thread, and sums add the pieces' partial sums in piece order, so a
// result depends on n only, not on which thread ran a piece or how many the
// pool has (a fit's bits are the same alone and beside other fits).
constexpr int kPieces = 4;

// n from which a loop is cut: 32768, or SAIGE_HOST_PIECE_MIN (tests lower it
// to run the pieces on small problems)
int64_t piece_min() {
    static const int64_t v = [] {
        const char* e = std::getenv("SAIGE_HOST_PIECE_MIN");
        const long long x = e ? std::atoll(e) : 0;
        return x > 0 ? (int64_t)x : (int64_t)32768;
    }();
    return v;
}

int pieces_for(int64_t n) { return n >= piece_min() ? kPieces : 1; }
int64_t piece_begin(int64_t n, int c, int C) { return n * c / C; }

class HostPool {
public:
    static HostPool& get() {
        static HostPool* pool = new HostPool();   // never destroyed: its threads outlive static destructors
        return *pool;
    }
    // fn(c) for c = 0..C-1; piece 0 on the calling thread
    void run(int C, const std::function<void(int)>& fn) {
        if (C <= 1 || nthreads_ == 0) {
            for (int c = 0; c < C; c++) fn(c);
            return;
        }
        std::mutex m;
        std::condition_variable cv;
        int left = C - 1;
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (int c = 1; c < C; c++)
                queue_.push_back([&, c] {
                    fn(c);
                    std::lock_guard<std::mutex> l2(m);
                    if (--left == 0) cv.notify_one();
                });
        }
        cv_.notify_all();
        fn(0);
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return left == 0; });
    }

private:
    HostPool() {
        const unsigned hw = std::thread::hardware_concurrency();
        nthreads_ = hw > 4 ? std::min(12u, hw - 4) : 0u;
        for (unsigned t = 0; t < nthreads_; t++) std::thread([this] { work(); }).detach();
    }
    void work() {
        for (;;) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return !queue_.empty(); });
                f = std::move(queue_.front());
                queue_.pop_front();
            }
            f();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> queue_;
    unsigned nthreads_ = 0;
};

// fn(i0, i1, c) over the pieces [i0, i1) of [0, n)
template <class F>
void for_pieces(int64_t n, F fn) {
    const int C = pieces_for(n);
    HostPool::get().run(C, [&](int c) { fn(piece_begin(n, c, C), piece_begin(n, c + 1, C), c); });
}

// ------------------------------------------------------------ dense algebra

// Solve A X = B (A p x p) by Gaussian elimination with partial pivoting, as
// LAPACK's dgesv; throws if A is singular.
Matrix solve_dense(Matrix A, Matrix B) {
    const int64_t p = A.rows;
    std::vector<int64_t> piv(p);
    for (int64_t k = 0; k < p; k++) {
        int64_t r = k;
        for (int64_t i = k + 1; i < p; i++)
            if (std::fabs(A(i, k)) > std::fabs(A(r, k))) r = i;
        if (A(r, k) == 0.0) throw std::runtime_error("singular matrix");
        if (r != k) {
            for (int64_t j = 0; j < p; j++) std::swap(A(k, j), A(r, j));
            for (int64_t j = 0; j < B.cols; j++) std::swap(B(k, j), B(r, j));
        }
        for (int64_t i = k + 1; i < p; i++) {
            const double f = A(i, k) / A(k, k);
            A(i, k) = 0.0;
            for (int64_t j = k + 1; j < p; j++) A(i, j) -= f * A(k, j);
            for (int64_t j = 0; j < B.cols; j++) B(i, j) -= f * B(k, j);
        }
    }
    for (int64_t j = 0; j < B.cols; j++)
        for (int64_t i = p - 1; i >= 0; i--) {
            double s = B(i, j);
            for (int64_t l = i + 1; l < p; l++) s -= A(i, l) * B(l, j);
            B(i, j) = s / A(i, i);
        }
    return B;
}

Matrix identity(int64_t p) {
    Matrix I(p, p);
    for (int64_t i = 0; i < p; i++) I(i, i) = 1.0;
    return I;
}

Matrix inverse(const Matrix& A) { return solve_dense(A, identity(A.rows)); }

bool cholesky_ok(const Matrix& A) {
    const int64_t p = A.rows;
    Matrix L(p, p);
    for (int64_t j = 0; j < p; j++) {
        double d = A(j, j);
        for (int64_t k = 0; k < j; k++) d -= L(j, k) * L(j, k);
        if (!(d > 0.0)) return false;
        L(j, j) = std::sqrt(d);
        for (int64_t i = j + 1; i < p; i++) {
            double s = A(i, j);
            for (int64_t k = 0; k < j; k++) s -= L(i, k) * L(j, k);
            L(i, j) = s / L(j, j);
        }
    }
    return true;
}

// Moore-Penrose inverse of a symmetric matrix from its Jacobi eigen
// decomposition; singular values below 1e-15 of the largest count as zero
// (numpy.linalg.pinv's default cutoff).
Matrix pinv_sym(const Matrix& A) {
    const int64_t p = A.rows;
    Matrix a = A, v = identity(p);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int64_t i = 0; i < p; i++)
            for (int64_t j = i + 1; j < p; j++) off += a(i, j) * a(i, j);
        if (off < 1e-30) break;
        for (int64_t i = 0; i < p; i++)
            for (int64_t j = i + 1; j < p; j++) {
                if (a(i, j) == 0.0) continue;
                const double th = (a(j, j) - a(i, i)) / (2.0 * a(i, j));
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int64_t k = 0; k < p; k++) {
                    const double aki = a(k, i), akj = a(k, j);
                    a(k, i) = c * aki - s * akj;
                    a(k, j) = s * aki + c * akj;
                }
                for (int64_t k = 0; k < p; k++) {
                    const double aik = a(i, k), ajk = a(j, k);
                    a(i, k) = c * aik - s * ajk;
                    a(j, k) = s * aik + c * ajk;
                }
                for (int64_t k = 0; k < p; k++) {
                    const double vki = v(k, i), vkj = v(k, j);
                    v(k, i) = c * vki - s * vkj;
                    v(k, j) = s * vki + c * vkj;
                }
            }
    }
    double smax = 0.0;
    for (int64_t i = 0; i < p; i++) smax = std::max(smax, std::fabs(a(i, i)));
    Matrix out(p, p);
    for (int64_t k = 0; k < p; k++) {
        const double l = a(k, k);
        if (std::fabs(l) <= 1e-15 * smax) continue;
        for (int64_t i = 0; i < p; i++)
            for (int64_t j = 0; j < p; j++) out(i, j) += v(i, k) * v(j, k) / l;
    }
    return out;
}

// C = A^T B (A n x p, B n x q)
Matrix tmul(const Matrix& A, const Matrix& B) {
    Matrix C(A.cols, B.cols);
    for (int64_t i = 0; i < A.cols; i++)
        for (int64_t j = 0; j < B.cols; j++) {
            const double* a = A.col(i);
            const double* b = B.col(j);
            double s = 0.0;
            for (int64_t r = 0; r < A.rows; r++) s += a[r] * b[r];
            C(i, j) = s;
        }
    return C;
}

// y = A x (A n x p)
std::vector<double> mul(const Matrix& A, const std::vector<double>& x) {
    std::vector<double> y((size_t)A.rows, 0.0);
    for (int64_t j = 0; j < A.cols; j++) {
        const double* a = A.col(j);
        const double xj = x[(size_t)j];
        for (int64_t r = 0; r < A.rows; r++) y[(size_t)r] += a[r] * xj;
    }
    return y;
}

// y = A^T x
std::vector<double> tmul(const Matrix& A, const std::vector<double>& x) {
    std::vector<double> y((size_t)A.cols, 0.0);
    for (int64_t j = 0; j < A.cols; j++) {
        const double* a = A.col(j);
        double s = 0.0;
        for (int64_t r = 0; r < A.rows; r++) s += a[r] * x[(size_t)r];
        y[(size_t)j] = s;
    }
    return y;
}

// four interleaved partial sums, added pairwise at the end: independent
// add chains instead of one
double dot(const double* a, const double* b, int64_t n) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t i = 0;
    for (; i + 4 <= n; i += 4)
        for (int k = 0; k < 4; k++) s[k] += a[i + k] * b[i + k];
    for (; i < n; i++) s[0] += a[i] * b[i];
    return (s[0] + s[1]) + (s[2] + s[3]);
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    return dot(a.data(), b.data(), (int64_t)a.size());
}

// ----------------------------------------------------------------- families

// binomial(logit) or gaussian(identity).  mu_eta and link_both run over n
// samples with the exponentials taken as arrays (ex: n values of scratch):
// logit mu = 1 / (1 + exp(-eta)), mu_eta = e / (1 + e)^2 with e = exp(-|eta|).
// Beyond |eta| > 30 R's binomial() clamps (its logit_linkinv and
// logit_mu_eta): exp(eta) is taken as 1/DBL_EPSILON or DBL_EPSILON, so mu
// stays DBL_EPSILON/(1 + DBL_EPSILON) away from 0 and 1, and mu_eta is
// DBL_EPSILON; quasi-separated phenotypes reach it.
constexpr double kLogitThresh = 30.0;

struct Family {
    bool binomial;
    double variance(double mu) const { return binomial ? mu * (1.0 - mu) : 1.0; }
    void mu_eta(const double* eta, int64_t n, double* me, double* ex) const {
        if (!binomial) {
            std::fill(me, me + n, 1.0);
            return;
        }
        vmath::exp_neg_abs(eta, ex, n);
        for (int64_t i = 0; i < n; i++)
            me[i] = std::fabs(eta[i]) > kLogitThresh ? DBL_EPSILON : ex[i] / ((1.0 + ex[i]) * (1.0 + ex[i]));
    }
    // both from one exponential e = exp(-|eta|): logit mu = 1 / (1 + e) for
    // eta >= 0 and e / (1 + e) below (the same value, rounded once more or
    // less), mu_eta = e / (1 + e)^2
    void link_both(const double* eta, int64_t n, double* mu, double* me, double* ex) const {
        if (!binomial) {
            std::copy(eta, eta + n, mu);
            std::fill(me, me + n, 1.0);
            return;
        }
        vmath::exp_neg_abs(eta, ex, n);
        for (int64_t i = 0; i < n; i++) {
            if (std::fabs(eta[i]) > kLogitThresh) {
                mu[i] = (eta[i] > 0.0 ? 1.0 : DBL_EPSILON) / (1.0 + DBL_EPSILON);
                me[i] = DBL_EPSILON;
                continue;
            }
            const double e = ex[i], d = 1.0 + e;
            mu[i] = (eta[i] >= 0.0 ? 1.0 : e) / d;
            me[i] = e / (d * d);
        }
    }
};

// ------------------------------------------------------------- the solvers

struct Coefficients {
    std::vector<double> Sigma_iY, alpha, eta;
    Matrix Sigma_iX, cov;
};

// getCoefficients (FG.cpp:3167-3199): Sigma^-1 [Y, X] in one request
Coefficients get_coefficients(Solver& s, const std::vector<double>& Y, const Matrix& X,
                              const std::vector<double>& w, const double tau[2]) {
    const int64_t n = X.rows, p = X.cols;
    Matrix B(n, p + 1);
    std::copy(Y.begin(), Y.end(), B.col(0));
    std::copy(X.a.begin(), X.a.end(), B.col(1));
    Matrix sol;
    s.solve(w, tau, B, sol);
    Coefficients c;
    c.Sigma_iY.assign(sol.col(0), sol.col(0) + n);
    c.Sigma_iX = Matrix(n, p);
    std::copy(sol.col(1), sol.col(1) + n * p, c.Sigma_iX.a.begin());
    c.cov = Matrix(p, p);
    c.alpha.resize((size_t)p);
    c.eta.resize((size_t)n);
    coefficients_from_solution(n, p, Y.data(), X.a.data(), w.data(), tau[0], sol.a.data(), c.cov.a.data(),
                               c.alpha.data(), c.eta.data());
    return c;
}

struct CoefState {
    std::vector<double> Y, alpha, eta, W, mu, Sigma_iY;
    Matrix cov, Sigma_iX;
};

// Get_Coef (FG.R:2-35)
CoefState get_coef(Solver& s, const std::vector<double>& y, const Matrix& X, const double tau[2], const Family& fam,
                   std::vector<double> alpha0, const std::vector<double>& eta0, const std::vector<double>& offset,
                   const Options& opt) {
    const double tol_coef = 0.1;
    const int64_t n = X.rows;
    std::vector<double> Y((size_t)n), W((size_t)n), mu((size_t)n), eta;
    const Trait trait = fam.binomial ? Trait::Binary : Trait::Quantitative;
    auto working = [&](const std::vector<double>& e) {
        glm_working(trait, n, y.data(), e.data(), offset.data(), mu.data(), Y.data(), W.data());
    };
    working(eta0);
    Coefficients rc;
    std::vector<double> alpha;
    for (int it = 0; it < opt.maxiter; it++) {
        rc = get_coefficients(s, Y, X, W, tau);
        alpha = rc.alpha;
        eta = rc.eta;
        for (int64_t i = 0; i < n; i++) eta[(size_t)i] += offset[(size_t)i];
        working(eta);
        double m = 0.0;
        for (size_t j = 0; j < alpha.size(); j++)
            m = std::max(m, std::fabs(alpha[j] - alpha0[j]) / (std::fabs(alpha[j]) + std::fabs(alpha0[j]) + tol_coef));
        if (m < tol_coef) break;
        alpha0 = alpha;
    }
    CoefState st;
    st.Y = Y;
    st.alpha = alpha;
    st.eta = eta;
    st.W = W;
    st.mu = mu;
    st.cov = rc.cov;
    st.Sigma_iY = rc.Sigma_iY;
    st.Sigma_iX = rc.Sigma_iX;
    return st;
}

// P v = Sigma^-1 v - Sigma^-1 X cov (Sigma^-1 X)^T v
std::vector<double> proj(const std::vector<double>& v, const double* Sigma_iv, const Matrix& Sigma_iX,
                         const Matrix& cov) {
    const std::vector<double> t = mul(Sigma_iX, mul(cov, tmul(Sigma_iX, v)));
    std::vector<double> out(t.size());
    for (size_t i = 0; i < t.size(); i++) out[i] = Sigma_iv[i] - t[i];
    return out;
}

// b^T cov c for p-vectors
double quad(const std::vector<double>& b, const Matrix& cov, const std::vector<double>& c) {
    double s = 0.0;
    for (int64_t i = 0; i < cov.rows; i++)
        for (int64_t j = 0; j < cov.cols; j++) s += b[(size_t)i] * cov(i, j) * c[(size_t)j];
    return s;
}

// Probe vectors are the same at every GetTrace call (FG.cpp:3114), so they
// are kept process-wide per (seed, probe index, length), up to 1 GiB.
const std::vector<double>& cached_probe(uint64_t seed, int64_t i, int64_t n) {
    static std::mutex mu;
    static std::map<std::tuple<uint64_t, int64_t, int64_t>, std::shared_ptr<std::vector<double>>> cache;
    static int64_t bytes = 0;
    thread_local std::vector<double> scratch;
    const auto key = std::make_tuple(seed, i, n);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return *it->second;
    }
    auto v = std::make_shared<std::vector<double>>((size_t)n);
    probe_vector(seed, i, n, v->data());
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return *it->second;
    if (bytes + 8 * n > (int64_t(1) << 30)) {
        scratch.swap(*v);
        return scratch;
    }
    bytes += 8 * n;
    return *cache.emplace(key, v).first->second;
}

struct AIScore {
    double YPAPY = 0.0, YPA0PY = 0.0;
    double trace[2] = {0.0, 0.0};   // binary: trace[1] only
    double AI[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    int nprobe = 0;
};

// getAIScore (FG.cpp:3279-3295) / getAIScore_q (FG.cpp:3479-3531) with
// GetTrace / GetTrace_q (FG.cpp:3113-3160, 3409-3472).  Each round of probes
// U is one product K [PY, U] (K U after the first round) and one solve
// Sigma^-1 [U, APY, (PY)]; the probe terms are t_i = (K u_i)^T P u_i and, for
// the quantitative trait, u_i^T P u_i.  A round that leaves the coefficient
// of variation above the cutoff adds ten probes (FG.cpp:3147-3154).
AIScore ai_score(Solver& s, const std::vector<double>& Y, const Matrix& X, const std::vector<double>& w,
                 const double tau[2], const std::vector<double>& Sigma_iY, const Matrix& Sigma_iX, Matrix cov,
                 const Options& opt, bool quantitative) {
    const int64_t n = X.rows, p = X.cols;
    if (quantitative) cov = inv_sympd(tmul(X, Sigma_iX));   // getAIScore_q recomputes cov (FG.cpp:3487-3494)
    std::vector<double> PY = proj(Y, Sigma_iY.data(), Sigma_iX, cov);
    std::vector<double> t1, t0, APY, SAPY, SPY;
    int64_t start = 0, end = opt.nrun;
    for (;;) {
        const int64_t k = end - start;
        const bool first = APY.empty();
        TraceTerms tt;
        s.trace_round(w, tau, Sigma_iX, first ? &PY : nullptr, quantitative, start, end, opt.probe_seed, tt);
        if (first) {
            APY = std::move(tt.APY);
            SAPY = std::move(tt.SAPY);
            if (quantitative) SPY = std::move(tt.SPY);
        }
        std::vector<double> h((size_t)p), g((size_t)p);
        for (int64_t j = 0; j < k; j++) {
            for (int64_t c = 0; c < p; c++) {
                h[(size_t)c] = tt.h(j, c);
                g[(size_t)c] = tt.g(j, c);
            }
            t1.push_back(tt.d[(size_t)j] - quad(h, cov, g));
            if (quantitative) t0.push_back(tt.e[(size_t)j] - quad(g, cov, g));
        }
        const double cv = cal_cv(t1);
        const double cv0 = quantitative ? cal_cv(t0) : 0.0;
        if (cv > opt.trace_cv_cutoff || (quantitative && cv0 > opt.trace_cv_cutoff)) {
            start = end;
            end = end + 10;
            continue;
        }
        break;
    }
    (void)n;
    AIScore r;
    r.nprobe = (int)t1.size();
    r.YPAPY = dot(PY, APY);
    const std::vector<double> PAPY = proj(APY, SAPY.data(), Sigma_iX, cov);
    const double mean1 = std::accumulate(t1.begin(), t1.end(), 0.0) / (double)t1.size();
    if (!quantitative) {
        r.trace[1] = mean1;
        r.AI[1][1] = dot(APY, PAPY);
    } else {
        const std::vector<double> PA0PY = proj(PY, SPY.data(), Sigma_iX, cov);
        r.YPA0PY = dot(PY, PY);
        r.trace[0] = std::accumulate(t0.begin(), t0.end(), 0.0) / (double)t0.size();
        r.trace[1] = mean1;
        r.AI[0][0] = dot(PY, PA0PY);
        r.AI[0][1] = r.AI[1][0] = dot(PY, PAPY);
        r.AI[1][1] = dot(APY, PAPY);
    }
    return r;
}

// fitglmmaiRPCG (FG.cpp:3302-3341)
void fitglmmai_rpcg(Solver& s, const CoefState& rc, const Matrix& X, double tau[2], const Options& opt) {
    const AIScore re = ai_score(s, rc.Y, X, rc.W, tau, rc.Sigma_iY, rc.Sigma_iX, rc.cov, opt, false);
    const double Dtau = (re.YPAPY - re.trace[1]) / re.AI[1][1];
    const double t0[2] = {tau[0], tau[1]};
    tau[1] = t0[1] + Dtau;
    for (int i = 0; i < 2; i++)
        if (tau[i] < opt.tol) tau[i] = 0.0;
    double step = 1.0;
    while (tau[1] < 0.0) {
        step *= 0.5;
        tau[1] = t0[1] + step * Dtau;
    }
    for (int i = 0; i < 2; i++)
        if (tau[i] < opt.tol) tau[i] = 0.0;
}

// fitglmmaiRPCG_q (FG.cpp:3610-3662): 2x2 AI system, GMMAT zeroing
void fitglmmai_rpcg_q(Solver& s, const CoefState& rc, const Matrix& X, double tau[2], const Options& opt) {
    const bool zero[2] = {tau[0] < opt.tol, tau[1] < opt.tol};
    const AIScore re = ai_score(s, rc.Y, X, rc.W, tau, rc.Sigma_iY, rc.Sigma_iX, rc.cov, opt, true);
    Matrix AI(2, 2), score(2, 1);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) AI(i, j) = re.AI[i][j];
    score(0, 0) = re.YPA0PY - re.trace[0];
    score(1, 0) = re.YPAPY - re.trace[1];
    const Matrix D = solve_dense(AI, score);
    const double t0[2] = {tau[0], tau[1]};
    auto apply = [&](double step) {
        for (int i = 0; i < 2; i++) {
            tau[i] = t0[i] + step * D(i, 0);
            if (zero[i] && tau[i] < opt.tol) tau[i] = 0.0;
        }
    };
    apply(1.0);
    double step = 1.0;
    while (tau[0] < 0.0 || tau[1] < 0.0) {
        step *= 0.5;
        apply(step);
    }
    for (int i = 0; i < 2; i++)
        if (tau[i] < opt.tol) tau[i] = 0.0;
}

double rel_change(const double a[2], const double b[2], double tol) {
    double m = 0.0;
    for (int i = 0; i < 2; i++) m = std::max(m, std::fabs(a[i] - b[i]) / (std::fabs(a[i]) + std::fabs(b[i]) + tol));
    return m;
}

}  // namespace

// ------------------------------------------------------------------ public

Matrix gather_columns(const Matrix& B, const Matrix& KB, const std::vector<int>& src) {
    Matrix C(B.rows, (int64_t)src.size());
    for (size_t c = 0; c < src.size(); c++) {
        const double* from = src[c] >= 0 ? B.col(src[c]) : KB.col(-1 - src[c]);
        std::copy(from, from + B.rows, C.col((int64_t)c));
    }
    return C;
}

void Solver::product_then_solve(const Matrix& B, Matrix& KB, const std::vector<double>& w, const double tau[2],
                                const std::vector<int>& src, Matrix& X) {
    product(B, KB);
    solve(w, tau, gather_columns(B, KB, src), X);
}

// The host trace round: K [PY, U] (K U after the first round), then
// Sigma^-1 [U, K PY, (PY)], and the dots in one pass over the samples.
void Solver::trace_round(const std::vector<double>& w, const double tau[2], const Matrix& Sigma_iX,
                         const std::vector<double>* PY, bool quantitative, int64_t start, int64_t end, uint64_t seed,
                         TraceTerms& out) {
    const int64_t n = Sigma_iX.rows, p = Sigma_iX.cols, k = end - start;
    const bool first = PY != nullptr;
    Matrix U(n, k);
    for (int64_t j = 0; j < k; j++) {
        const std::vector<double>& pv = cached_probe(seed, start + j, n);
        std::copy(pv.begin(), pv.end(), U.col(j));
    }
    Matrix KB, sol;
    std::vector<int> src;
    if (first) {
        Matrix B(n, k + 1);
        std::copy(PY->begin(), PY->end(), B.col(0));
        std::copy(U.a.begin(), U.a.end(), B.col(1));
        for (int64_t j = 0; j < k; j++) src.push_back((int)(1 + j));
        src.push_back(-1);                      // K PY
        if (quantitative) src.push_back(0);     // PY
        product_then_solve(B, KB, w, tau, src, sol);
    } else {
        for (int64_t j = 0; j < k; j++) src.push_back((int)j);
        product_then_solve(U, KB, w, tau, src, sol);
    }
    const int64_t u0 = first ? 1 : 0;   // first column of K U in KB
    if (first) {
        out.APY.assign(KB.col(0), KB.col(0) + n);
        out.SAPY.assign(sol.col(k), sol.col(k) + n);
        if (quantitative) out.SPY.assign(sol.col(k + 1), sol.col(k + 1) + n);
    }
    out.d.assign((size_t)k, 0.0);
    out.e.assign((size_t)k, 0.0);
    out.h = Matrix(k, p);
    out.g = Matrix(k, p);
    std::vector<double> h((size_t)p), g((size_t)p);
    for (int64_t j = 0; j < k; j++) {
        const double* au = KB.col(u0 + j);
        const double* su = sol.col(j);
        const double* u = U.col(j);
        double d = 0.0, e = 0.0;
        std::fill(h.begin(), h.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        for (int64_t r = 0; r < n; r++) {
            d += au[r] * su[r];
            e += u[r] * su[r];
            for (int64_t c = 0; c < p; c++) {
                const double sx = Sigma_iX.a[(size_t)(c * n + r)];
                h[(size_t)c] += au[r] * sx;
                g[(size_t)c] += u[r] * sx;
            }
        }
        out.d[(size_t)j] = d;
        out.e[(size_t)j] = e;
        for (int64_t c = 0; c < p; c++) {
            out.h(j, c) = h[(size_t)c];
            out.g(j, c) = g[(size_t)c];
        }
    }
}

void probe_vector(uint64_t seed, int64_t i, int64_t n, double* out) {
    auto mix = [](uint64_t z) {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const uint64_t ic = (uint64_t)i * 0x9E3779B97F4A7C15ull;
    for (int64_t r = 0; r < n; r++) {
        const uint64_t inner = mix((uint64_t)r * 0xC2B2AE3D27D4EB4Full + 0x5851F42Dull);
        const uint64_t h = mix(seed ^ mix(ic + inner));
        out[r] = (h >> 63) ? -1.0 : 1.0;
    }
}

double cal_cv(const std::vector<double>& x) {
    const double n = (double)x.size();
    const double mean = std::accumulate(x.begin(), x.end(), 0.0) / n;
    double ss = 0.0;
    for (double v : x) ss += (v - mean) * (v - mean);
    const double sd = std::sqrt(ss / (n - 1.0));
    return (sd / mean) / n;
}

Matrix inv_sympd(const Matrix& A0) {
    Matrix A = A0;
    for (int64_t i = 0; i < A.rows; i++)
        for (int64_t j = 0; j < i; j++) A(i, j) = A(j, i);   // symmatu
    if (cholesky_ok(A)) return inverse(A);
    return pinv_sym(A);
}

namespace {

// A += X^T W X, b += X^T W z over rows i0 + [0, len) (X's p columns xc; wt
// and z hold those rows' values), lower triangle of A only; P > 0 is p fixed
// at compile time, so the sums stay in registers
template <int P>
void normal_equations(const double* const* xc, int64_t p, int64_t i0, int64_t len, const double* wt,
                      const double* z, Matrix& A, Matrix& b) {
    if (P > 0) {
        constexpr int Q = P > 0 ? P : 1;
        double a[Q][Q] = {}, bb[Q] = {};
        const double* x[Q];
        for (int j = 0; j < Q; j++) x[j] = xc[j] + i0;
        for (int64_t i = 0; i < len; i++) {
            double xw[Q];
            for (int j = 0; j < Q; j++) xw[j] = x[j][i] * wt[i];
            for (int j = 0; j < Q; j++) {
                bb[j] += xw[j] * z[i];
                for (int k = 0; k <= j; k++) a[j][k] += xw[j] * x[k][i];
            }
        }
        for (int j = 0; j < Q; j++) {
            b.a[(size_t)j] += bb[j];
            for (int k = 0; k <= j; k++) A(j, k) += a[j][k];
        }
        return;
    }
    std::vector<double> xw((size_t)p);
    for (int64_t i = 0; i < len; i++) {
        for (int64_t j = 0; j < p; j++) xw[(size_t)j] = xc[j][i0 + i] * wt[i];
        for (int64_t j = 0; j < p; j++) {
            b.a[(size_t)j] += xw[(size_t)j] * z[i];
            for (int64_t k = 0; k <= j; k++) A(j, k) += xw[(size_t)j] * xc[k][i0 + i];
        }
    }
}

}  // namespace

void glm_working(Trait trait, int64_t n, const double* y, const double* eta, const double* offset, double* mu,
                 double* Y, double* W) {
    const Family fam{trait == Trait::Binary};
    for_pieces(n, [&](int64_t i0, int64_t i1, int) {
        // W is the exponentials' scratch and Y holds mu_eta until it is read
        fam.link_both(eta + i0, i1 - i0, mu + i0, Y + i0, W + i0);
        for (int64_t i = i0; i < i1; i++) {
            const double m = mu[i], me = Y[i];
            Y[i] = eta[i] - (offset ? offset[i] : 0.0) + (y[i] - m) / me;
            const double r = me / std::sqrt(fam.variance(m));
            W[i] = r * r;
        }
    });
}

namespace {

// A = X^T S (p x p) and b = S^T Y (p) over n rows (X, S column-major with p
// columns) in one pass; P > 0 is p fixed at compile time, so the sums stay in
// registers, else one dot per entry
template <int P>
void cross_products(const double* X, const double* S, const double* Y, int64_t n, int64_t p, int64_t i0, int64_t i1,
                    Matrix& A, std::vector<double>& b) {
    if (P > 0) {
        constexpr int Q = P > 0 ? P : 1;
        double a[Q][Q] = {}, bb[Q] = {};
        for (int64_t i = i0; i < i1; i++) {
            double x[Q], sv[Q];
            for (int j = 0; j < Q; j++) {
                x[j] = X[j * n + i];
                sv[j] = S[j * n + i];
            }
            for (int j = 0; j < Q; j++) {
                bb[j] += sv[j] * Y[i];
                for (int k = 0; k < Q; k++) a[j][k] += x[j] * sv[k];
            }
        }
        for (int j = 0; j < Q; j++) {
            b[(size_t)j] = bb[j];
            for (int k = 0; k < Q; k++) A(j, k) = a[j][k];
        }
        return;
    }
    for (int64_t j = 0; j < p; j++) {
        b[(size_t)j] = dot(S + j * n + i0, Y + i0, i1 - i0);
        for (int64_t k = 0; k < p; k++) A(j, k) = dot(X + j * n + i0, S + k * n + i0, i1 - i0);
    }
}

}  // namespace

void coefficients_from_solution(int64_t n, int64_t p, const double* Y, const double* X, const double* w, double tau0,
                                const double* sol, double* cov, double* alpha, double* eta) {
    const double* siY = sol;
    const double* siX = sol + n;
    // X^T Sigma^-1 X and (Sigma^-1 X)^T Y, per piece of the samples, the
    // pieces' sums added in order
    const int P = pieces_for(n);
    std::vector<Matrix> Ap((size_t)P, Matrix(p, p));
    std::vector<std::vector<double>> syp((size_t)P, std::vector<double>((size_t)p));
    for_pieces(n, [&](int64_t i0, int64_t i1, int c) {
        Matrix& Ac = Ap[(size_t)c];
        std::vector<double>& sc = syp[(size_t)c];
        switch (p) {
            case 1: cross_products<1>(X, siX, Y, n, p, i0, i1, Ac, sc); break;
            case 2: cross_products<2>(X, siX, Y, n, p, i0, i1, Ac, sc); break;
            case 3: cross_products<3>(X, siX, Y, n, p, i0, i1, Ac, sc); break;
            case 4: cross_products<4>(X, siX, Y, n, p, i0, i1, Ac, sc); break;
            default: cross_products<0>(X, siX, Y, n, p, i0, i1, Ac, sc);
        }
    });
    Matrix A = Ap[0];
    std::vector<double> sy = syp[0];
    for (int c = 1; c < P; c++) {
        for (size_t k = 0; k < A.a.size(); k++) A.a[k] += Ap[(size_t)c].a[k];
        for (int64_t j = 0; j < p; j++) sy[(size_t)j] += syp[(size_t)c][(size_t)j];
    }
    const Matrix C = inv_sympd(A);
    for (int64_t j = 0; j < p; j++) {
        double a = 0.0;
        for (int64_t k = 0; k < p; k++) a += C(j, k) * sy[(size_t)k];
        alpha[j] = a;
    }
    std::copy(C.a.begin(), C.a.end(), cov);
    for_pieces(n, [&](int64_t i0, int64_t i1, int) {
        for (int64_t i = i0; i < i1; i++) {
            double sxa = 0.0;
            for (int64_t k = 0; k < p; k++) sxa += siX[k * n + i] * alpha[k];
            eta[i] = Y[i] - tau0 * (siY[i] - sxa) / w[i];
        }
    });
}

GlmStart glm_start(Trait trait, const std::vector<double>& y, const Matrix& X, const std::vector<double>& offset,
                   int maxit, double epsilon) {
    const int64_t n = X.rows, p = X.cols;
    const Family fam{trait == Trait::Binary};
    const double* off = offset.empty() ? nullptr : offset.data();
    std::vector<double> mu((size_t)n), eta((size_t)n);
    bool binary01 = fam.binomial;
    for (int64_t i = 0; i < n; i++) {
        if (fam.binomial) {
            mu[(size_t)i] = (y[(size_t)i] + 0.5) / 2.0;
            eta[(size_t)i] = mu[(size_t)i] / (1.0 - mu[(size_t)i]);
            if (y[(size_t)i] != 0.0 && y[(size_t)i] != 1.0) binary01 = false;
        } else {
            mu[(size_t)i] = eta[(size_t)i] = y[(size_t)i];
        }
    }
    if (fam.binomial) vmath::log(eta.data(), eta.data(), n);
    // the samples go in pieces (one per thread) and, within a piece, in tiles
    // of T, so each pass's temporaries stay in cache; me_all holds mu_eta at
    // the current eta, made with mu from one exponential
    constexpr int64_t T = 2048;
    const int P = pieces_for(n);
    std::vector<std::vector<double>> tmp((size_t)P, std::vector<double>((size_t)(3 * T)));
    std::vector<double> me_all((size_t)n);
    for_pieces(n, [&](int64_t p0, int64_t p1, int c) {
        double* ex = tmp[(size_t)c].data();
        for (int64_t i0 = p0; i0 < p1; i0 += T)
            fam.mu_eta(eta.data() + i0, std::min(T, p1 - i0), me_all.data() + i0, ex);
    });
    // per piece: four interleaved partial sums (binary y) or one, then the
    // pieces in order
    auto deviance = [&](const double* m) {
        std::vector<double> part((size_t)P, 0.0);
        for_pieces(n, [&](int64_t p0, int64_t p1, int c) {
            double* ex = tmp[(size_t)c].data();
            if (binary01) {
                // y in {0, 1}: the unit deviance is -2 log(mu) or -2 log(1 - mu);
                // y m + (1 - y)(1 - m) is m or 1 - m exactly
                double d4[4] = {0.0, 0.0, 0.0, 0.0};
                for (int64_t i0 = p0; i0 < p1; i0 += T) {
                    const int64_t len = std::min(T, p1 - i0);
                    for (int64_t k = 0; k < len; k++) {
                        const double yi = y[(size_t)(i0 + k)], mi = m[i0 + k];
                        ex[k] = yi * mi + (1.0 - yi) * (1.0 - mi);
                    }
                    vmath::log(ex, ex, len);
                    int64_t k = 0;
                    for (; k + 4 <= len; k += 4)
                        for (int q = 0; q < 4; q++) d4[q] += ex[k + q];
                    for (; k < len; k++) d4[0] += ex[k];
                }
                part[(size_t)c] = (d4[0] + d4[1]) + (d4[2] + d4[3]);
                return;
            }
            double d = 0.0;
            for (int64_t i = p0; i < p1; i++) {
                const double yi = y[(size_t)i], mi = m[i];
                if (fam.binomial) {
                    const double a = yi > 0 ? yi * std::log(yi / mi) : 0.0;
                    const double b = yi < 1 ? (1 - yi) * std::log((1 - yi) / (1 - mi)) : 0.0;
                    d += a + b;
                } else {
                    d += (yi - mi) * (yi - mi);
                }
            }
            part[(size_t)c] = d;
        });
        double d = 0.0;
        for (int c = 0; c < P; c++) d += part[(size_t)c];
        if (binary01) return -2.0 * d;
        return fam.binomial ? 2.0 * d : d;
    };
    double dev_old = deviance(mu.data());
    std::vector<double> beta((size_t)p, 0.0);
    std::vector<const double*> xc((size_t)p);
    for (int64_t j = 0; j < p; j++) xc[(size_t)j] = X.col(j);
    for (int it = 0; it < maxit; it++) {
        // per tile: the working response z and weights wt, then the tile's
        // share of the weighted normal equations X^T W X beta = X^T W z; the
        // pieces' sums added in order
        std::vector<Matrix> Ap((size_t)P, Matrix(p, p)), bp((size_t)P, Matrix(p, 1));
        for_pieces(n, [&](int64_t p0, int64_t p1, int c) {
            double* z = tmp[(size_t)c].data() + T;
            double* wt = z + T;
            Matrix& A = Ap[(size_t)c];
            Matrix& b = bp[(size_t)c];
            for (int64_t i0 = p0; i0 < p1; i0 += T) {
                const int64_t len = std::min(T, p1 - i0);
                const double* me = me_all.data() + i0;
                for (int64_t k = 0; k < len; k++) {
                    const int64_t i = i0 + k;
                    const double mi = mu[(size_t)i];
                    z[k] = (eta[(size_t)i] - (off ? off[i] : 0.0)) + (y[(size_t)i] - mi) / me[k];
                    wt[k] = me[k] * me[k] / fam.variance(mi);
                }
                switch (p) {
                    case 1: normal_equations<1>(xc.data(), p, i0, len, wt, z, A, b); break;
                    case 2: normal_equations<2>(xc.data(), p, i0, len, wt, z, A, b); break;
                    case 3: normal_equations<3>(xc.data(), p, i0, len, wt, z, A, b); break;
                    case 4: normal_equations<4>(xc.data(), p, i0, len, wt, z, A, b); break;
                    default: normal_equations<0>(xc.data(), p, i0, len, wt, z, A, b);
                }
            }
        });
        Matrix A = Ap[0], b = bp[0];
        for (int c = 1; c < P; c++) {
            for (size_t k = 0; k < A.a.size(); k++) A.a[k] += Ap[(size_t)c].a[k];
            for (size_t k = 0; k < b.a.size(); k++) b.a[k] += bp[(size_t)c].a[k];
        }
        for (int64_t j = 0; j < p; j++)
            for (int64_t k = 0; k < j; k++) A(k, j) = A(j, k);
        const Matrix sol = solve_dense(A, b);
        beta.assign(sol.a.begin(), sol.a.end());
        // eta = X beta + offset and mu, per tile
        for_pieces(n, [&](int64_t p0, int64_t p1, int c) {
            double* ex = tmp[(size_t)c].data();
            for (int64_t i0 = p0; i0 < p1; i0 += T) {
                const int64_t len = std::min(T, p1 - i0);
                for (int64_t i = i0; i < i0 + len; i++) {
                    double e = off ? off[i] : 0.0;
                    for (int64_t j = 0; j < p; j++) e += xc[(size_t)j][i] * beta[(size_t)j];
                    eta[(size_t)i] = e;
                }
                fam.link_both(eta.data() + i0, len, mu.data() + i0, me_all.data() + i0, ex);
            }
        });
        const double dev = deviance(mu.data());
        if (std::fabs(dev - dev_old) / (std::fabs(dev) + 0.1) < epsilon) break;
        dev_old = dev;
    }
    return {beta, eta, mu};
}

Result fit_null(Solver& s, Trait trait, const std::vector<double>& y, const Matrix& X,
                const std::vector<double>& offset_in, const Options& opt) {
    const int64_t n = X.rows;
    if ((int64_t)y.size() != n) throw std::invalid_argument("y and X differ in length");
    const std::vector<double> offset = offset_in.empty() ? std::vector<double>((size_t)n, 0.0) : offset_in;
    const bool quant = trait == Trait::Quantitative;
    const Family fam{!quant};
    const GlmStart fit0 = glm_start(trait, y, X, offset);
    std::vector<double> eta = fit0.eta, alpha = fit0.coef;
    Result res;
    // starting values (FG.R:148-157, 389-398)
    double tau[2] = {1.0, quant ? 0.0 : 0.1};
    double tau0[2] = {tau[0], tau[1]};
    CoefState rc = get_coef(s, y, X, tau, fam, alpha, eta, offset, opt);
    const AIScore re = ai_score(s, rc.Y, X, rc.W, tau, rc.Sigma_iY, rc.Sigma_iX, rc.cov, opt, quant);
    tau[1] = std::max(0.0, tau0[1] + tau0[1] * tau0[1] * (re.YPAPY - re.trace[1]) / (double)n);   // FG.R:165, 405
    if (quant) tau[0] = std::max(0.0, tau0[0] + tau0[0] * tau0[0] * (re.YPA0PY - re.trace[0]) / (double)n);
    res.theta_history.insert(res.theta_history.end(), tau, tau + 2);
    int it = 0;
    for (it = 1; it <= opt.maxiter; it++) {
        const std::vector<double> alpha0 = quant ? alpha : rc.alpha;
        tau0[0] = tau[0];
        tau0[1] = tau[1];
        rc = get_coef(s, y, X, tau, fam, alpha0, eta, offset, opt);
        if (quant)
            fitglmmai_rpcg_q(s, rc, X, tau, opt);
        else
            fitglmmai_rpcg(s, rc, X, tau, opt);
        alpha = rc.alpha;
        eta = rc.eta;
        res.theta_history.insert(res.theta_history.end(), tau, tau + 2);
        if (quant && tau[0] <= 0)
            throw std::runtime_error("ERROR! The first variance component parameter estimate is 0");   // FG.R:449-451
        if (tau[1] <= 0) break;                                                    // FG.R:200-206, 452-454
        if (rel_change(tau, tau0, opt.tol) < opt.tol) break;                       // FG.R:206, 456
        if (std::max(tau[0], tau[1]) > std::pow(opt.tol, -2)) {
            it = opt.maxiter;
            break;
        }
    }
    if (it > opt.maxiter) it = opt.maxiter;   // the loop ran out
    // final coefficients at the estimate (FG.R:218, 475)
    rc = get_coef(s, y, X, tau, fam, quant ? alpha : rc.alpha, eta, offset, opt);
    res.theta[0] = tau[0];
    res.theta[1] = tau[1];
    res.coefficients = rc.alpha;
    res.linear_predictors = rc.eta;
    res.fitted_values = rc.mu;
    res.working_y = rc.Y;
    res.cov = rc.cov;
    res.converged = it < opt.maxiter;
    res.iterations = it;
    return res;
}

CovariateTransform covariate_transform(const Matrix& X0) {
    auto qr = [](const Matrix& A, Matrix& Q, Matrix& R) {
        const int64_t n = A.rows, p = A.cols;
        Matrix a = A;
        std::vector<double> taus((size_t)p, 0.0);
        for (int64_t j = 0; j < p && j < n; j++) {
            double xnorm = 0.0;
            for (int64_t i = j + 1; i < n; i++) xnorm += a(i, j) * a(i, j);
            xnorm = std::sqrt(xnorm);
            const double alpha = a(j, j);
            if (xnorm == 0.0) {
                taus[(size_t)j] = 0.0;
                continue;
            }
            const double beta = -std::copysign(std::hypot(alpha, xnorm), alpha);
            taus[(size_t)j] = (beta - alpha) / beta;
            const double scale = 1.0 / (alpha - beta);
            for (int64_t i = j + 1; i < n; i++) a(i, j) *= scale;
            a(j, j) = beta;
            // apply H = I - tau v v^T (v = [1, a(j+1:, j)]) to the columns right of j
            for (int64_t c = j + 1; c < p; c++) {
                double s = a(j, c);
                for (int64_t i = j + 1; i < n; i++) s += a(i, j) * a(i, c);
                s *= taus[(size_t)j];
                a(j, c) -= s;
                for (int64_t i = j + 1; i < n; i++) a(i, c) -= s * a(i, j);
            }
        }
        R = Matrix(p, p);
        for (int64_t j = 0; j < p; j++)
            for (int64_t i = 0; i <= j; i++) R(i, j) = a(i, j);
        // Q = H_1 ... H_p applied to the first p columns of the identity
        Q = Matrix(n, p);
        for (int64_t j = 0; j < p; j++) Q(j, j) = 1.0;
        for (int64_t j = p - 1; j >= 0; j--) {
            const double t = taus[(size_t)j];
            if (t == 0.0) continue;
            for (int64_t c = 0; c < p; c++) {
                double s = Q(j, c);
                for (int64_t i = j + 1; i < n; i++) s += a(i, j) * Q(i, c);
                s *= t;
                Q(j, c) -= s;
                for (int64_t i = j + 1; i < n; i++) Q(i, c) -= s * a(i, j);
            }
        }
    };
    CovariateTransform out;
    Matrix Q, R;
    qr(X0, Q, R);
    double dmax = 1.0;
    for (int64_t j = 0; j < R.rows; j++) dmax = std::max(dmax, std::fabs(R(j, j)));
    std::vector<int> keep;
    for (int64_t j = 0; j < R.rows; j++)
        if (std::fabs(R(j, j)) > 1e-10 * dmax) keep.push_back((int)j);
    if ((int64_t)keep.size() < X0.cols) {
        Matrix Xk(X0.rows, (int64_t)keep.size());
        for (size_t c = 0; c < keep.size(); c++) std::copy(X0.col(keep[c]), X0.col(keep[c]) + X0.rows, Xk.col((int64_t)c));
        qr(Xk, Q, R);
    }
    const double sq = std::sqrt((double)X0.rows);
    for (double& v : Q.a) v *= sq;
    out.X = Q;
    out.R = R;
    out.kept = keep;
    return out;
}

ScoreNull score_test_null(const std::vector<double>& mu, const std::vector<double>& V, const std::vector<double>& y,
                          const Matrix& X) {
    const int64_t n = X.rows, p = X.cols;
    ScoreNull o;
    o.XV = Matrix(p, n);
    Matrix XVt(n, p);
    for (int64_t j = 0; j < p; j++)
        for (int64_t i = 0; i < n; i++) {
            const double v = X(i, j) * V[(size_t)i];
            o.XV(j, i) = v;
            XVt(i, j) = v;
        }
    o.XVX = tmul(X, XVt);
    o.XVX_inv = inverse(o.XVX);
    o.XXVX_inv = Matrix(n, p);
    for (int64_t j = 0; j < p; j++)
        for (int64_t i = 0; i < n; i++) {
            double s = 0.0;
            for (int64_t k = 0; k < p; k++) s += X(i, k) * o.XVX_inv(k, j);
            o.XXVX_inv(i, j) = s;
        }
    o.XVX_inv_XV = o.XXVX_inv;
    for (int64_t j = 0; j < p; j++)
        for (int64_t i = 0; i < n; i++) o.XVX_inv_XV(i, j) *= V[(size_t)i];
    o.S_a.assign((size_t)p, 0.0);
    for (int64_t j = 0; j < p; j++)
        for (int64_t i = 0; i < n; i++) o.S_a[(size_t)j] += X(i, j) * (y[(size_t)i] - mu[(size_t)i]);
    return o;
}

}  // namespace fit
}  // namespace saige
