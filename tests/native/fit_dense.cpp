// Test harness only: the C++ null-fit driver (saige-gpu_amd/csrc/fit_core.cpp)
// run against a dense FP64 Solver on the CPU (exact solves of
// Sigma = tau0 diag(1/w) + tau1 K by Cholesky), exported for ctypes so the
// tests can compare it with the Python driver on the same requests.
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "fit_core.hpp"

using saige::fit::Matrix;

namespace {

class DenseSolver : public saige::fit::Solver {
public:
    DenseSolver(const double* K, int64_t n) : K_(K), n_(n) {}
    void solve(const std::vector<double>& w, const double tau[2], const Matrix& B, Matrix& X) override {
        const int64_t n = n_;
        std::vector<double> L((size_t)(n * n));
        for (int64_t j = 0; j < n; j++)
            for (int64_t i = 0; i < n; i++)
                L[(size_t)(j * n + i)] = tau[1] * K_[j * n + i] + (i == j ? tau[0] / w[(size_t)i] : 0.0);
        // Cholesky, lower triangle in place (column-major)
        for (int64_t j = 0; j < n; j++) {
            double d = L[(size_t)(j * n + j)];
            for (int64_t k = 0; k < j; k++) d -= L[(size_t)(k * n + j)] * L[(size_t)(k * n + j)];
            if (!(d > 0)) throw std::runtime_error("Sigma not positive definite");
            d = std::sqrt(d);
            L[(size_t)(j * n + j)] = d;
            for (int64_t i = j + 1; i < n; i++) {
                double s = L[(size_t)(j * n + i)];
                for (int64_t k = 0; k < j; k++) s -= L[(size_t)(k * n + i)] * L[(size_t)(k * n + j)];
                L[(size_t)(j * n + i)] = s / d;
            }
        }
        X = Matrix(n, B.cols);
        for (int64_t c = 0; c < B.cols; c++) {
            std::vector<double> y(B.col(c), B.col(c) + n);
            for (int64_t i = 0; i < n; i++) {
                double s = y[(size_t)i];
                for (int64_t k = 0; k < i; k++) s -= L[(size_t)(k * n + i)] * y[(size_t)k];
                y[(size_t)i] = s / L[(size_t)(i * n + i)];
            }
            for (int64_t i = n - 1; i >= 0; i--) {
                double s = y[(size_t)i];
                for (int64_t k = i + 1; k < n; k++) s -= L[(size_t)(i * n + k)] * y[(size_t)k];
                y[(size_t)i] = s / L[(size_t)(i * n + i)];
            }
            std::memcpy(X.col(c), y.data(), sizeof(double) * (size_t)n);
        }
    }
    void product(const Matrix& B, Matrix& KB) override {
        const int64_t n = n_;
        KB = Matrix(n, B.cols);
        for (int64_t c = 0; c < B.cols; c++)
            for (int64_t j = 0; j < n; j++) {
                const double b = B(j, c);
                for (int64_t i = 0; i < n; i++) KB(i, c) += K_[j * n + i] * b;
            }
    }

private:
    const double* K_;
    int64_t n_;
};

}  // namespace

extern "C" int fitnull_dense(int trait, int64_t n, int p, const double* K, const double* y, const double* X,
                             int maxiter, double tol, int nrun, double cv_cutoff, uint64_t seed, double* theta,
                             double* coef, double* eta, double* mu, double* Yw, double* cov, int* iters,
                             int* converged, double* history, int* nhist) {
    try {
        DenseSolver s(K, n);
        saige::fit::Options opt;
        opt.maxiter = maxiter;
        opt.tol = tol;
        opt.nrun = nrun;
        opt.trace_cv_cutoff = cv_cutoff;
        opt.probe_seed = seed;
        Matrix Xm(n, p);
        std::memcpy(Xm.a.data(), X, sizeof(double) * (size_t)(n * p));
        const std::vector<double> yv(y, y + n);
        const auto r = saige::fit::fit_null(s, trait == 0 ? saige::fit::Trait::Binary : saige::fit::Trait::Quantitative,
                                            yv, Xm, {}, opt);
        theta[0] = r.theta[0];
        theta[1] = r.theta[1];
        std::memcpy(coef, r.coefficients.data(), sizeof(double) * (size_t)p);
        std::memcpy(eta, r.linear_predictors.data(), sizeof(double) * (size_t)n);
        std::memcpy(mu, r.fitted_values.data(), sizeof(double) * (size_t)n);
        std::memcpy(Yw, r.working_y.data(), sizeof(double) * (size_t)n);
        std::memcpy(cov, r.cov.a.data(), sizeof(double) * (size_t)(p * p));
        *iters = r.iterations;
        *converged = r.converged;
        *nhist = (int)r.theta_history.size() / 2;
        std::memcpy(history, r.theta_history.data(), sizeof(double) * r.theta_history.size());
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

extern "C" void probes(uint64_t seed, int64_t i, int64_t n, double* out) { saige::fit::probe_vector(seed, i, n, out); }

extern "C" double calcv(const double* x, int64_t n) { return saige::fit::cal_cv(std::vector<double>(x, x + n)); }

extern "C" int covtrans(const double* X, int64_t n, int p, double* Xout, double* R, int* kept, int* pk) {
    Matrix Xm(n, p);
    std::memcpy(Xm.a.data(), X, sizeof(double) * (size_t)(n * p));
    const auto t = saige::fit::covariate_transform(Xm);
    *pk = (int)t.kept.size();
    std::memcpy(Xout, t.X.a.data(), sizeof(double) * t.X.a.size());
    std::memcpy(R, t.R.a.data(), sizeof(double) * t.R.a.size());
    for (size_t i = 0; i < t.kept.size(); i++) kept[i] = t.kept[i];
    return 0;
}

extern "C" int scorenull(const double* mu, const double* V, const double* y, const double* X, int64_t n, int p,
                         double* XVX, double* XVX_inv, double* S_a, double* XVX_inv_XV) {
    Matrix Xm(n, p);
    std::memcpy(Xm.a.data(), X, sizeof(double) * (size_t)(n * p));
    const auto o = saige::fit::score_test_null(std::vector<double>(mu, mu + n), std::vector<double>(V, V + n),
                                               std::vector<double>(y, y + n), Xm);
    std::memcpy(XVX, o.XVX.a.data(), sizeof(double) * (size_t)(p * p));
    std::memcpy(XVX_inv, o.XVX_inv.a.data(), sizeof(double) * (size_t)(p * p));
    std::memcpy(S_a, o.S_a.data(), sizeof(double) * (size_t)p);
    std::memcpy(XVX_inv_XV, o.XVX_inv_XV.a.data(), sizeof(double) * (size_t)(n * p));
    return 0;
}
