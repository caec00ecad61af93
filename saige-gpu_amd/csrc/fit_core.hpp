// Null-GLMM fitting logic of SAIGE step 1 in C++ (host only, no HIP): the
// R drivers glmmkin.ai_PCG_Rcpp_Binary / _Quantitative (FG.R:79-304,
// 309-549), Get_Coef (FG.R:2-35), the start values of glm() (FG.R:1119),
// Covariate_Transform (FG.R:1612-1647) and ScoreTest_NULL_Model
// (FG.R:579-591), with the Rcpp pieces they call: getCoefficients
// (FG.cpp:3167-3199), getAIScore / getAIScore_q with GetTrace / GetTrace_q
// (FG.cpp:3113-3160, 3279-3295, 3409-3531) and fitglmmaiRPCG / _q
// (FG.cpp:3302-3341, 3610-3662).
//
// Every GRM operation goes through a Solver: Sigma^-1 B (PCG) and K B.  The
// library's Solver batches the requests of many fits into the block-PCG
// engine (fit_engine.cpp); a test backend answers them with dense FP64
// algebra.  All host arithmetic is FP64.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace saige {

// y[i] = exp(-|x[i]|), log(x[i]) over arrays by the vector math library
// (vmath.cpp)
namespace vmath {
void exp_neg_abs(const double* x, double* y, int64_t n);
void log(const double* x, double* y, int64_t n);
}  // namespace vmath

namespace fit {

struct Options {
    int maxiter = 20;              // AI-REML and Get_Coef iterations
    double tol = 0.02;             // tau convergence / zeroing threshold
    int nrun = 30;                 // probes in the first trace round
    double tol_pcg = 1e-5;
    int maxiter_pcg = 500;
    double trace_cv_cutoff = 0.0025;
    uint64_t probe_seed = 200;
};

// Column-major dense matrix.
struct Matrix {
    int64_t rows = 0, cols = 0;
    std::vector<double> a;
    Matrix() = default;
    Matrix(int64_t r, int64_t c) : rows(r), cols(c), a((size_t)(r * c), 0.0) {}
    double& operator()(int64_t i, int64_t j) { return a[(size_t)(j * rows + i)]; }
    double operator()(int64_t i, int64_t j) const { return a[(size_t)(j * rows + i)]; }
    double* col(int64_t j) { return a.data() + j * rows; }
    const double* col(int64_t j) const { return a.data() + j * rows; }
};

// One round of GetTrace / GetTrace_q (FG.cpp:3113-3160, 3409-3472) over
// probes start..end-1 (k of them), as FP64 scalars:
//   d_j = (K u_j)^T Sigma^-1 u_j,  e_j = u_j^T Sigma^-1 u_j (quantitative),
//   h_j = (Sigma^-1 X)^T K u_j,    g_j = (Sigma^-1 X)^T u_j,
// so that (K u)^T P u = d - h^T cov g and u^T P u = e - g^T cov g.  With PY
// (the first round) also K PY, Sigma^-1 K PY and (quantitative) Sigma^-1 PY.
struct TraceTerms {
    std::vector<double> d, e;          // k
    Matrix h, g;                       // k x p
    std::vector<double> APY, SAPY, SPY;
};

// What a fit needs from the GRM, over the fit's own samples (n of them).
// Calls block until the result is ready and throw on failure.
class Solver {
public:
    virtual ~Solver() = default;
    // X = Sigma^-1 B, Sigma = tau0 diag(1/w) + tau1 K (PCG, tol_pcg, maxiter_pcg)
    virtual void solve(const std::vector<double>& w, const double tau[2], const Matrix& B, Matrix& X) = 0;
    // KB = K B
    virtual void product(const Matrix& B, Matrix& KB) = 0;
    // KB = K B, then X = Sigma^-1 C where column c of C is B[:, src[c]] for
    // src[c] >= 0 and KB[:, -1 - src[c]] otherwise (a trace round: K [PY, U]
    // then Sigma^-1 [U, K PY, PY]).  The default runs the two in turn; the
    // engine's solver starts the solves as soon as the products are done.
    virtual void product_then_solve(const Matrix& B, Matrix& KB, const std::vector<double>& w, const double tau[2],
                                    const std::vector<int>& src, Matrix& X);
    // One trace round (TraceTerms) with the probes probe_vector(seed, i, n),
    // i = start..end-1.  The default forms the probes on the host and runs
    // product_then_solve; the engine's solver keeps the probes, K u and
    // Sigma^-1 u on the device and returns only the dots.
    virtual void trace_round(const std::vector<double>& w, const double tau[2], const Matrix& Sigma_iX,
                             const std::vector<double>* PY, bool quantitative, int64_t start, int64_t end,
                             uint64_t seed, TraceTerms& out);
};

// C from B, KB and src as product_then_solve defines it.
Matrix gather_columns(const Matrix& B, const Matrix& KB, const std::vector<int>& src);

enum class Trait { Binary = 0, Quantitative = 1 };

struct Result {
    double theta[2] = {0.0, 0.0};
    std::vector<double> coefficients;       // p
    std::vector<double> linear_predictors;  // n
    std::vector<double> fitted_values;      // n
    std::vector<double> working_y;          // n: Y of the last Get_Coef
    Matrix cov;                             // p x p
    bool converged = false;
    int iterations = 0;
    std::vector<double> theta_history;      // 2 per AI-REML step, starting values first
};

// The fit: y (n), X (n x p, already covariate-transformed), offset (n or empty).
Result fit_null(Solver& solver, Trait trait, const std::vector<double>& y, const Matrix& X,
                const std::vector<double>& offset, const Options& opt);

// glm() start values (FG.R:1119, 1370): R's glm.fit IRLS, glm.control defaults.
struct GlmStart {
    std::vector<double> coef, eta, mu;
};
GlmStart glm_start(Trait trait, const std::vector<double>& y, const Matrix& X, const std::vector<double>& offset,
                   int maxit = 25, double epsilon = 1e-8);

// Get_Coef's working response and weights at eta (FG.R:9-12, 24-27):
// mu = linkinv(eta), Y = eta - offset + (y - mu) / mu_eta(eta),
// W = (mu_eta(eta) / sqrt(variance(mu)))^2.  offset may be null (0).
void glm_working(Trait trait, int64_t n, const double* y, const double* eta, const double* offset, double* mu,
                 double* Y, double* W);

// getCoefficients (FG.cpp:3167-3199) after its solves, sol = Sigma^-1 [Y, X]
// (n x (1+p), column-major): cov = inv_sympd(X^T Sigma^-1 X), alpha = cov
// (Sigma^-1 X)^T Y, eta = Y - tau0 (Sigma^-1 Y - Sigma^-1 X alpha) / w.
// X is n x p column-major; cov is p x p.
void coefficients_from_solution(int64_t n, int64_t p, const double* Y, const double* X, const double* w, double tau0,
                                const double* sol, double* cov, double* alpha, double* eta);

// Covariate_Transform (FG.R:1612-1647): X -> Q sqrt(n) of its QR
// decomposition, after dropping columns that are linear combinations of
// earlier ones (|R_jj| <= 1e-10 max(1, max |R_ii|)).  Householder QR with
// R_jj = -sign(x_jj) ||x||, the convention of LAPACK's dgeqrf.
struct CovariateTransform {
    Matrix X;                    // n x p_kept
    Matrix R;                    // p_kept x p_kept
    std::vector<int> kept;       // indices of the kept columns
};
CovariateTransform covariate_transform(const Matrix& X);

// ScoreTest_NULL_Model (FG.R:579-591) for fitted mu, variance mu2 = V.
struct ScoreNull {
    Matrix XV;          // p x n
    Matrix XVX;         // p x p
    Matrix XVX_inv;     // p x p
    Matrix XXVX_inv;    // n x p
    Matrix XVX_inv_XV;  // n x p
    std::vector<double> S_a;   // p
};
ScoreNull score_test_null(const std::vector<double>& mu, const std::vector<double>& mu2,
                          const std::vector<double>& y, const Matrix& X);

// Probe i of length n: entry r is -1 when the top bit of
// mix(seed ^ mix(i*C1 + mix(r*C2 + C3))) is set, else +1 (the counter hash
// the device's saige_probes_device evaluates).
void probe_vector(uint64_t seed, int64_t i, int64_t n, double* out);

// calCV (FG.cpp:3104-3110): (sd / mean) / length, sd with n-1.
double cal_cv(const std::vector<double>& x);

// arma::inv_sympd(symmatu(A)) with the pinv fallback (FG.cpp:3183-3191).
Matrix inv_sympd(const Matrix& A);

}  // namespace fit
}  // namespace saige
