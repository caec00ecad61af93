// Array exponentials and logarithms for the host IRLS loops (glm start
// values, Get_Coef's working response), which evaluate one exp or log per
// sample per iteration.  This file is built with -ffast-math so that the
// loops call glibc's vector math library (libmvec: 4 doubles per call with
// AVX2, 2 with SSE2; at most 4 ulp from the correctly rounded value) in
// place of one scalar call per element; it holds nothing else, so the flag
// reaches no other arithmetic.  The AVX2 variants are chosen at run time.
#include "fit_core.hpp"

#include <cmath>
#include <cstdint>

namespace saige {
namespace vmath {

namespace {

#define SAIGE_VLOOPS(SUFFIX, ATTR)                                                                         \
    ATTR void exp_neg_abs_##SUFFIX(const double* __restrict x, double* __restrict y, int64_t n) {          \
        _Pragma("omp simd") for (int64_t i = 0; i < n; i++) y[i] = std::exp(-std::fabs(x[i]));            \
    }                                                                                                      \
    ATTR void log_##SUFFIX(const double* __restrict x, double* __restrict y, int64_t n) {                  \
        _Pragma("omp simd") for (int64_t i = 0; i < n; i++) y[i] = std::log(x[i]);                        \
    }

SAIGE_VLOOPS(base, )
SAIGE_VLOOPS(avx2, __attribute__((target("avx2,fma"))))

#undef SAIGE_VLOOPS

bool use_avx2() {
    static const bool v = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
    return v;
}

}  // namespace

void exp_neg_abs(const double* x, double* y, int64_t n) {
    use_avx2() ? exp_neg_abs_avx2(x, y, n) : exp_neg_abs_base(x, y, n);
}

void log(const double* x, double* y, int64_t n) { use_avx2() ? log_avx2(x, y, n) : log_base(x, y, n); }

}  // namespace vmath
}  // namespace saige
