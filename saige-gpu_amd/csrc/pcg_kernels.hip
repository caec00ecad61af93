// PCG kernels on the resident GRM S = Z Z^T (K = S / Mqc).
//
// Column blocks are [Npad][kpad] row-major (kpad = 1 for the plain
// single-RHS path, 64 for the block path).  Rows >= Nnm are zero in every
// vector and in S, so padded rows never contribute.
//
// Products:
//   k_gemv  (kpad 1): one wave per row, float4 streaming of S; FP32 dot.
//   k_gemm  (kpad 64): FP32 MFMA 32x32x2; a workgroup owns 128 rows x 64
//                      columns and walks the full row panel of S, so S is
//                      read from HBM exactly once per product.
// Both fuse the Sigma epilogue of getCrossprod (FG.cpp:2397-2425):
//   Ap = tau0 * (p * (1/w)) + tau1 * ((S p) / Mqc), masked to the system's
//   samples, and emit per-workgroup FP64 partials of p^T Ap.
// Reductions of partials always run in one fixed order, so a column's
// result does not depend on which other columns share the launch.
#include "saige_internal.hpp"

#pragma clang fp contract(off)

namespace saige {

namespace {

typedef __attribute__((ext_vector_type(16))) float f32x16;

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sigma p for one element (getCrossprod, FG.cpp:2412): the reference forms
// 1/w, then b % (1/w), then tau0 * that, plus tau1 * (K b); tau1 == 0
// returns the first term alone (FG.cpp:2401-2404).
__device__ __forceinline__ float sigma_elem(float p, float y, float winv, float tau0, float tau1,
                                            float mqc) {
    float t = __fmul_rn(tau0, __fmul_rn(p, winv));
    if (tau1 == 0.0f) return t;
    return __fadd_rn(t, __fmul_rn(tau1, div_rn(y, mqc)));
}

// --------------------------------------------------------------- GEMV
constexpr int GEMV_ROWS = 4;  // rows (waves) per workgroup

__global__ __launch_bounds__(256) void k_gemv(const float* __restrict__ S, int64_t ld, int64_t Nnm,
                                              const float* __restrict__ in, float* __restrict__ out,
                                              SigmaArgs a, double* __restrict__ partial) {
    __shared__ double s_dot[GEMV_ROWS];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t row = (int64_t)blockIdx.x * GEMV_ROWS + wv;
    const float4* Srow = reinterpret_cast<const float4*>(S + row * ld);
    const float4* x4 = reinterpret_cast<const float4*>(in);
    const int64_t n4 = ld / 4;
    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, acc3 = 0.f;
    int64_t j = lane;
    for (; j + 192 < n4; j += 256) {
        float4 s0 = Srow[j], s1 = Srow[j + 64], s2 = Srow[j + 128], s3 = Srow[j + 192];
        float4 x0 = x4[j], x1 = x4[j + 64], x2 = x4[j + 128], x3 = x4[j + 192];
        acc0 += s0.x * x0.x + s0.y * x0.y + s0.z * x0.z + s0.w * x0.w;
        acc1 += s1.x * x1.x + s1.y * x1.y + s1.z * x1.z + s1.w * x1.w;
        acc2 += s2.x * x2.x + s2.y * x2.y + s2.z * x2.z + s2.w * x2.w;
        acc3 += s3.x * x3.x + s3.y * x3.y + s3.z * x3.z + s3.w * x3.w;
    }
    for (; j < n4; j += 64) {
        float4 s0 = Srow[j], x0 = x4[j];
        acc0 += s0.x * x0.x + s0.y * x0.y + s0.z * x0.z + s0.w * x0.w;
    }
    float y = wave_sum_f((acc0 + acc1) + (acc2 + acc3));
    double d = 0.0;
    if (lane == 0) {
        float o;
        if (a.mode == 0) {
            o = div_rn(y, a.mqc);
        } else {
            const int s = a.sys[0];
            const bool on = a.live[0] && row < Nnm && a.mask[(int64_t)s * ld + row];
            const float p = in[row];
            o = on ? sigma_elem(p, y, a.winv[(int64_t)s * ld + row], a.tau[2 * s], a.tau[2 * s + 1], a.mqc)
                   : 0.0f;
            d = (double)p * (double)o;
        }
        out[row] = o;
        s_dot[wv] = d;
    }
    if (a.mode == 1) {
        __syncthreads();
        if (threadIdx.x == 0)
            partial[(int64_t)blockIdx.x * 2] = (s_dot[0] + s_dot[1]) + (s_dot[2] + s_dot[3]);
    }
}

// --------------------------------------------------------------- GEMM
constexpr int GM = 128;       // rows per workgroup
constexpr int GKC = 32;       // S columns per K chunk
constexpr int GP = GKC + 4;   // LDS pitch (floats), k-permuted like the FP32 SYRK

__global__ __launch_bounds__(256) void k_gemm(const float* __restrict__ S, int64_t ld, int64_t Nnm,
                                              const float* __restrict__ in, float* __restrict__ out,
                                              SigmaArgs a, double* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) float sS[GM * GP];
    __shared__ __attribute__((aligned(16))) float sP[EW * GP];
    __shared__ double s_dot[4][EW];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * GM;

    // S chunk: 16-byte chunk c = tid + 256 q -> row c/8 (+32q), k (c%8)*4
    const int lrow = tid >> 3, lk = (tid & 7) * 4, hk = lk >> 1;
    const float* gS = S + (r0 + lrow) * ld + lk;
    // P chunk: [32 j][64 n]; chunk c = tid + 256 q -> j c/16, n (c%16)*4
    const int pj0 = tid >> 4, pn = (tid & 15) * 4;
    const float* gP = in + (int64_t)pj0 * EW + pn;

    float4 rs[4], rp[2];
    const int64_t nk = ld / GKC;
#pragma unroll
    for (int q = 0; q < 4; q++) rs[q] = *reinterpret_cast<const float4*>(gS + (int64_t)(32 * q) * ld);
#pragma unroll
    for (int q = 0; q < 2; q++) rp[q] = *reinterpret_cast<const float4*>(gP + (int64_t)(16 * q) * EW);

    f32x16 acc[2];
#pragma unroll
    for (int ni = 0; ni < 2; ni++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[ni][r] = 0.0f;

    const int h = (lane >> 5) * 16;
    for (int64_t kc = 0; kc < nk; kc++) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float* d = sS + (lrow + 32 * q) * GP;
            *reinterpret_cast<float2*>(d + hk) = make_float2(rs[q].x, rs[q].z);
            *reinterpret_cast<float2*>(d + 16 + hk) = make_float2(rs[q].y, rs[q].w);
        }
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int j = pj0 + 16 * q;              // k index within the chunk
            const int off = (j & 1) * 16 + (j >> 1);
            sP[(pn + 0) * GP + off] = rp[q].x;
            sP[(pn + 1) * GP + off] = rp[q].y;
            sP[(pn + 2) * GP + off] = rp[q].z;
            sP[(pn + 3) * GP + off] = rp[q].w;
        }
        __syncthreads();
        if (kc + 1 < nk) {
            const int64_t off = (kc + 1) * GKC;
#pragma unroll
            for (int q = 0; q < 4; q++)
                rs[q] = *reinterpret_cast<const float4*>(gS + (int64_t)(32 * q) * ld + off);
#pragma unroll
            for (int q = 0; q < 2; q++)
                rp[q] = *reinterpret_cast<const float4*>(gP + (off + 16 * q) * EW);
        }
#pragma unroll
        for (int sq = 0; sq < 4; sq++) {
            const float4 fa = *reinterpret_cast<const float4*>(sS + (wv * 32 + (lane & 31)) * GP + h + 4 * sq);
            float4 fb[2];
#pragma unroll
            for (int ni = 0; ni < 2; ni++)
                fb[ni] = *reinterpret_cast<const float4*>(sP + (ni * 32 + (lane & 31)) * GP + h + 4 * sq);
#pragma unroll
            for (int e = 0; e < 4; e++)
#pragma unroll
                for (int ni = 0; ni < 2; ni++)
                    acc[ni] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[e], fb[ni][e], acc[ni], 0, 0, 0);
        }
    }

    // epilogue: lane holds column n = ni*32 + lane%32 and rows
    // 8*(r/4) + 4*(lane/32) + r%4 of the wave's 32-row block
#pragma unroll
    for (int ni = 0; ni < 2; ni++) {
        const int n = ni * 32 + (lane & 31);
        double dsum = 0.0;
        int s = 0;
        float tau0 = 0.f, tau1 = 0.f;
        bool col_on = false;
        if (a.mode == 1) {
            s = a.sys[n];
            tau0 = a.tau[2 * s];
            tau1 = a.tau[2 * s + 1];
            col_on = a.live[n] != 0;
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int64_t row = r0 + wv * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const float y = acc[ni][r];
            float o;
            if (a.mode == 0) {
                o = div_rn(y, a.mqc);
            } else {
                const bool on = col_on && row < Nnm && a.mask[(int64_t)s * ld + row];
                const float p = in[row * EW + n];
                o = on ? sigma_elem(p, y, a.winv[(int64_t)s * ld + row], tau0, tau1, a.mqc) : 0.0f;
                dsum += (double)p * (double)o;
            }
            out[row * EW + n] = o;
        }
        if (a.mode == 1) {
            dsum += __shfl_xor(dsum, 32, 64);
            if (lane < 32) s_dot[wv][n] = dsum;
        }
    }
    if (a.mode == 1) {
        __syncthreads();
        if (tid < EW) {
            double t = (s_dot[0][tid] + s_dot[1][tid]) + (s_dot[2][tid] + s_dot[3][tid]);
            partial[((int64_t)blockIdx.x * EW + tid) * 2] = t;
        }
    }
}

// ------------------------------------------------------ vector kernels
constexpr int VROWS = 256;   // rows per workgroup in the elementwise kernels

// In-workgroup reduction of per-thread values to per-column sums
// (threads tid and tid + s share a column when kpad divides s).
__device__ __forceinline__ void col_reduce(double* sm, int tid, int kpad) {
    for (int s = 128; s >= kpad; s >>= 1) {
        __syncthreads();
        if (tid < s) sm[tid] += sm[tid + s];
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_sigma_setup(const float* __restrict__ w,
                                                     const uint8_t* __restrict__ mask_in,
                                                     const float* __restrict__ diag, int64_t Nnm,
                                                     int64_t Npad, float tau0, float tau1, float mqc,
                                                     int diag_one, float* __restrict__ winv,
                                                     float* __restrict__ minv,
                                                     uint8_t* __restrict__ mask_out,
                                                     float* __restrict__ dout) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= Npad) return;
    const bool on = i < Nnm && (mask_in == nullptr || mask_in[i]);
    if (!on) {
        winv[i] = 0.f;
        minv[i] = 0.f;
        mask_out[i] = 0;
        if (dout) dout[i] = 0.f;
        return;
    }
    const float wi = w[i];
    // getDiagOfSigma (FG.cpp:2335-2358): (tau1 * diag) / Mqc + tau0 / w, >= 1e-4
    float d = diag_one ? __fadd_rn(tau1, div_rn(tau0, wi))
                       : __fadd_rn(div_rn(__fmul_rn(tau1, diag[i]), mqc), div_rn(tau0, wi));
    if (d < 1e-4f) d = 1e-4f;
    if (dout) dout[i] = d;
    winv[i] = div_rn(1.0f, wi);
    minv[i] = div_rn(1.0f, d);   // minvVec = 1/getDiagOfSigma (FG.cpp:2637)
    mask_out[i] = 1;
}

// Start column `slot`: x = 0, r = b, z = minv r, p = z (FG.cpp:2615-2690).
__global__ __launch_bounds__(256) void k_col_init(const float* __restrict__ b, int slot, int kpad,
                                                  int64_t Nnm, const float* __restrict__ minv,
                                                  const uint8_t* __restrict__ mask, float* X, float* R,
                                                  float* Z, float* P, double* __restrict__ partial) {
    __shared__ double sm[2][256];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * VROWS + tid;
    double rz = 0.0, rr = 0.0;
    const bool on = i < Nnm && mask[i];
    const float r = on ? b[i] : 0.0f;
    const float z = __fmul_rn(minv[i], r);
    const int64_t o = i * kpad + slot;
    X[o] = 0.0f;
    R[o] = r;
    Z[o] = z;
    P[o] = z;
    rz = (double)z * (double)r;
    rr = (double)r * (double)r;
    sm[0][tid] = rz;
    sm[1][tid] = rr;
    for (int s = 128; s >= 1; s >>= 1) {
        __syncthreads();
        if (tid < s) {
            sm[0][tid] += sm[0][tid + s];
            sm[1][tid] += sm[1][tid + s];
        }
    }
    if (tid == 0) {
        partial[(int64_t)blockIdx.x * 2] = sm[0][0];
        partial[(int64_t)blockIdx.x * 2 + 1] = sm[1][0];
    }
}

// Sum partial[p][col][comp] over p in a fixed order: one workgroup per
// column, thread t takes parts t, t+256, ..., then a tree.
__device__ __forceinline__ double sum_parts(const double* partial, int64_t nparts, int kpad, int col,
                                            int comp, double* sm) {
    const int tid = threadIdx.x;
    double v = 0.0;
    for (int64_t p = tid; p < nparts; p += 256) v += partial[(p * kpad + col) * 2 + comp];
    sm[tid] = v;
    for (int s = 128; s >= 1; s >>= 1) {
        __syncthreads();
        if (tid < s) sm[tid] += sm[tid + s];
    }
    __syncthreads();
    double r = sm[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_col_init_finish(int slot, const double* __restrict__ partial,
                                                         int64_t nparts, double* rz, double* sumr2,
                                                         const int32_t* maxiter, const double* tol,
                                                         int32_t* iter, uint8_t* done, uint8_t* live) {
    __shared__ double sm[256];
    const double a = sum_parts(partial, nparts, 1, 0, 0, sm);
    const double b = sum_parts(partial, nparts, 1, 0, 1, sm);
    if (threadIdx.x == 0) {
        rz[slot] = a;
        sumr2[slot] = b;
        iter[slot] = 0;
        live[slot] = 1;
        // while (sumr2 > tolPCG && iter < maxiterPCG) (FG.cpp:2696)
        done[slot] = (b > tol[slot] && 0 < maxiter[slot]) ? 0 : 1;
    }
}

// alpha = r^T z / p^T A p (FG.cpp:2710)
__global__ __launch_bounds__(256) void k_alpha(int kpad, const double* __restrict__ partial,
                                               int64_t nparts, const double* __restrict__ rz,
                                               const uint8_t* __restrict__ done,
                                               const uint8_t* __restrict__ live, float* alpha) {
    __shared__ double sm[256];
    const int c = blockIdx.x;
    const double pap = sum_parts(partial, nparts, kpad, c, 0, sm);
    if (threadIdx.x == 0) alpha[c] = (live[c] && !done[c]) ? (float)(rz[c] / pap) : 0.0f;
}

// x += a p; r1 = r - a Ap; z1 = minv r1 (FG.cpp:2734-2767), partial z1.r1, r1.r1
__global__ __launch_bounds__(256) void k_update(int kpad, int64_t Npad, const float* __restrict__ alpha,
                                                const uint8_t* __restrict__ done,
                                                const uint8_t* __restrict__ live,
                                                const float* __restrict__ minv,
                                                const int32_t* __restrict__ sys, float* X, float* R,
                                                float* Z, const float* __restrict__ P,
                                                const float* __restrict__ AP,
                                                double* __restrict__ partial) {
    __shared__ double sm[2][256];
    const int tid = threadIdx.x;
    const int c = tid % kpad;
    const int rstep = 256 / kpad;
    const bool act = live[c] && !done[c];
    const float a = alpha[c];
    const float* mv = minv + (int64_t)sys[c] * Npad;
    double rz = 0.0, rr = 0.0;
    const int64_t rbase = (int64_t)blockIdx.x * VROWS;
    if (act) {
        for (int rl = tid / kpad; rl < VROWS; rl += rstep) {
            const int64_t i = rbase + rl;
            const int64_t o = i * kpad + c;
            const float x = __fadd_rn(X[o], __fmul_rn(a, P[o]));
            const float r1 = __fsub_rn(R[o], __fmul_rn(a, AP[o]));
            const float z1 = __fmul_rn(mv[i], r1);
            X[o] = x;
            R[o] = r1;
            Z[o] = z1;
            rz += (double)z1 * (double)r1;
            rr += (double)r1 * (double)r1;
        }
    }
    sm[0][tid] = rz;
    sm[1][tid] = rr;
    for (int s = 128; s >= kpad; s >>= 1) {
        __syncthreads();
        if (tid < s) {
            sm[0][tid] += sm[0][tid + s];
            sm[1][tid] += sm[1][tid + s];
        }
    }
    __syncthreads();
    if (tid < kpad) {
        partial[((int64_t)blockIdx.x * kpad + tid) * 2] = sm[0][tid];
        partial[((int64_t)blockIdx.x * kpad + tid) * 2 + 1] = sm[1][tid];
    }
}

// beta = z1^T r1 / z^T r; sumr2 = r1^T r1; stop test (FG.cpp:2770-2792)
__global__ __launch_bounds__(256) void k_beta(int kpad, const double* __restrict__ partial,
                                              int64_t nparts, double* rz, double* sumr2,
                                              const int32_t* __restrict__ maxiter,
                                              const double* __restrict__ tol, int32_t* iter,
                                              uint8_t* done, const uint8_t* __restrict__ live,
                                              float* beta) {
    __shared__ double sm[256];
    const int c = blockIdx.x;
    const double rzn = sum_parts(partial, nparts, kpad, c, 0, sm);
    const double rr = sum_parts(partial, nparts, kpad, c, 1, sm);
    if (threadIdx.x == 0) {
        if (live[c] && !done[c]) {
            beta[c] = (float)(rzn / rz[c]);
            rz[c] = rzn;
            sumr2[c] = rr;
            const int it = iter[c] + 1;
            iter[c] = it;
            done[c] = (rr > tol[c] && it < maxiter[c]) ? 0 : 1;
        } else {
            beta[c] = 0.0f;
        }
    }
}

// p = z1 + beta p (FG.cpp:2772)
__global__ __launch_bounds__(256) void k_pupdate(int kpad, int64_t Npad, const float* __restrict__ beta,
                                                 const uint8_t* __restrict__ done,
                                                 const uint8_t* __restrict__ live,
                                                 const float* __restrict__ Z, float* P) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= Npad * kpad) return;
    const int c = (int)(t % kpad);
    if (!live[c] || done[c]) return;
    P[t] = __fadd_rn(Z[t], __fmul_rn(beta[c], P[t]));
}

__global__ void k_gather_col(const float* __restrict__ src, int slot, int kpad, int64_t Nnm,
                             float* __restrict__ dst) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < Nnm) dst[i] = src[i * kpad + slot];
}

// column-major [k][Nnm] (host layout) <-> [Npad][kpad] row-major blocks
__global__ void k_scatter_cols(const float* __restrict__ src, int64_t k, int64_t Nnm, int64_t Npad,
                               int kpad, float* __restrict__ dst) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= Npad * kpad) return;
    const int64_t i = t / kpad;
    const int c = (int)(t % kpad);
    dst[t] = (c < k && i < Nnm) ? src[(int64_t)c * Nnm + i] : 0.0f;
}

__global__ void k_gather_cols(const float* __restrict__ src, int64_t k, int64_t Nnm, int kpad,
                              float* __restrict__ dst) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= k * Nnm) return;
    const int64_t c = t / Nnm, i = t - c * Nnm;
    dst[t] = src[i * kpad + c];
}

__global__ void k_fill_zero(uint32_t* p, int64_t n) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0u;
}

inline unsigned nblk(int64_t n, int64_t b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

int64_t gemv_max_parts(int64_t Npad) { return Npad / GEMV_ROWS; }
int64_t gemm_max_parts(int64_t Npad) { return Npad / GM; }
int64_t vec_max_parts(int64_t Npad, int) { return Npad / VROWS; }

hipError_t launch_grm_gemv(hipStream_t s, const float* S, int64_t Npad, int64_t Nnm, const float* in,
                           float* out, SigmaArgs a, double* partial, int64_t* nparts) {
    const int64_t nb = Npad / GEMV_ROWS;
    hipLaunchKernelGGL(k_gemv, dim3((unsigned)nb), dim3(256), 0, s, S, Npad, Nnm, in, out, a, partial);
    if (nparts) *nparts = nb;
    return hipGetLastError();
}

hipError_t launch_grm_gemm(hipStream_t s, const float* S, int64_t Npad, int64_t Nnm, const float* in,
                           float* out, SigmaArgs a, double* partial, int64_t* nparts) {
    const int64_t nb = Npad / GM;
    hipLaunchKernelGGL(k_gemm, dim3((unsigned)nb), dim3(256), 0, s, S, Npad, Nnm, in, out, a, partial);
    if (nparts) *nparts = nb;
    return hipGetLastError();
}

hipError_t launch_sigma_setup(hipStream_t s, const float* w, const uint8_t* mask_in, const float* diag,
                              int64_t Nnm, int64_t Npad, float tau0, float tau1, float mqc, int diag_one,
                              float* winv, float* minv, uint8_t* mask_out, float* dout) {
    hipLaunchKernelGGL(k_sigma_setup, dim3(nblk(Npad, 256)), dim3(256), 0, s, w, mask_in, diag, Nnm, Npad,
                       tau0, tau1, mqc, diag_one, winv, minv, mask_out, dout);
    return hipGetLastError();
}

hipError_t launch_col_init(hipStream_t s, const float* b, int slot, int kpad, int64_t Nnm, int64_t Npad,
                           const float* minv_sys, const uint8_t* mask_sys, float* X, float* R, float* Z,
                           float* P, double* partial, int64_t* nparts) {
    const int64_t nb = Npad / VROWS;
    hipLaunchKernelGGL(k_col_init, dim3((unsigned)nb), dim3(256), 0, s, b, slot, kpad, Nnm, minv_sys,
                       mask_sys, X, R, Z, P, partial);
    if (nparts) *nparts = nb;
    return hipGetLastError();
}

hipError_t launch_col_init_finish(hipStream_t s, int slot, int kpad, const double* partial,
                                  int64_t nparts, double* rz, double* sumr2, const int32_t* maxiter,
                                  const double* tol, int32_t* iter, uint8_t* done, uint8_t* live) {
    (void)kpad;
    hipLaunchKernelGGL(k_col_init_finish, dim3(1), dim3(256), 0, s, slot, partial, nparts, rz, sumr2,
                       maxiter, tol, iter, done, live);
    return hipGetLastError();
}

hipError_t launch_alpha(hipStream_t s, int kpad, const double* partial, int64_t nparts, const double* rz,
                        const uint8_t* done, const uint8_t* live, float* alpha) {
    hipLaunchKernelGGL(k_alpha, dim3((unsigned)kpad), dim3(256), 0, s, kpad, partial, nparts, rz, done,
                       live, alpha);
    return hipGetLastError();
}

hipError_t launch_update(hipStream_t s, int kpad, int64_t Nnm, int64_t Npad, const float* alpha,
                         const uint8_t* done, const uint8_t* live, const float* minv, const int32_t* sys,
                         float* X, float* R, float* Z, const float* P, const float* AP, double* partial,
                         int64_t* nparts) {
    (void)Nnm;
    const int64_t nb = Npad / VROWS;
    hipLaunchKernelGGL(k_update, dim3((unsigned)nb), dim3(256), 0, s, kpad, Npad, alpha, done, live, minv,
                       sys, X, R, Z, P, AP, partial);
    if (nparts) *nparts = nb;
    return hipGetLastError();
}

hipError_t launch_beta(hipStream_t s, int kpad, const double* partial, int64_t nparts, double* rz,
                       double* sumr2, const int32_t* maxiter, const double* tol, int32_t* iter,
                       uint8_t* done, const uint8_t* live, float* beta) {
    hipLaunchKernelGGL(k_beta, dim3((unsigned)kpad), dim3(256), 0, s, kpad, partial, nparts, rz, sumr2,
                       maxiter, tol, iter, done, live, beta);
    return hipGetLastError();
}

hipError_t launch_pupdate(hipStream_t s, int kpad, int64_t Nnm, int64_t Npad, const float* beta,
                          const uint8_t* done, const uint8_t* live, const float* Z, float* P) {
    (void)Nnm;
    hipLaunchKernelGGL(k_pupdate, dim3(nblk(Npad * kpad, 256)), dim3(256), 0, s, kpad, Npad, beta, done,
                       live, Z, P);
    return hipGetLastError();
}

hipError_t launch_gather_col(hipStream_t s, const float* src, int slot, int kpad, int64_t Nnm,
                             float* dst) {
    hipLaunchKernelGGL(k_gather_col, dim3(nblk(Nnm, 256)), dim3(256), 0, s, src, slot, kpad, Nnm, dst);
    return hipGetLastError();
}

hipError_t launch_scatter_cols(hipStream_t s, const float* src, int64_t k, int64_t Nnm, int64_t Npad,
                               int kpad, float* dst) {
    hipLaunchKernelGGL(k_scatter_cols, dim3(nblk(Npad * kpad, 256)), dim3(256), 0, s, src, k, Nnm, Npad,
                       kpad, dst);
    return hipGetLastError();
}

hipError_t launch_gather_cols(hipStream_t s, const float* src, int64_t k, int64_t Nnm, int kpad,
                              float* dst) {
    hipLaunchKernelGGL(k_gather_cols, dim3(nblk(k * Nnm, 256)), dim3(256), 0, s, src, k, Nnm, kpad, dst);
    return hipGetLastError();
}

hipError_t launch_fill_zero(hipStream_t s, void* p, int64_t bytes) {
    const int64_t n = bytes / 4;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_zero, dim3(nblk(n, 256)), dim3(256), 0, s, (uint32_t*)p, n);
    return hipGetLastError();
}

}  // namespace saige
