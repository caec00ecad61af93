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

#include <algorithm>
#include <type_traits>

#pragma clang fp contract(off)

namespace saige {

namespace {

typedef __attribute__((ext_vector_type(16))) float f32x16;

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// wait for this wave's LDS traffic, then the workgroup barrier; outstanding
// global loads stay in flight
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

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

// ------------------------------------------------ lower-triangle GEMV
// k_gemv_tri: y = S x read from the lower triangle of S only (half the
// bytes).  A workgroup of eight waves owns a row set R of 256 rows (wave w:
// the strip 256R + 32w + [0, 32)) and walks one item: a range of 256-column
// chunks left of the row set's diagonal block, each element serving
//   use 1: y_i += S_ij x_j    and    use 2: y_j += S_ij x_i,
// or the diagonal 256 x 256 block itself (use 1 only, upper half included).
// Lane l owns columns 4l + [0, 4) of a chunk.  Use 1 accumulates per lane and
// row over the item and is summed across the wave's lanes at its end (one
// partial per item and row); use 2 accumulates per lane over the strip's 32
// rows, the eight strips' sums are added in LDS in wave order, and the sum
// is stored as the partial of (chunk, row set).  k_gemv_tri_epi sums each
// row's partials in a fixed order and applies k_gemv's epilogue.
constexpr int TV_ROWS = 256;           // rows per row set (eight strips of 32)
constexpr int TV_COLS = 256;           // columns per chunk (64 lanes x float4)

// index of the first use-2 partial of chunk c: chunk c' has NR - 1 - c'
__host__ __device__ __forceinline__ int64_t triv_y2_first(int64_t c, int64_t NR) {
    return c * (NR - 1) - c * (c - 1) / 2;
}

__global__ __launch_bounds__(512) void k_gemv_tri(const float* __restrict__ S, int64_t ld,
                                                  const float* __restrict__ x, const int4* __restrict__ items,
                                                  float* __restrict__ y1p, float* __restrict__ y2p) {
    __shared__ float4 sz[2][8][64];    // use-2 strip sums, two chunks
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int4 item = items[blockIdx.x];
    const int64_t R = item.x, ca = item.y, cb = item.z, slot = item.w;
    const int64_t NR = ld / TV_ROWS;
    const int64_t r0 = R * TV_ROWS + 32 * wv;   // the wave's first row (ld is a multiple of 256)
    const bool diag = ca >= R;
    const float* Sb = S + r0 * ld + 4 * lane;
    float xr[32];
#pragma unroll
    for (int r = 0; r < 32; r++) xr[r] = x[r0 + r];
    float acc[32];
#pragma unroll
    for (int r = 0; r < 32; r++) acc[r] = 0.0f;
    for (int64_t c = ca; c < cb; c++) {
        const float4 xj = *reinterpret_cast<const float4*>(x + TV_COLS * c + 4 * lane);
        float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int r = 0; r < 32; r++) {
            const float4 v = *reinterpret_cast<const float4*>(Sb + r * ld + TV_COLS * c);
            acc[r] += v.x * xj.x + v.y * xj.y + v.z * xj.z + v.w * xj.w;
            z.x += v.x * xr[r];
            z.y += v.y * xr[r];
            z.z += v.z * xr[r];
            z.w += v.w * xr[r];
        }
        if (!diag) {
            // buffer c & 1 was last read for chunk c-2, before chunk c-1's barrier
            sz[c & 1][wv][lane] = z;
            lds_barrier();
            if (wv == 0) {
                float4 sum = sz[c & 1][0][lane];
#pragma unroll
                for (int w = 1; w < 8; w++) {
                    const float4 v = sz[c & 1][w][lane];
                    sum.x += v.x;
                    sum.y += v.y;
                    sum.z += v.z;
                    sum.w += v.w;
                }
                const int64_t ps = triv_y2_first(c, NR) + (R - c - 1);
                *reinterpret_cast<float4*>(y2p + ps * TV_COLS + 4 * lane) = sum;
            }
        }
    }
    float* yp = y1p + slot * TV_ROWS + 32 * wv;
#pragma unroll
    for (int r = 0; r < 32; r++) {
        const float t = wave_sum_f(acc[r]);
        if (lane == r) yp[r] = t;
    }
}

// Y[row] = the row set's use-1 partials in slot order, then the row's chunk's
// use-2 partials in row-set order; then k_gemv's epilogue and the block's
// FP64 dot partial (mode 1), rows = threads
__global__ __launch_bounds__(256) void k_gemv_tri_epi(const float* __restrict__ y1p, const float* __restrict__ y2p,
                                                      const int2* __restrict__ rtab, int64_t ld, int64_t Nnm,
                                                      const float* __restrict__ in, float* __restrict__ out,
                                                      SigmaArgs a, double* __restrict__ partial) {
    __shared__ double sm[256];
    const int tid = threadIdx.x;
    const int64_t R = blockIdx.x, row = R * TV_ROWS + tid, NR = ld / TV_ROWS;
    const int2 rt = rtab[R];
    float y = y1p[(int64_t)rt.x * TV_ROWS + tid];
    for (int k = 1; k < rt.y; k++) y += y1p[((int64_t)rt.x + k) * TV_ROWS + tid];
    const float* p2 = y2p + triv_y2_first(R, NR) * TV_COLS + tid;
    const int64_t n2 = NR - 1 - R;
#pragma unroll 8
    for (int64_t k = 0; k < n2; k++) y += p2[k * TV_COLS];
    double d = 0.0;
    float o;
    if (a.mode == 0) {
        o = div_rn(y, a.mqc);
    } else {
        const int s = a.sys[0];
        const bool on = a.live[0] && row < Nnm && a.mask[(int64_t)s * ld + row];
        const float p = in[row];
        o = on ? sigma_elem(p, y, a.winv[(int64_t)s * ld + row], a.tau[2 * s], a.tau[2 * s + 1], a.mqc) : 0.0f;
        d = (double)p * (double)o;
    }
    out[row] = o;
    if (a.mode == 1) {
        sm[tid] = d;
        for (int st = 128; st >= 1; st >>= 1) {
            __syncthreads();
            if (tid < st) sm[tid] += sm[tid + st];
        }
        if (tid == 0) partial[R * 2] = sm[0];
    }
}

// --------------------------------------------------------------- GEMM
// Block product Y = S P for EW = 64 columns, FP32 MFMA 32x32x2.
//
// k_gemm_part<RT>: a workgroup owns 128*RT rows (4 waves x RT tiles of 32
// rows) and one of gridDim.y slices of the column range of S.  Each wave
// streams its rows of S from HBM straight into MFMA operand registers (lane
// l holds row l%32 of each tile and the 16 consecutive columns starting at
// 16*(l/32) of each 32-column chunk; the next chunk is prefetched while the
// current one is multiplied).  The matching 128 x 64 superchunk of P is
// staged in LDS as [n][k] so each lane reads its 16 B-operand values with
// four b128 loads, each feeding 2*RT MFMAs.  S is read from HBM exactly once;
// P (25.6 MB at N = 100k) is re-read per row block from L2/MALL, so RT = 2
// halves that traffic at the cost of registers.  The raw sums go to
// Ypart[slice][row][n].  RT and the slice count are picked at run time by
// saige_autotune.
//
// k_gemm_epi: Y = sum over slices in fixed order, then the Sigma epilogue
// and the per-column FP64 partial dot products.
constexpr int GKS = 128;                // columns of S per P superchunk
constexpr int GKP = GKS + 4;            // LDS pitch of sP (floats)
constexpr int GSPLIT_MAX = 8;           // most column slices

template <int RT, int PD>
__global__ __launch_bounds__(256) void k_gemm_part(const float* __restrict__ S, int64_t ld,
                                                   const float* __restrict__ in, float* __restrict__ ypart,
                                                   int halves) {
    constexpr int GM = 128 * RT;        // rows per workgroup
    constexpr int CPS = GKS / 32;       // 32-column chunks per superchunk
    constexpr int RING = PD + 1;        // S register buffers; RING divides CPS, so a
    static_assert(CPS % RING == 0, "");  // superchunk always starts at ring slot 0
    __shared__ __attribute__((aligned(16))) float sP[EW * GKP];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * GM;
    const int split = blockIdx.y, nsplit = gridDim.y;
    const int64_t nsc = ld / GKS;                           // superchunks in a row
    const int64_t sc0 = nsc * split / nsplit, sc1 = nsc * (split + 1) / nsplit;
    const int64_t kend = sc1 * GKS;                         // end column of this slice
    const int hl = lane >> 5;                               // k half of this lane
    // wave wv owns rows r0 + wv*32*RT + t*32 + [0, 32) for tiles t < RT
    const float* srow = S + (r0 + wv * 32 * RT + (lane & 31)) * ld + hl * 16;
    // P superchunk staging: thread -> column n, 32 consecutive k
    const int pn = tid & 63, pq = tid >> 6;

    f32x16 acc[RT][2];
#pragma unroll
    for (int t = 0; t < RT; t++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[t][ni][r] = 0.0f;

    // ring[slot][t][q]: the chunk starting at column kc lives in slot
    // (kc / 32) % RING; chunk kc + 32*PD is requested while chunk kc is used
    float4 ring[RING][RT][4];
    float pr[32];
    if (sc0 < sc1) {
        const int64_t k0 = sc0 * GKS;
#pragma unroll
        for (int d = 0; d < PD; d++) {
            if (k0 + 32 * d < kend) {
#pragma unroll
                for (int t = 0; t < RT; t++)
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        ring[d][t][q] =
                            *reinterpret_cast<const float4*>(srow + (int64_t)t * 32 * ld + k0 + 32 * d + 4 * q);
            }
        }
#pragma unroll
        for (int q = 0; q < 32; q++) pr[q] = in[(k0 + pq * 32 + q) * EW + pn];
    }
    for (int64_t sc = sc0; sc < sc1; sc++) {
        const int64_t k0 = sc * GKS;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; q++)
            *reinterpret_cast<float4*>(sP + pn * GKP + pq * 32 + 4 * q) =
                make_float4(pr[4 * q], pr[4 * q + 1], pr[4 * q + 2], pr[4 * q + 3]);
        __syncthreads();
        if (sc + 1 < sc1) {
            const int64_t kn = k0 + GKS;
#pragma unroll
            for (int q = 0; q < 32; q++) pr[q] = in[(kn + pq * 32 + q) * EW + pn];
        }
#pragma unroll
        for (int c = 0; c < CPS; c++) {
            // request chunk c + PD of this wave's S rows (it may lie in the next superchunk)
            const int64_t kf = k0 + 32 * (c + PD);
            if (kf < kend) {
#pragma unroll
                for (int t = 0; t < RT; t++)
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        ring[(c + PD) % RING][t][q] =
                            *reinterpret_cast<const float4*>(srow + (int64_t)t * 32 * ld + kf + 4 * q);
            }
            const float* pb = sP + c * 32 + hl * 16;
            if (halves == 3) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float4 b0 = *reinterpret_cast<const float4*>(pb + (lane & 31) * GKP + 4 * q);
                    const float4 b1 = *reinterpret_cast<const float4*>(pb + (32 + (lane & 31)) * GKP + 4 * q);
#pragma unroll
                    for (int t = 0; t < RT; t++) {
                        const float4 a = ring[c % RING][t][q];
                        acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b0.x, acc[t][0], 0, 0, 0);
                        acc[t][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b1.x, acc[t][1], 0, 0, 0);
                        acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b0.y, acc[t][0], 0, 0, 0);
                        acc[t][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b1.y, acc[t][1], 0, 0, 0);
                        acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, b0.z, acc[t][0], 0, 0, 0);
                        acc[t][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, b1.z, acc[t][1], 0, 0, 0);
                        acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, b0.w, acc[t][0], 0, 0, 0);
                        acc[t][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, b1.w, acc[t][1], 0, 0, 0);
                    }
                }
            } else {
                // one 32-column half live: half the MFMA work (the other half's
                // partial sums are left unwritten and ignored by the epilogue)
                const int h = halves >> 1;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float4 b = *reinterpret_cast<const float4*>(pb + (32 * h + (lane & 31)) * GKP + 4 * q);
#pragma unroll
                    for (int t = 0; t < RT; t++) {
                        const float4 a = ring[c % RING][t][q];
                        acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b.x, acc[t][0], 0, 0, 0);
                        acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b.y, acc[t][0], 0, 0, 0);
                        acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, b.z, acc[t][0], 0, 0, 0);
                        acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, b.w, acc[t][0], 0, 0, 0);
                    }
                }
            }
        }
    }
    // raw sums: lane holds column ni*32 + lane%32, rows 8*(r/4) + 4*(lane/32) + r%4
    // (with one live half, acc[t][0] holds that half)
    float* yp = ypart + (int64_t)split * (gridDim.x * (int64_t)GM) * EW;
    const int nh = halves == 3 ? 2 : 1;
    const int h0 = halves == 2 ? 1 : 0;
#pragma unroll
    for (int t = 0; t < RT; t++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++) {
            if (ni >= nh) break;
            const int n = (h0 + ni) * 32 + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int64_t row = r0 + wv * 32 * RT + t * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                yp[row * EW + n] = acc[t][ni][r];
            }
        }
}

// ------------------------------------------- GEMM, three-piece BF16 split
// The same product Y = S P on the BF16 MFMA pipe (32x32x16) at FP32
// accuracy.  Every FP32 operand x is split exactly into three BF16 pieces,
// x = hi + mid + lo: hi keeps the top 8 significand bits (truncation), the
// remainder x - hi is exact in FP32 and is split the same way, and the last
// remainder has at most 8 significant bits, so lo holds it exactly.  Of the
// nine piece products the six with weight >= 2^-16 are summed,
//   hi*hi + hi*mid + mid*hi + hi*lo + lo*hi + mid*mid,
// each exact in FP32 and accumulated by the MFMA in FP32; the three dropped
// terms (mid*lo, lo*mid, lo*lo) are below 2^-23 of |s p| together.  Six
// 32x32x16 BF16 MFMAs take 96 cycles per 32x32x16 block against 512 for
// eight 32x32x2 FP32 MFMAs, so the product is bound by the HBM stream of S
// rather than by the FP32 matrix pipe.
//
// k_split_p: P [Npad][EW] FP32 -> pb[3][Npad/8][EW] x 8 BF16 (piece, group of
// 8 consecutive rows g, column n): one 16-byte B operand per (g, n).
// k_gemm_x3<RT, PD, W>: rows and slices as k_gemm_part; S is split in
// registers (lane l holds row l%32 and columns 16*(l/32) + [0, 16) of each
// 32-column chunk: MFMA step s uses its 8 columns 16*(l/32) + 8s + [0, 8), so
// the B operand of step s is group 4c + 2*(l/32) + s).  Each chunk's B pieces
// (12 KB) are loaded once per workgroup (thread -> group 4c + wave, column
// lane, three pieces) into a register ring at the same distance PD as the S
// chunk, stored to a double-buffered LDS tile and read back by all four
// waves.  The chunk loop is straight-line (prefetches past the slice end are
// clamped to its last chunk and never used), so the compiler's wait counts
// stay exact and loads remain PD chunks in flight; the barrier is an
// LDS-only one (lgkmcnt, s_barrier), since __syncthreads' release fence
// would also drain every outstanding global load.  W is the waves-per-SIMD
// bound each shape is compiled for.  Raw sums go to Ypart exactly as
// k_gemm_part writes them.
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;

__device__ __forceinline__ void split3_pair(float x0, float x1, uint32_t& h, uint32_t& m, uint32_t& l) {
    const uint32_t b0 = __float_as_uint(x0), b1 = __float_as_uint(x1);
    const float r0 = x0 - __uint_as_float(b0 & 0xffff0000u);
    const float r1 = x1 - __uint_as_float(b1 & 0xffff0000u);
    const uint32_t c0 = __float_as_uint(r0), c1 = __float_as_uint(r1);
    const float q0 = r0 - __uint_as_float(c0 & 0xffff0000u);
    const float q1 = r1 - __uint_as_float(c1 & 0xffff0000u);
    h = (b1 & 0xffff0000u) | (b0 >> 16);
    m = (c1 & 0xffff0000u) | (c0 >> 16);
    l = (__float_as_uint(q1) & 0xffff0000u) | (__float_as_uint(q0) >> 16);
}

__device__ __forceinline__ void split3_8(float4 a, float4 b, u32x4& h, u32x4& m, u32x4& l) {
    uint32_t h0, h1, h2, h3, m0, m1, m2, m3, l0, l1, l2, l3;
    split3_pair(a.x, a.y, h0, m0, l0);
    split3_pair(a.z, a.w, h1, m1, l1);
    split3_pair(b.x, b.y, h2, m2, l2);
    split3_pair(b.z, b.w, h3, m3, l3);
    h = u32x4{h0, h1, h2, h3};
    m = u32x4{m0, m1, m2, m3};
    l = u32x4{l0, l1, l2, l3};
}

// The first two pieces of split3_pair: h + m is x exactly when x has at most
// 16 significant bits (the split GRM's storage), whose third piece is zero.
__device__ __forceinline__ void split2_pair(float x0, float x1, uint32_t& h, uint32_t& m) {
    const uint32_t b0 = __float_as_uint(x0), b1 = __float_as_uint(x1);
    const float r0 = x0 - __uint_as_float(b0 & 0xffff0000u);
    const float r1 = x1 - __uint_as_float(b1 & 0xffff0000u);
    h = (b1 & 0xffff0000u) | (b0 >> 16);
    m = (__float_as_uint(r1) & 0xffff0000u) | (__float_as_uint(r0) >> 16);
}

__device__ __forceinline__ void split2_8(float4 a, float4 b, u32x4& h, u32x4& m) {
    uint32_t h0, h1, h2, h3, m0, m1, m2, m3;
    split2_pair(a.x, a.y, h0, m0);
    split2_pair(a.z, a.w, h1, m1);
    split2_pair(b.x, b.y, h2, m2);
    split2_pair(b.z, b.w, h3, m3);
    h = u32x4{h0, h1, h2, h3};
    m = u32x4{m0, m1, m2, m3};
}

__device__ __forceinline__ f32x16 mfma_bf(u32x4 a, u32x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                   c, 0, 0, 0);
}

// the six terms, smallest first
__device__ __forceinline__ f32x16 mfma_x3(u32x4 ah, u32x4 am, u32x4 al, u32x4 bh, u32x4 bm, u32x4 bl,
                                          f32x16 c) {
    c = mfma_bf(al, bh, c);
    c = mfma_bf(ah, bl, c);
    c = mfma_bf(am, bm, c);
    c = mfma_bf(am, bh, c);
    c = mfma_bf(ah, bm, c);
    return mfma_bf(ah, bh, c);
}


__global__ __launch_bounds__(256) void k_split_p(const float* __restrict__ in, int64_t G,
                                                 u32x4* __restrict__ pb) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= G * EW) return;
    const int64_t g = t / EW;
    const int n = (int)(t % EW);
    float x[8];
#pragma unroll
    for (int e = 0; e < 8; e++) x[e] = in[(8 * g + e) * EW + n];
    u32x4 h, m, l;
    split3_8(make_float4(x[0], x[1], x[2], x[3]), make_float4(x[4], x[5], x[6], x[7]), h, m, l);
    pb[g * EW + n] = h;
    pb[(G + g) * EW + n] = m;
    pb[(2 * G + g) * EW + n] = l;
}

template <int RT, int PD, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W))) void k_gemm_x3(
    const float* __restrict__ S, int64_t ld, const u32x4* __restrict__ pb, float* __restrict__ ypart, int halves) {
    constexpr int GM = 128 * RT;
    constexpr int RING = PD + 1;
    __shared__ u32x4 sB[2][12 * 64];                        // [parity][(g*3 + piece)*64 + n]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, hl = lane >> 5, ln = lane & 31;
    const int64_t r0 = (int64_t)blockIdx.x * GM;
    const int split = blockIdx.y, nsplit = gridDim.y;
    const int64_t nch = ld / 32, G = ld / 8;
    const int64_t c0 = nch * split / nsplit, c1 = nch * (split + 1) / nsplit;   // >= 4 chunks
    const float* srow = S + (r0 + wv * 32 * RT + ln) * ld + hl * 16;
    const u32x4* bsrc = pb + wv * EW + lane;

    f32x16 acc[RT][2];
#pragma unroll
    for (int t = 0; t < RT; t++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[t][ni][r] = 0.0f;
    float4 ring[RING][RT][4];                               // S chunks in flight
    u32x4 bring[RING][3];                                   // B pieces in flight
    auto issue = [&](float4 (&rs)[RT][4], u32x4 (&br)[3], int64_t ch) {
#pragma unroll
        for (int p = 0; p < 3; p++) br[p] = bsrc[(p * G + 4 * ch) * EW];
#pragma unroll
        for (int t = 0; t < RT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                rs[t][q] = *reinterpret_cast<const float4*>(srow + (int64_t)t * 32 * ld + ch * 32 + 4 * q);
    };
    auto stage = [&](u32x4* sb, const u32x4 (&br)[3]) {
#pragma unroll
        for (int p = 0; p < 3; p++) sb[(wv * 3 + p) * 64 + lane] = br[p];
    };
    const int h1 = halves >> 1;   // the live half when only one is live
    auto compute = [&](const u32x4* sb, const float4 (&rs)[RT][4]) {
#pragma unroll
        for (int s = 0; s < 2; s++) {
            u32x4 ah[RT], am[RT], al[RT];
#pragma unroll
            for (int t = 0; t < RT; t++) split3_8(rs[t][2 * s], rs[t][2 * s + 1], ah[t], am[t], al[t]);
            if (halves == 3) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const u32x4* bb = sb + (2 * hl + s) * 3 * 64 + 32 * h + ln;
                    const u32x4 bh = bb[0], bm = bb[64], bl = bb[128];
#pragma unroll
                    for (int t = 0; t < RT; t++) acc[t][h] = mfma_x3(ah[t], am[t], al[t], bh, bm, bl, acc[t][h]);
                }
            } else {
                const u32x4* bb = sb + (2 * hl + s) * 3 * 64 + 32 * h1 + ln;
                const u32x4 bh = bb[0], bm = bb[64], bl = bb[128];
#pragma unroll
                for (int t = 0; t < RT; t++) acc[t][0] = mfma_x3(ah[t], am[t], al[t], bh, bm, bl, acc[t][0]);
            }
        }
    };
#pragma unroll
    for (int d = 0; d < PD; d++) issue(ring[d], bring[d], c0 + d);
    int64_t cb = c0;
    for (; cb + RING <= c1; cb += RING) {
#pragma unroll
        for (int u = 0; u < RING; u++) {
            const int64_t c = cb + u;
            u32x4* sb = sB[(c - c0) & 1];
            stage(sb, bring[u]);
            lds_barrier();
            issue(ring[(u + PD) % RING], bring[(u + PD) % RING], c + PD < c1 ? c + PD : c1 - 1);
            compute(sb, ring[u]);
        }
    }
    // the last (c1 - cb) < RING chunks are already in slots 0, 1, ...
#pragma unroll
    for (int u = 0; u < RING - 1; u++) {
        const int64_t c = cb + u;
        if (c < c1) {
            u32x4* sb = sB[(c - c0) & 1];
            stage(sb, bring[u]);
            lds_barrier();
            compute(sb, ring[u]);
        }
    }
    float* yp = ypart + (int64_t)split * (gridDim.x * (int64_t)GM) * EW;
    const int nh = halves == 3 ? 2 : 1;
    const int hh0 = halves == 2 ? 1 : 0;
#pragma unroll
    for (int t = 0; t < RT; t++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++) {
            if (ni >= nh) break;
            const int n = (hh0 + ni) * 32 + ln;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int64_t row = r0 + wv * 32 * RT + t * 32 + 8 * (r >> 2) + 4 * hl + (r & 3);
                yp[row * EW + n] = acc[t][ni][r];
            }
        }
}

// k_gemm_x3w<PD>: the same per-wave work as k_gemm_x3<1, PD, 3> (one 32-row
// tile per wave, identical arithmetic and bits) in workgroups of twelve
// waves (384 rows), so each chunk's B pieces (12 KB, one 16-byte piece per
// thread: wave w moves group 4c + w/3, piece w%3) serve three times as many
// rows of S: B travels from L2 at a quarter of S's bytes instead of three
// quarters.  A workgroup needs three waves per SIMD, so registers allow at
// most PD = 2.  Waves whose rows lie past Npad (the last workgroup) read row
// Npad-32 in place of theirs, take their part in staging and barriers, and
// store nothing.  FULL: both halves live, fixed at compile time (faster for
// full launches); otherwise `halves` is read at run time (faster for
// one-half launches, measured).
template <int PD, bool FULL>
__global__ __launch_bounds__(768) void k_gemm_x3w(const float* __restrict__ S, int64_t ld,
                                                 const u32x4* __restrict__ pb, float* __restrict__ ypart,
                                                 int halves) {
    constexpr int NWV = 12;
    constexpr int GM = 32 * NWV;
    constexpr int RING = PD + 1;
    __shared__ u32x4 sB[2][12 * 64];                        // [parity][(g*3 + piece)*64 + n]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, hl = lane >> 5, ln = lane & 31;
    const int64_t r0 = (int64_t)blockIdx.x * GM;
    const int split = blockIdx.y, nsplit = gridDim.y;
    const int64_t nch = ld / 32, G = ld / 8;
    const int64_t c0 = nch * split / nsplit, c1 = nch * (split + 1) / nsplit;   // >= 4 chunks
    const int64_t rw = r0 + wv * 32;                        // the wave's first row
    const bool valid = rw < ld;                             // ld = Npad, a multiple of 32
    const float* srow = S + ((valid ? rw : ld - 32) + ln) * ld + hl * 16;
    const int bg = wv / 3, bp = wv % 3;                     // staging: group 4c + bg, piece bp
    const u32x4* bsrc = pb + (bp * G + bg) * EW + lane;
    u32x4* bdst0 = &sB[0][(bg * 3 + bp) * 64 + lane];
    u32x4* bdst1 = &sB[1][(bg * 3 + bp) * 64 + lane];

    f32x16 acc[2];
#pragma unroll
    for (int ni = 0; ni < 2; ni++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[ni][r] = 0.0f;
    float4 ring[RING][4];
    u32x4 bring[RING];
    auto issue = [&](float4 (&rs)[4], u32x4& br, int64_t ch) {
        br = bsrc[4 * ch * EW];
#pragma unroll
        for (int q = 0; q < 4; q++) rs[q] = *reinterpret_cast<const float4*>(srow + ch * 32 + 4 * q);
    };
    const int h1 = halves >> 1;
    auto compute = [&](const u32x4* sb, const float4 (&rs)[4]) {
#pragma unroll
        for (int s = 0; s < 2; s++) {
            u32x4 ah, am, al;
            split3_8(rs[2 * s], rs[2 * s + 1], ah, am, al);
            if (FULL || halves == 3) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const u32x4* bb = sb + (2 * hl + s) * 3 * 64 + 32 * h + ln;
                    acc[h] = mfma_x3(ah, am, al, bb[0], bb[64], bb[128], acc[h]);
                }
            } else {
                const u32x4* bb = sb + (2 * hl + s) * 3 * 64 + 32 * h1 + ln;
                acc[0] = mfma_x3(ah, am, al, bb[0], bb[64], bb[128], acc[0]);
            }
        }
    };
#pragma unroll
    for (int d = 0; d < PD; d++) issue(ring[d], bring[d], c0 + d);
    int64_t cb = c0;
    for (; cb + RING <= c1; cb += RING) {
#pragma unroll
        for (int u = 0; u < RING; u++) {
            const int64_t c = cb + u;
            const int par = (int)((c - c0) & 1);
            *(par ? bdst1 : bdst0) = bring[u];
            lds_barrier();
            issue(ring[(u + PD) % RING], bring[(u + PD) % RING], c + PD < c1 ? c + PD : c1 - 1);
            compute(sB[par], ring[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < RING - 1; u++) {
        const int64_t c = cb + u;
        if (c < c1) {
            const int par = (int)((c - c0) & 1);
            *(par ? bdst1 : bdst0) = bring[u];
            lds_barrier();
            compute(sB[par], ring[u]);
        }
    }
    if (!valid) return;
    float* yp = ypart + (int64_t)split * ld * EW;
    const int nh = (FULL || halves == 3) ? 2 : 1;
    const int hh0 = (!FULL && halves == 2) ? 1 : 0;
#pragma unroll
    for (int ni = 0; ni < 2; ni++) {
        if (ni >= nh) break;
        const int n = (hh0 + ni) * 32 + ln;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int64_t row = rw + 8 * (r >> 2) + 4 * hl + (r & 3);
            yp[row * EW + n] = acc[ni][r];
        }
    }
}

// ------------------------------------------------- lower-triangle product
// k_gemm_t16<NWV, PD, NQ, SP>: the three-piece product read from the lower
// triangle of S only (half the bytes), on the 16x16x32 BF16 MFMA, over NQ
// tiles of 16 right-hand sides: NQ = 4 for all 64 columns, 2 for 32 and 1
// for 16 block columns that a map names.  S is symmetric, so a 32x32 block S_IJ left
// of the diagonal serves two products:
//   use 1: Y_I += S_IJ P_J
//   use 2: Y_J += S_IJ^T P_I
// A workgroup of NWV waves owns a row set R of 32*NWV rows (wave w: rows
// 32*(NWV*R + w) + [0, 32)) and walks one item: a range of 32-column chunks.
// Chunks left of the row set's diagonal block (c < NWV*R) are used twice;
// the item over the diagonal block (chunks NWV*R + [0, NWV)) runs use 1
// alone over the whole square, its upper half included.  Per wave and
// chunk:
// * the lane (l%16, l/16) loads rows 16r + l%16 (row tiles r = 0, 1),
//   columns 8*(l/16) + [0, 8): the A operand of use 1 for row tile r, K = the
//   chunk's columns, split into BF16 pieces in registers;
// * use 1: Y[16r + .][tile t] += S_r P_J, B = the chunk's P pieces, staged
//   once per workgroup in LDS (lane: column l%16 of the tile, rows
//   8*(l/16) + [0, 8) of the chunk);
// * use 2 needs S in the B layout with K = rows.  T_r = S_r E_jg (E_jg the 0/1
//   matrix picking columns 16*jg + [0, 16)) is exact, one value and zeros per
//   element, and its result layout (lane: column l%16, rows 4*(l/16) +
//   [0, 4) of the tile) gives lane group g rows 4g + [0, 4) of each row
//   tile.  The B operand's eight K values of lane group g are rows 4g +
//   [0, 4) of tile 0 and of tile 1 (T's high halves), and the A operand
//   P_I^T (lane: rhs l%16, three pieces split once per item) is loaded in
//   that row order;
// * use 2: Y_J^T[tile t][16*jg + .] = P_I^T T_jg per j-group jg, stored to
//   the wave's use-2 tile [column j][rhs] in LDS.
// The NWV use-2 tiles of a chunk are summed in LDS in wave order and stored
// as the partial of (chunk, row set) in y2p.  With eight waves the LDS holds
// two such tiles, and chunk c's sum is formed in chunk c+1's iteration,
// beside its MFMA work, so one barrier per chunk suffices; twelve waves'
// tiles fit once, and the sum follows a second barrier in the chunk's own
// iteration.  Use 1 accumulates in registers over the item and is stored as
// the item's partial in y1p.  k_tri_reduce sums, for each row, its row set's
// use-1 partials in slot order and then its chunk's use-2 partials in
// row-set order: a fixed order, so results are reproducible, and a column's
// arithmetic is the same whatever NQ, so a launch can be as narrow as its
// live columns (a 16-column launch does a quarter of a full one's MFMA work)
// and a column's bits still do not depend on its co-tenants.  A 16- or
// 32-column launch takes any 16 or 32 block columns (a map of launch column
// -> block column), so at most 16 or 32 live columns anywhere in the block
// run in one.  The arithmetic of an output column depends only on its own
// column of P, in whatever position of the MFMA operands it sits.
// Eight waves have registers for a software-pipelined loop: chunk c's pieces
// are split during chunk c-1, beside that chunk's MFMAs.  Twelve waves split
// each chunk at its own turn, PD chunks of raw data ahead.
// SP = 2: S holds values of at most 16 significant bits (the split GRM's
// storage), so its third piece is zero: S is split into two pieces, and the
// products and the transpose of the zero piece are left out (five piece
// products per use instead of six, two transposes instead of three).  Adding
// a zero product leaves an accumulator unchanged, so on such an S both forms
// give the same bits.

// a use-2 tile in LDS: [column j][64 rhs], float4 slot s of row j stored at
// slot s ^ (j % 16), so the column-per-lane writes and row-wise reads both
// spread over all banks
__device__ __forceinline__ int tri_sr_index(int j, int q) { return j * 64 + 4 * ((q >> 2) ^ (j & 15)) + (q & 3); }

// index of the first use-2 partial of chunk c: chunk c' has NR - 1 - c'/NWV
template <int NWV>
__host__ __device__ __forceinline__ int64_t tri_y2_first(int64_t c, int64_t NR) {
    const int64_t b = c / NWV;
    return NWV * (b * (NR - 1) - b * (b - 1) / 2) + (c - b * NWV) * (NR - 1 - b);
}

typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ f32x4 mfma16(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                   c, 0, 0, 0);
}

// the piece products in mfma_x3's order (smallest first); _a2 / _b2: the
// third A / B piece is zero (its products add exact zeros to the
// accumulator) and they are left out
__device__ __forceinline__ f32x4 mfma16_x3(u32x4 ah, u32x4 am, u32x4 al, u32x4 bh, u32x4 bm, u32x4 bl, f32x4 c) {
    c = mfma16(al, bh, c);
    c = mfma16(ah, bl, c);
    c = mfma16(am, bm, c);
    c = mfma16(am, bh, c);
    c = mfma16(ah, bm, c);
    return mfma16(ah, bh, c);
}

__device__ __forceinline__ f32x4 mfma16_a2(u32x4 ah, u32x4 am, u32x4 bh, u32x4 bm, u32x4 bl, f32x4 c) {
    c = mfma16(ah, bl, c);
    c = mfma16(am, bm, c);
    c = mfma16(am, bh, c);
    c = mfma16(ah, bm, c);
    return mfma16(ah, bh, c);
}

__device__ __forceinline__ f32x4 mfma16_b2(u32x4 ah, u32x4 am, u32x4 al, u32x4 bh, u32x4 bm, f32x4 c) {
    c = mfma16(al, bh, c);
    c = mfma16(am, bm, c);
    c = mfma16(am, bh, c);
    c = mfma16(ah, bm, c);
    return mfma16(ah, bh, c);
}

// one piece of both row tiles, transposed for use 2: operand jg holds
// S[rows 4g + [0, 4) of tile 0, then of tile 1][column 16*jg + l%16]
__device__ __forceinline__ void t16_transpose(const u32x4 (&a)[2], const u32x4 (&e)[2], u32x4 (&t)[2]) {
    const f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int jg = 0; jg < 2; jg++) {
        const f32x4 T0 = mfma16(a[0], e[jg], z);
        const f32x4 T1 = mfma16(a[1], e[jg], z);
        t[jg][0] = __builtin_amdgcn_perm(__float_as_uint(T0[1]), __float_as_uint(T0[0]), 0x07060302u);
        t[jg][1] = __builtin_amdgcn_perm(__float_as_uint(T0[3]), __float_as_uint(T0[2]), 0x07060302u);
        t[jg][2] = __builtin_amdgcn_perm(__float_as_uint(T1[1]), __float_as_uint(T1[0]), 0x07060302u);
        t[jg][3] = __builtin_amdgcn_perm(__float_as_uint(T1[3]), __float_as_uint(T1[2]), 0x07060302u);
    }
}

// block column of launch column k (< 32) of a narrow launch: byte k of the
// map, picked with selects (no indexing into the kernel argument)
__device__ __forceinline__ int map_col(const ColMap32& m, int k) {
    const uint32_t lo = (k & 8) ? ((k & 4) ? m.w[3] : m.w[2]) : ((k & 4) ? m.w[1] : m.w[0]);
    const uint32_t hi = (k & 8) ? ((k & 4) ? m.w[7] : m.w[6]) : ((k & 4) ? m.w[5] : m.w[4]);
    return (int)((((k & 16) ? hi : lo) >> (8 * (k & 3))) & 0xFFu);
}

// cm: the block columns of a 16- or 32-column launch (NQ 1 or 2), which
// holds its P pieces, partials and LDS tiles at launch columns [0, 16*NQ)
template <int NWV, int PD, int NQ, int SP>
__global__ __launch_bounds__(64 * NWV) void k_gemm_t16(const float* __restrict__ S, int64_t ld,
                                                       const u32x4* __restrict__ pb, const float* __restrict__ P,
                                                       const int4* __restrict__ items, float* __restrict__ y1p,
                                                       float* __restrict__ y2p, ColMap32 cm) {
    constexpr int GM = 32 * NWV;
    constexpr bool PIPE = NWV == 8;                  // software-pipelined chunk loop (register room)
    constexpr int RING = PD + 1;                     // raw chunks held
    constexpr int NTH = 64 * NWV;
    constexpr int NSTG = 192 * NQ;                   // P pieces a chunk needs: 4 groups x 3 pieces x 16*NQ columns
    constexpr int STG = (NSTG + NTH - 1) / NTH;      // P pieces staged per thread and chunk
    constexpr int YS = 16 * NQ;                      // row stride of the partials: the launch's columns
    __shared__ u32x4 sB0[12 * 64], sB1[12 * 64];     // a chunk's P pieces, [(g*3 + piece)*64 + n]
    constexpr int SRB = NWV == 8 ? 2 : 1;            // use-2 tile buffers
    __shared__ float sR[SRB][NWV * 32 * 64];         // use-2 results, [wave][column j][launch column]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, g = lane >> 4, ln = lane & 15;
    const int4 item = items[blockIdx.x];
    const int64_t R = item.x, ca = item.y, cb = item.z, slot = item.w;
    const int64_t nch = ld / 32, G = ld / 8, NR = (nch + NWV - 1) / NWV;
    const int64_t rw = R * GM + wv * 32;             // the wave's first row
    const bool valid = rw < ld;                      // ld = Npad, a multiple of 32
    const float* srow = S + ((valid ? rw : ld - 32) + ln) * ld + 8 * g;
    constexpr int c0 = 0;                            // the launch's first column in LDS and partials
    // the block column of launch column k (of 16 * NQ)
    auto gcol = [&](int k) { return NQ < 4 ? map_col(cm, k) : k; };
    // staging: element k of a chunk's NSTG pieces is launch column k % (16*NQ)
    // of (group, piece) k / (16*NQ); threads past the count repeat an
    // element of the first round
    const u32x4* bsrc[STG];
    int bdi[STG];
#pragma unroll
    for (int s = 0; s < STG; s++) {
        const int k = tid + s * NTH < NSTG ? tid + s * NTH : tid % NSTG;
        const int gp = k / (16 * NQ), kk = k % (16 * NQ);
        bsrc[s] = pb + ((int64_t)(gp % 3) * G + gp / 3) * EW + gcol(kk);
        bdi[s] = gp * 64 + c0 + kk;
    }
    // E_jg as B operand: lane (ln, g) holds K = columns 8g + [0, 8) of column
    // ln, E_jg[k][n] = 1 when k == 16*jg + n
    u32x4 esel[2];
#pragma unroll
    for (int jg = 0; jg < 2; jg++) {
        const int t1 = 16 * jg + ln - 8 * g;
#pragma unroll
        for (int k = 0; k < 4; k++)
            esel[jg][k] = (t1 == 2 * k ? 0x3F80u : 0u) | (t1 == 2 * k + 1 ? 0x3F800000u : 0u);
    }
    f32x4 acc1[2][NQ];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int t = 0; t < NQ; t++) acc1[r][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float4 ring[RING][2][2];                         // [slot][row tile][half of the lane's 8 columns]
    u32x4 bring[RING][STG];
    auto issue = [&](float4 (&rs)[2][2], u32x4 (&br)[STG], int64_t ch) {
#pragma unroll
        for (int s = 0; s < STG; s++) br[s] = bsrc[s][4 * ch * EW];
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int h = 0; h < 2; h++)
                rs[r][h] = *reinterpret_cast<const float4*>(srow + (int64_t)r * 16 * ld + ch * 32 + 4 * h);
    };
    // the run over the item; U2: chunks left of the diagonal (use 1 and 2)
    auto run = [&](auto u2tag) {
        constexpr bool U2 = decltype(u2tag)::value;
        // the A operand of use 2: P rows rw + 4g + tt (tt < 4) and
        // rw + 16 + 4g + tt - 4 (tt >= 4), launch column 16t + ln, as three
        // pieces (zero for rows past ld)
        u32x4 pah[NQ], pam[NQ], pal[NQ];
        if (U2) {
#pragma unroll
            for (int t = 0; t < NQ; t++) {
                const int col = gcol(16 * t + ln);
                float v[8];
#pragma unroll
                for (int tt = 0; tt < 8; tt++)
                    v[tt] = valid ? P[(rw + (tt < 4 ? 4 * g + tt : 12 + 4 * g + tt)) * EW + col] : 0.0f;
                split3_8(make_float4(v[0], v[1], v[2], v[3]), make_float4(v[4], v[5], v[6], v[7]), pah[t], pam[t],
                         pal[t]);
            }
        }
        // the sum over the waves of chunk cc's use-2 tiles (32 columns x
        // 16*NQ rhs, NQ values per lane of waves [0, 8)), stored as its
        // partial for this row set
        auto reduce = [&](int64_t cc, const float* sr) {
            if (wv < 8) {
                const int o = (wv * 64 + lane) * NQ;
                const int j = o / (16 * NQ), ql = o % (16 * NQ);
                const float* src = sr + tri_sr_index(j, ql);
                float* dst = y2p + (tri_y2_first<NWV>(cc, NR) + (R - cc / NWV - 1)) * (32 * YS) + j * YS + ql;
                if constexpr (NQ == 4) {
                    float4 sum = *reinterpret_cast<const float4*>(src);
#pragma unroll
                    for (int w = 1; w < NWV; w++) {
                        const float4 v = *reinterpret_cast<const float4*>(src + w * 32 * 64);
                        sum.x += v.x;
                        sum.y += v.y;
                        sum.z += v.z;
                        sum.w += v.w;
                    }
                    *reinterpret_cast<float4*>(dst) = sum;
                } else if constexpr (NQ == 2) {
                    float2 sum = *reinterpret_cast<const float2*>(src);
#pragma unroll
                    for (int w = 1; w < NWV; w++) {
                        const float2 v = *reinterpret_cast<const float2*>(src + w * 32 * 64);
                        sum.x += v.x;
                        sum.y += v.y;
                    }
                    *reinterpret_cast<float2*>(dst) = sum;
                } else {
                    float sum = src[0];
#pragma unroll
                    for (int w = 1; w < NWV; w++) sum += src[w * 32 * 64];
                    *dst = sum;
                }
            }
        };
        // chunk c's MFMA work from its pieces ([.][r]: row tile r), with its P
        // pieces in LDS buffer par: use 1, then (U2) the transposes, use 2 and
        // the wave's use-2 tile to LDS
        auto chunk_mfma = [&](int par, int64_t c, const u32x4 (&ah)[2], const u32x4 (&am)[2],
                              const u32x4 (&al)[2]) {
            const u32x4* sb = (par ? sB1 : sB0) + g * 3 * 64 + c0 + ln;
#pragma unroll
            for (int r = 0; r < 2; r++) {
#pragma unroll
                for (int t = 0; t < NQ; t++) {
                    const u32x4* bb = sb + 16 * t;
                    if (SP == 3)
                        acc1[r][t] = mfma16_x3(ah[r], am[r], al[r], bb[0], bb[64], bb[128], acc1[r][t]);
                    else
                        acc1[r][t] = mfma16_a2(ah[r], am[r], bb[0], bb[64], bb[128], acc1[r][t]);
                }
            }
            if (U2) {
                u32x4 th[2], tm[2], tl[2];
                t16_transpose(ah, esel, th);
                t16_transpose(am, esel, tm);
                if (SP == 3) t16_transpose(al, esel, tl);
                // with two buffers, chunk c's tile goes to buffer c & 1, last
                // read for chunk c-2 before this chunk's barrier
                float* dst = sR[SRB == 2 ? (c & 1) : 0] + wv * 32 * 64;
#pragma unroll
                for (int t = 0; t < NQ; t++) {
#pragma unroll
                    for (int jg = 0; jg < 2; jg++) {
                        f32x4 c2 = {0.0f, 0.0f, 0.0f, 0.0f};
                        if (SP == 3)
                            c2 = mfma16_x3(pah[t], pam[t], pal[t], th[jg], tm[jg], tl[jg], c2);
                        else
                            c2 = mfma16_b2(pah[t], pam[t], pal[t], th[jg], tm[jg], c2);
                        // c2[a]: rhs 16t + 4g + a, column 16*jg + ln
                        *reinterpret_cast<float4*>(dst + tri_sr_index(16 * jg + ln, 16 * t + 4 * g)) =
                            make_float4(c2[0], c2[1], c2[2], c2[3]);
                    }
                }
                if (SRB == 1) {
                    lds_barrier();
                    reduce(c, sR[0]);
                }
            }
        };
        auto step = [&](int u, int64_t c, bool pre) {
            const int par = (int)((c - ca) & 1);
            u32x4* sbw = par ? sB1 : sB0;
#pragma unroll
            for (int s = 0; s < STG; s++) sbw[bdi[s]] = bring[u][s];
            lds_barrier();
            if (SRB == 2 && U2 && c > ca) reduce(c - 1, sR[(c - 1) & 1]);
            if (pre) issue(ring[(u + PD) % RING], bring[(u + PD) % RING], c + PD < cb ? c + PD : cb - 1);
            u32x4 ah[2], am[2], al[2];
#pragma unroll
            for (int r = 0; r < 2; r++) {
                if (SP == 3)
                    split3_8(ring[u][r][0], ring[u][r][1], ah[r], am[r], al[r]);
                else
                    split2_8(ring[u][r][0], ring[u][r][1], ah[r], am[r]);
            }
            chunk_mfma(par, c, ah, am, al);
        };
        if (PIPE) {
            // software-pipelined: chunk c's pieces were split during chunk
            // c-1, so this chunk's MFMAs run beside the next chunk's split.
            // Raw chunks c+1 .. c+RING are resident or in flight (slot (c-ca)
            // % RING), and so are their P pieces; chunk ids past the item's
            // end are clamped to its last chunk, whose data is then split or
            // staged unused.  The loop is unrolled by UNR, a multiple of RING
            // and of 2, so slots and the pieces' set (the chunk's parity) are
            // compile-time indices.
            constexpr int UNR = RING == 3 ? 6 : 2;
            u32x4 pcs[2][SP][2];                      // [set][piece][row tile]
            auto split_into = [&](const float4 (&rs)[2][2], u32x4 (&p)[SP][2]) {
#pragma unroll
                for (int r = 0; r < 2; r++) {
                    if constexpr (SP == 3)
                        split3_8(rs[r][0], rs[r][1], p[0][r], p[1][r], p[2][r]);
                    else
                        split2_8(rs[r][0], rs[r][1], p[0][r], p[1][r]);
                }
            };
            // k: the chunk's offset from ca modulo UNR
            auto pstep = [&](int k, int64_t c) {
                const int u = k % RING, par = k & 1;
                u32x4* sbw = par ? sB1 : sB0;
#pragma unroll
                for (int st = 0; st < STG; st++) sbw[bdi[st]] = bring[u][st];
                lds_barrier();
                if (SRB == 2 && U2 && c > ca) reduce(c - 1, sR[(c - 1) & 1]);
                split_into(ring[(k + 1) % RING], pcs[par ^ 1]);
                issue(ring[u], bring[u], c + RING < cb ? c + RING : cb - 1);
                chunk_mfma(par, c, pcs[par][0], pcs[par][1], pcs[par][SP - 1]);
            };
#pragma unroll
            for (int d = 0; d < RING; d++) issue(ring[d], bring[d], ca + d < cb ? ca + d : cb - 1);
            split_into(ring[0], pcs[0]);
            int64_t c = ca;
            if (RING == 2) {
                // written out: the generic form below allocates more registers here
                for (; c + 2 <= cb; c += 2) {
                    pstep(0, c);
                    pstep(1, c + 1);
                }
                if (c < cb) pstep(0, c);
            } else {
                for (; c + UNR <= cb; c += UNR) {
#pragma unroll
                    for (int k = 0; k < UNR; k++) pstep(k, c + k);
                }
#pragma unroll
                for (int k = 0; k < UNR - 1; k++)
                    if (c + k < cb) pstep(k, c + k);
            }
        } else {
#pragma unroll
            for (int d = 0; d < PD; d++) issue(ring[d], bring[d], ca + d < cb ? ca + d : cb - 1);
            int64_t cbase = ca;
            for (; cbase + RING <= cb; cbase += RING) {
#pragma unroll
                for (int u = 0; u < RING; u++) step(u, cbase + u, true);
            }
#pragma unroll
            for (int u = 0; u < RING - 1; u++)
                if (cbase + u < cb) step(u, cbase + u, false);
        }
        if (SRB == 2 && U2) {
            lds_barrier();
            reduce(cb - 1, sR[(cb - 1) & 1]);
        }
    };
    if (ca >= R * NWV)
        run(std::integral_constant<bool, false>());
    else
        run(std::integral_constant<bool, true>());
    if (!valid) return;
    float* yp = y1p + slot * GM * YS;
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int t = 0; t < NQ; t++)
#pragma unroll
            for (int a = 0; a < 4; a++) yp[(int64_t)(wv * 32 + 16 * r + 4 * g + a) * YS + 16 * t + ln] = acc1[r][t][a];
}

// k_tri_reduce: Y[row][q] = the row set's use-1 partials (rtab[R] = {first
// slot, count}) in slot order, then the chunk's use-2 partials in row-set
// order; float4 per thread, QW threads per row (16: all 64 columns; 8 and 4:
// the 32 or 16 columns of a narrow launch, stored to the block columns of
// cm).  The partials hold the launch's 4 * QW columns only (row stride
// 4 * QW).
template <int NWV, int QW>
__global__ __launch_bounds__(256) void k_tri_reduce(const float* __restrict__ y1p, const float* __restrict__ y2p,
                                                    const int2* __restrict__ rtab, int64_t ld, ColMap32 cm,
                                                    float* __restrict__ y) {
    constexpr int GM = 32 * NWV, YS = 4 * QW;
    const int64_t row = (int64_t)blockIdx.x * (256 / QW) + threadIdx.x / QW;
    const int q = 4 * (threadIdx.x % QW);               // the launch's column
    const int64_t nch = ld / 32, NR = (nch + NWV - 1) / NWV;
    const int64_t c = row / 32, R = c / NWV;
    const int2 rt = rtab[R];
    const int64_t rr = row - R * GM;
    const float* p1 = y1p + ((int64_t)rt.x * GM + rr) * YS + q;
    float4 s = *reinterpret_cast<const float4*>(p1);
    for (int k = 1; k < rt.y; k++) {
        const float4 v = *reinterpret_cast<const float4*>(p1 + (int64_t)k * GM * YS);
        s.x += v.x;
        s.y += v.y;
        s.z += v.z;
        s.w += v.w;
    }
    const float* p2 = y2p + tri_y2_first<NWV>(c, NR) * (32 * YS) + (row % 32) * YS + q;
    const int64_t n2 = NR - 1 - R;
#pragma unroll 8
    for (int64_t k = 0; k < n2; k++) {
        const float4 v = *reinterpret_cast<const float4*>(p2 + k * (32 * YS));
        s.x += v.x;
        s.y += v.y;
        s.z += v.z;
        s.w += v.w;
    }
    if constexpr (QW < 16) {
        float* yr = y + row * EW;
        yr[map_col(cm, q)] = s.x;
        yr[map_col(cm, q + 1)] = s.y;
        yr[map_col(cm, q + 2)] = s.z;
        yr[map_col(cm, q + 3)] = s.w;
    } else {
        *reinterpret_cast<float4*>(y + row * EW + q) = s;
    }
}

constexpr int EROWS = 256;   // rows per workgroup of the epilogue

__global__ __launch_bounds__(256) void k_gemm_epi(const float* __restrict__ ypart, int nsplit, int64_t Npad,
                                                  int64_t Nnm, const float* __restrict__ in,
                                                  float* __restrict__ out, SigmaArgs a,
                                                  double* __restrict__ partial) {
    __shared__ double sm[256];
    const int tid = threadIdx.x;
    const int n = tid & 63;
    int s = 0;
    float tau0 = 0.f, tau1 = 0.f;
    bool col_on = false, product = false;
    if (a.mode == 1) {
        s = a.sys[n];
        tau0 = a.tau[2 * s];
        tau1 = a.tau[2 * s + 1];
        col_on = a.live[n] != 0;
        product = a.kind != nullptr && a.kind[n] != 0;
    }
    double dsum = 0.0;
    const int64_t rb = (int64_t)blockIdx.x * EROWS;
    for (int rl = tid >> 6; rl < EROWS; rl += 4) {
        const int64_t row = rb + rl;
        const int64_t o = row * EW + n;
        float y = ypart[o];
#pragma unroll
        for (int sp = 1; sp < nsplit; sp++) y += ypart[(int64_t)sp * Npad * EW + o];
        float r;
        if (a.mode == 0) {
            r = div_rn(y, a.mqc);
        } else if (product) {
            // a product slot: K p on the system's samples, stored as the result
            const bool on = col_on && row < Nnm && a.mask[(int64_t)s * Npad + row];
            r = on ? div_rn(y, a.mqc) : 0.0f;
            a.xout[o] = r;
        } else {
            const bool on = col_on && row < Nnm && a.mask[(int64_t)s * Npad + row];
            const float p = in[o];
            r = on ? sigma_elem(p, y, a.winv[(int64_t)s * Npad + row], tau0, tau1, a.mqc) : 0.0f;
            dsum += (double)p * (double)r;
        }
        out[o] = r;
    }
    if (a.mode == 1) {
        sm[tid] = dsum;
        for (int st = 128; st >= EW; st >>= 1) {
            __syncthreads();
            if (tid < st) sm[tid] += sm[tid + st];
        }
        __syncthreads();
        if (tid < EW) partial[((int64_t)blockIdx.x * EW + tid) * 2] = sm[tid];
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
                                               const uint8_t* __restrict__ live,
                                               const uint8_t* __restrict__ kind, float* alpha) {
    __shared__ double sm[256];
    const int c = blockIdx.x;
    const double pap = sum_parts(partial, nparts, kpad, c, 0, sm);
    if (threadIdx.x == 0)
        alpha[c] = (live[c] && !done[c] && !kind[c]) ? (float)(rz[c] / pap) : 0.0f;
}

// x += a p; r1 = r - a Ap; z1 = minv r1 (FG.cpp:2734-2767), partial z1.r1, r1.r1
__global__ __launch_bounds__(256) void k_update(int kpad, int64_t Npad, const float* __restrict__ alpha,
                                                const uint8_t* __restrict__ done,
                                                const uint8_t* __restrict__ live,
                                                const uint8_t* __restrict__ kind,
                                                const float* __restrict__ minv,
                                                const int32_t* __restrict__ sys, float* X, float* R,
                                                float* Z, const float* __restrict__ P,
                                                const float* __restrict__ AP,
                                                double* __restrict__ partial) {
    __shared__ double sm[2][256];
    const int tid = threadIdx.x;
    const int c = tid % kpad;
    const int rstep = 256 / kpad;
    const bool act = live[c] && !done[c] && !kind[c];
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
                                              const uint8_t* __restrict__ kind, float* beta) {
    __shared__ double sm[256];
    const int c = blockIdx.x;
    const double rzn = sum_parts(partial, nparts, kpad, c, 0, sm);
    const double rr = sum_parts(partial, nparts, kpad, c, 1, sm);
    if (threadIdx.x == 0) {
        if (live[c] && !done[c] && kind[c]) {
            // a product slot finishes with the product it rode along in
            beta[c] = 0.0f;
            iter[c] = 1;
            done[c] = 1;
        } else if (live[c] && !done[c]) {
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
                                                 const uint8_t* __restrict__ kind,
                                                 const float* __restrict__ Z, float* P) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= Npad * kpad) return;
    const int c = (int)(t % kpad);
    if (!live[c] || done[c] || kind[c]) return;
    P[t] = __fadd_rn(Z[t], __fmul_rn(beta[c], P[t]));
}

// Start the listed columns together: column j of the staged block B
// ([k][Nnm]) goes to slot slots[j] with system sys[j] (FG.cpp:2615-2690).
// product != 0 starts GRM-product slots instead: P = b on the system's
// samples, and the slot is done after the next GEMM pass.
__global__ __launch_bounds__(256) void k_cols_init(const float* __restrict__ B, const int32_t* __restrict__ slots,
                                                   const int32_t* __restrict__ sys, int kpad, int64_t Nnm,
                                                   int64_t Npad, const float* __restrict__ minv,
                                                   const uint8_t* __restrict__ mask, float* X, float* R, float* Z,
                                                   float* P, double* __restrict__ partial, int product) {
    __shared__ double sm[2][256];
    const int tid = threadIdx.x;
    const int j = blockIdx.y;
    const int slot = slots[j];
    const int64_t so = (int64_t)sys[j] * Npad;
    const int64_t i = (int64_t)blockIdx.x * VROWS + tid;
    const bool on = i < Nnm && mask[so + i];
    const float r = on ? B[(int64_t)j * Nnm + i] : 0.0f;
    const float z = __fmul_rn(minv[so + i], r);
    const int64_t o = i * kpad + slot;
    X[o] = 0.0f;
    R[o] = r;
    Z[o] = z;
    P[o] = product ? r : z;
    sm[0][tid] = (double)z * (double)r;
    sm[1][tid] = (double)r * (double)r;
    for (int s = 128; s >= 1; s >>= 1) {
        __syncthreads();
        if (tid < s) {
            sm[0][tid] += sm[0][tid + s];
            sm[1][tid] += sm[1][tid + s];
        }
    }
    if (tid == 0) {
        partial[((int64_t)blockIdx.x * kpad + slot) * 2] = sm[0][0];
        partial[((int64_t)blockIdx.x * kpad + slot) * 2 + 1] = sm[1][0];
    }
}

__global__ __launch_bounds__(256) void k_cols_init_finish(const int32_t* __restrict__ slots, int kpad,
                                                          const double* __restrict__ partial, int64_t nparts,
                                                          double* rz, double* sumr2, const int32_t* maxiter,
                                                          const double* tol, int32_t* iter, uint8_t* done,
                                                          int product) {
    __shared__ double sm[256];
    const int slot = slots[blockIdx.x];
    const double a = sum_parts(partial, nparts, kpad, slot, 0, sm);
    const double b = sum_parts(partial, nparts, kpad, slot, 1, sm);
    if (threadIdx.x == 0) {
        rz[slot] = a;
        sumr2[slot] = b;
        iter[slot] = 0;
        done[slot] = product ? 0 : ((b > tol[slot] && 0 < maxiter[slot]) ? 0 : 1);   // FG.cpp:2696
    }
}

__global__ void k_cols_gather(const float* __restrict__ src, const int32_t* __restrict__ slots, int kpad,
                              int64_t Nnm, float* __restrict__ dst) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < Nnm) dst[(int64_t)j * Nnm + i] = src[i * kpad + slots[j]];
}

__global__ void k_gather_col(const float* __restrict__ src, int slot, int kpad, int64_t Nnm,
                             float* __restrict__ dst) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < Nnm) dst[i] = src[i * kpad + slot];
}

// slot results to chosen columns of a column-major buffer: dst[:, cols[j]]
__global__ void k_cols_gather_to(const float* __restrict__ src, const int32_t* __restrict__ slots,
                                 const int32_t* __restrict__ cols, int kpad, int64_t Nnm, int64_t ld,
                                 float* __restrict__ dst) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < Nnm) dst[(int64_t)cols[j] * ld + i] = src[i * kpad + slots[j]];
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

// Deterministic values in [-1, 1) (a 32-bit integer hash of the index), so
// that timing runs see operands with realistic bit activity.
__global__ void k_fill_hash(float* p, int64_t n, uint32_t seed) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t h = (uint32_t)t * 0x9E3779B1u ^ seed;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    p[t] = (float)(int32_t)h * (1.0f / 2147483648.0f);
}

// ------------------------------------------------------ trace probes
// Rademacher probe i, entry r: +1 or -1 from the top bit of a counter-based
// 64-bit hash of (seed, i, r) (splitmix64 finaliser).  The same function is
// restated in numpy on the host, so every executor sees identical probes.
__device__ __forceinline__ uint64_t probe_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_rademacher(float* __restrict__ dst, int64_t ld, int64_t N, uint64_t seed, int64_t i0) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (r >= N) return;
    const uint64_t h = probe_mix(seed ^ probe_mix((uint64_t)(i0 + j) * 0x9E3779B97F4A7C15ull +
                                                  probe_mix((uint64_t)r * 0xC2B2AE3D27D4EB4Full + 0x5851F42Dull)));
    dst[j * ld + r] = (h >> 63) ? -1.0f : 1.0f;
}

// The probes of a fit on a sample subset: entry i of probe i0 + j (the hash
// of the position i within the subset, as probe_vector(seed, i0 + j, n))
// goes to row rows[i] of column j; rows outside [0, N) are skipped.
__global__ void k_rademacher_rows(float* __restrict__ dst, int64_t ld, int64_t N, const int64_t* __restrict__ rows,
                                  int64_t n, uint64_t seed, int64_t i0) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= n) return;
    const int64_t r = rows[i];
    if (r < 0 || r >= N) return;
    const uint64_t h = probe_mix(seed ^ probe_mix((uint64_t)(i0 + j) * 0x9E3779B97F4A7C15ull +
                                                  probe_mix((uint64_t)i * 0xC2B2AE3D27D4EB4Full + 0x5851F42Dull)));
    dst[j * ld + r] = (h >> 63) ? -1.0f : 1.0f;
}

// FP64 dot products of k columns x_j with partner columns: P1[:, j], P2[:, j]
// (pairwise, optional) and Q[:, r] for r < m (shared).  x_j is either a slot
// of an engine block ([Npad][kpad] row-major, slots != nullptr) or column j
// of a column-major buffer.  Workgroup (b, j) sums rows [256 b, 256 b + 256)
// in a fixed tree order into partial[b][j][d]; k_col_dots_finish adds the
// blocks in order, so the result does not depend on the launch.
constexpr int DOT_ROWS = 256;
constexpr int DOT_MAX = 16;   // 2 pairwise + at most 14 shared columns

__global__ __launch_bounds__(256) void k_col_dots(const float* __restrict__ X, int64_t ldx, int kpad,
                                                  const int32_t* __restrict__ slots, int k,
                                                  const float* __restrict__ P1, const float* __restrict__ P2,
                                                  int64_t ldp, const int32_t* __restrict__ pcols,
                                                  const float* __restrict__ Q, int m, int64_t ldq,
                                                  int64_t Nnm, double* __restrict__ partial) {
    __shared__ double sm[256];
    const int tid = threadIdx.x;
    const int j = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * DOT_ROWS + tid;
    const bool on = i < Nnm;
    float x = 0.0f;
    if (on) x = slots ? X[i * kpad + slots[j]] : X[(int64_t)j * ldx + i];
    const int nd = (P1 ? 1 : 0) + (P2 ? 1 : 0) + m;
    double* out = partial + ((int64_t)blockIdx.x * k + j) * nd;
    int d = 0;
    for (int q = -2; q < m; q++) {
        const float* src = q == -2 ? P1 : (q == -1 ? P2 : Q + (int64_t)q * ldq);
        if (q < 0 && !src) continue;
        const int64_t pc = pcols ? pcols[j] : j;   // partner column of x_j
        const float* col = q == -2 ? P1 + pc * ldp : (q == -1 ? P2 + pc * ldp : src);
        sm[tid] = on ? (double)x * (double)col[i] : 0.0;
        for (int s = 128; s >= 1; s >>= 1) {
            __syncthreads();
            if (tid < s) sm[tid] += sm[tid + s];
        }
        __syncthreads();
        if (tid == 0) out[d] = sm[0];
        d++;
    }
}

__global__ void k_col_dots_finish(const double* __restrict__ partial, int64_t nb, int k, int nd,
                                  double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k * nd) return;
    const int j = t / nd, d = t - j * nd;
    double s = 0.0;
    for (int64_t b = 0; b < nb; b++) s += partial[(b * k + j) * nd + d];
    out[t] = s;
}

inline unsigned nblk(int64_t n, int64_t b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

int64_t gemv_max_parts(int64_t Npad) { return Npad / GEMV_ROWS; }
int64_t gemm_max_parts(int64_t Npad) { return Npad / EROWS; }
int64_t vec_max_parts(int64_t Npad, int) { return Npad / VROWS; }

hipError_t launch_grm_gemv(hipStream_t s, const float* S, int64_t Npad, int64_t Nnm, const float* in,
                           float* out, SigmaArgs a, double* partial, int64_t* nparts, const GemvPlan* tri) {
    if (tri) {
        if (tri->Npad != Npad || !tri->items || Npad % TV_ROWS != 0) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_gemv_tri, dim3((unsigned)tri->nitems), dim3(512), 0, s, S, Npad, in, tri->items,
                           tri->y1p, tri->y2p);
        const int64_t nb = Npad / TV_ROWS;
        hipLaunchKernelGGL(k_gemv_tri_epi, dim3((unsigned)nb), dim3(256), 0, s, tri->y1p, tri->y2p, tri->rtab, Npad,
                           Nnm, in, out, a, partial);
        if (nparts) *nparts = nb;
        return hipGetLastError();
    }
    const int64_t nb = Npad / GEMV_ROWS;
    hipLaunchKernelGGL(k_gemv, dim3((unsigned)nb), dim3(256), 0, s, S, Npad, Nnm, in, out, a, partial);
    if (nparts) *nparts = nb;
    return hipGetLastError();
}

bool gemm_config_valid(GemmConfig g) {
    return (g.rt == 1 || g.rt == 2) && (g.pd == 1 || g.pd == 3) &&
           (g.nsplit == 1 || g.nsplit == 2 || g.nsplit == 4 || g.nsplit == 8) &&
           (g.mode == GEMM_FP32 || g.mode == GEMM_BF16X3 || g.mode == GEMM_BF16X3_TRI);
}

void gemv_plan_free(GemvPlan& t) {
    if (t.items) (void)hipFree(t.items);
    if (t.rtab) (void)hipFree(t.rtab);
    if (t.y1p) (void)hipFree(t.y1p);
    if (t.y2p) (void)hipFree(t.y2p);
    t = GemvPlan{};
}

hipError_t gemv_plan_build(GemvPlan& t, int64_t Npad, int ch) {
    if (t.items && t.Npad == Npad && t.ch == ch) return hipSuccess;
    gemv_plan_free(t);
    if (ch < 1 || Npad <= 0 || Npad % TV_ROWS != 0) return hipErrorInvalidValue;
    const int64_t nr = Npad / TV_ROWS;     // row sets, and 256-column chunks
    std::vector<int4> items;
    std::vector<int2> rtab((size_t)nr);
    int64_t slot = 0;
    for (int64_t R = 0; R < nr; R++) {
        // the R chunks left of the diagonal block cut into k near-equal items,
        // then the diagonal block
        const int64_t k = (R + ch - 1) / ch;
        rtab[(size_t)R] = make_int2((int)slot, (int)(k + 1));
        for (int64_t i = 0; i < k; i++)
            items.push_back(make_int4((int)R, (int)(R * i / k), (int)(R * (i + 1) / k), (int)slot++));
        items.push_back(make_int4((int)R, (int)R, (int)(R + 1), (int)slot++));
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const int4& a, const int4& b) { return a.z - a.y > b.z - b.y; });
    const int64_t y2 = triv_y2_first(nr, nr);
    t.Npad = Npad;
    t.ch = ch;
    t.nitems = (int64_t)items.size();
    t.nr = nr;
    t.y1_slots = slot;
    t.y2_slots = y2;
    hipError_t e = hipMalloc(&t.items, sizeof(int4) * items.size());
    if (e == hipSuccess) e = hipMalloc(&t.rtab, sizeof(int2) * rtab.size());
    if (e == hipSuccess) e = hipMalloc(&t.y1p, sizeof(float) * (size_t)(slot * TV_ROWS));
    if (e == hipSuccess) e = hipMalloc(&t.y2p, sizeof(float) * (size_t)(std::max<int64_t>(y2, 1) * TV_COLS));
    if (e == hipSuccess) e = hipMemcpy(t.items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t.rtab, rtab.data(), sizeof(int2) * rtab.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) gemv_plan_free(t);
    return e;
}

void tri_plan_free(TriPlan& t) {
    if (t.items) (void)hipFree(t.items);
    if (t.rtab) (void)hipFree(t.rtab);
    if (t.y1p) (void)hipFree(t.y1p);
    if (t.y2p) (void)hipFree(t.y2p);
    t = TriPlan{};
}

// The dispatch order of the triangle's items.  A launch's workgroups go to
// the eight XCDs in turn (workgroup b to XCD b % 8), one per CU at a time,
// 32 CUs per XCD, and each XCD has its own L2.  Items of different row sets
// over the same chunk range read the same staged P pieces, so full-length
// items are grouped by range into tiles of 32; a round of eight tiles takes
// 256 consecutive positions, tile x at positions x, x + 8, ...: the 32
// workgroups of one XCD then walk the same chunks together and share their
// pieces in its L2.  The rest (partial tiles, shorter and diagonal items)
// follows, longest first, to fill the tail.
static std::vector<int4> tri_dispatch_order(const std::vector<int4>& items, int nwv, int ch) {
    constexpr int XCD = 8, PER_XCD = 32;
    std::vector<std::vector<int4>> groups;
    std::vector<int4> rest;
    for (const int4& it : items) {
        const bool left = it.y < nwv * it.x;         // left of the row set's diagonal block
        if (left && it.z - it.y == ch) {
            const size_t g = (size_t)(it.y / ch);    // full items start at multiples of ch
            if (groups.size() <= g) groups.resize(g + 1);
            groups[g].push_back(it);
        } else {
            rest.push_back(it);
        }
    }
    std::vector<std::vector<int4>> tiles;
    for (const auto& v : groups) {
        size_t i = 0;
        for (; i + PER_XCD <= v.size(); i += PER_XCD) tiles.emplace_back(v.begin() + i, v.begin() + i + PER_XCD);
        rest.insert(rest.end(), v.begin() + i, v.end());
    }
    const size_t rounds = tiles.size() / XCD;
    for (size_t t = rounds * XCD; t < tiles.size(); t++) rest.insert(rest.end(), tiles[t].begin(), tiles[t].end());
    std::vector<int4> out;
    out.reserve(items.size());
    for (size_t r = 0; r < rounds; r++)
        for (int i = 0; i < PER_XCD; i++)
            for (int x = 0; x < XCD; x++) out.push_back(tiles[r * XCD + x][(size_t)i]);
    std::stable_sort(rest.begin(), rest.end(), [](const int4& a, const int4& b) { return a.z - a.y > b.z - b.y; });
    out.insert(out.end(), rest.begin(), rest.end());
    return out;
}

hipError_t tri_plan_build(TriPlan& t, int64_t Npad, int nwv, int ch) {
    if (t.items && t.Npad == Npad && t.nwv == nwv && t.ch == ch) return hipSuccess;
    tri_plan_free(t);
    if ((nwv != 8 && nwv != 12) || ch < 1 || Npad <= 0 || Npad % 256 != 0) return hipErrorInvalidValue;
    const int64_t nch = Npad / 32, nr = (nch + nwv - 1) / nwv;
    std::vector<int4> items;
    std::vector<int2> rtab((size_t)nr);
    int64_t slot = 0;
    for (int64_t R = 0; R < nr; R++) {
        // chunks left of the diagonal block, cut at multiples of ch (the last
        // piece shorter), then the diagonal block itself; slots in chunk order
        const int64_t off = nwv * R, first = slot;
        for (int64_t a = 0; a < off; a += ch)
            items.push_back(make_int4((int)R, (int)a, (int)std::min<int64_t>(a + ch, off), (int)slot++));
        items.push_back(make_int4((int)R, (int)off, (int)std::min(off + nwv, nch), (int)slot++));
        rtab[(size_t)R] = make_int2((int)first, (int)(slot - first));
    }
    items = tri_dispatch_order(items, nwv, ch);
    const int64_t y2 = nwv == 8 ? tri_y2_first<8>(nch, nr) : tri_y2_first<12>(nch, nr);
    t.Npad = Npad;
    t.nwv = nwv;
    t.ch = ch;
    t.nitems = (int64_t)items.size();
    t.nr = nr;
    t.y1_slots = slot;
    t.y2_slots = y2;
    hipError_t e = hipMalloc(&t.items, sizeof(int4) * items.size());
    if (e == hipSuccess) e = hipMalloc(&t.rtab, sizeof(int2) * rtab.size());
    if (e == hipSuccess) e = hipMalloc(&t.y1p, sizeof(float) * (size_t)(slot * 32 * nwv * EW));
    if (e == hipSuccess) e = hipMalloc(&t.y2p, sizeof(float) * (size_t)(std::max<int64_t>(y2, 1) * 32 * EW));
    if (e == hipSuccess) e = hipMemcpy(t.items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t.rtab, rtab.data(), sizeof(int2) * rtab.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) tri_plan_free(t);
    return e;
}

namespace {
template <int NWV, int PD, int NQ, int SP>
void launch_t16(hipStream_t s, dim3 grid, const float* S, int64_t Npad, const u32x4* pb, const float* in,
                const TriPlan& t, ColMap32 cm, float* ypart) {
    hipLaunchKernelGGL((k_gemm_t16<NWV, PD, NQ, SP>), grid, dim3(64 * NWV), 0, s, S, Npad, pb, in, t.items, t.y1p,
                       t.y2p, cm);
    hipLaunchKernelGGL((k_tri_reduce<NWV, 4 * NQ>), dim3((unsigned)(Npad * 4 * NQ / 256)), dim3(256), 0, s, t.y1p,
                       t.y2p, t.rtab, Npad, cm, ypart);
}

// the lower-triangle product over the launch's 64 columns, or the 16 or 32
// block columns narrow lists (nullptr and halves 1 or 2: the half's 32)
template <int NWV, int SP>
hipError_t launch_tri(hipStream_t s, const float* S, int64_t Npad, const float* in, float* ypart, int pd,
                      int halves, const NarrowCols* narrow, const TriPlan& t) {
    const int64_t G = Npad / 8;
    u32x4* pb = reinterpret_cast<u32x4*>(ypart + (int64_t)GSPLIT_MAX * Npad * EW);
    hipLaunchKernelGGL(k_split_p, dim3((unsigned)(G * EW / 256)), dim3(256), 0, s, in, G, pb);
    const dim3 grid((unsigned)t.nitems);
    const bool deep = pd == 3;
    const NarrowCols half = NarrowCols::tile(32, halves == 2 ? 32 : 0);
    if (!narrow && halves != 3) narrow = &half;
    ColMap32 cm{};
    if (narrow)
        for (int k = 0; k < narrow->width; k++) cm.w[k / 4] |= (uint32_t)narrow->col[k] << (8 * (k % 4));
    if (narrow && narrow->width == 16) {
        if (deep) launch_t16<NWV, 2, 1, SP>(s, grid, S, Npad, pb, in, t, cm, ypart);
        else launch_t16<NWV, 1, 1, SP>(s, grid, S, Npad, pb, in, t, cm, ypart);
    } else if (narrow) {
        if (deep) launch_t16<NWV, 2, 2, SP>(s, grid, S, Npad, pb, in, t, cm, ypart);
        else launch_t16<NWV, 1, 2, SP>(s, grid, S, Npad, pb, in, t, cm, ypart);
    } else {
        if (deep) launch_t16<NWV, 2, 4, SP>(s, grid, S, Npad, pb, in, t, cm, ypart);
        else launch_t16<NWV, 1, 4, SP>(s, grid, S, Npad, pb, in, t, cm, ypart);
    }
    return hipGetLastError();
}
}  // namespace

hipError_t launch_grm_gemm(hipStream_t s, const float* S, int64_t Npad, int64_t Nnm, const float* in,
                           float* out, SigmaArgs a, double* partial, int64_t* nparts, float* ypart, GemmConfig g,
                           int halves, const TriPlan* tri, bool s16, const NarrowCols* narrow) {
    // Npad is a multiple of 256 = 128 * RT for both RT; a slice holds at least
    // one superchunk when Npad / GKS >= nsplit
    if (!gemm_config_valid(g) || Npad % 256 != 0 || halves < 1 || halves > 3) return hipErrorInvalidValue;
    if (g.mode == GEMM_BF16X3_TRI) {
        // the lower-triangle product sums its partials into slice 0
        if (!tri || tri->Npad != Npad || !tri->items) return hipErrorInvalidValue;
        hipError_t e;
        if (narrow) {
            // 16 or 32 distinct block columns
            if (narrow->width != 16 && narrow->width != 32) return hipErrorInvalidValue;
            uint64_t seen = 0;
            for (int k = 0; k < narrow->width; k++) {
                const int col = narrow->col[k];
                if (col >= EW || (seen >> col & 1)) return hipErrorInvalidValue;
                seen |= 1ull << col;
            }
        }
        if (s16)
            e = tri->nwv == 8 ? launch_tri<8, 2>(s, S, Npad, in, ypart, g.pd, halves, narrow, *tri)
                              : launch_tri<12, 2>(s, S, Npad, in, ypart, g.pd, halves, narrow, *tri);
        else
            e = tri->nwv == 8 ? launch_tri<8, 3>(s, S, Npad, in, ypart, g.pd, halves, narrow, *tri)
                              : launch_tri<12, 3>(s, S, Npad, in, ypart, g.pd, halves, narrow, *tri);
        if (e != hipSuccess) return e;
        const int64_t nb = Npad / EROWS;
        hipLaunchKernelGGL(k_gemm_epi, dim3((unsigned)nb), dim3(256), 0, s, ypart, 1, Npad, Nnm, in, out, a,
                           partial);
        if (nparts) *nparts = nb;
        return hipGetLastError();
    }
    const int nsplit = (int)std::min<int64_t>(g.nsplit, Npad / GKS);
    const dim3 grid((unsigned)(Npad / (128 * g.rt)), nsplit);
    if (g.mode == GEMM_BF16X3) {
        // split P into the three BF16 pieces behind the slices in ypart
        const int64_t G = Npad / 8;
        u32x4* pb = reinterpret_cast<u32x4*>(ypart + (int64_t)GSPLIT_MAX * Npad * EW);
        hipLaunchKernelGGL(k_split_p, dim3((unsigned)(G * EW / 256)), dim3(256), 0, s, in, G, pb);
        // rt 1: twelve-wave workgroups (pd 3 -> two chunks ahead, the register
        // limit at three waves per SIMD); rt 2: four waves of two tiles, at the
        // most waves per SIMD each prefetch depth allows without spills
        const dim3 grid12((unsigned)((Npad + 383) / 384), nsplit);
        if (g.rt == 1 && halves == 3) {
            if (g.pd == 3)
                hipLaunchKernelGGL((k_gemm_x3w<2, true>), grid12, dim3(768), 0, s, S, Npad, pb, ypart, halves);
            else
                hipLaunchKernelGGL((k_gemm_x3w<1, true>), grid12, dim3(768), 0, s, S, Npad, pb, ypart, halves);
        } else if (g.rt == 1 && g.pd == 3)
            hipLaunchKernelGGL((k_gemm_x3w<2, false>), grid12, dim3(768), 0, s, S, Npad, pb, ypart, halves);
        else if (g.rt == 1)
            hipLaunchKernelGGL((k_gemm_x3w<1, false>), grid12, dim3(768), 0, s, S, Npad, pb, ypart, halves);
        else if (g.pd == 3)
            hipLaunchKernelGGL((k_gemm_x3<2, 3, 1>), grid, dim3(256), 0, s, S, Npad, pb, ypart, halves);
        else
            hipLaunchKernelGGL((k_gemm_x3<2, 1, 2>), grid, dim3(256), 0, s, S, Npad, pb, ypart, halves);
    } else if (g.rt == 2 && g.pd == 3)
        hipLaunchKernelGGL((k_gemm_part<2, 3>), grid, dim3(256), 0, s, S, Npad, in, ypart, halves);
    else if (g.rt == 2)
        hipLaunchKernelGGL((k_gemm_part<2, 1>), grid, dim3(256), 0, s, S, Npad, in, ypart, halves);
    else if (g.pd == 3)
        hipLaunchKernelGGL((k_gemm_part<1, 3>), grid, dim3(256), 0, s, S, Npad, in, ypart, halves);
    else
        hipLaunchKernelGGL((k_gemm_part<1, 1>), grid, dim3(256), 0, s, S, Npad, in, ypart, halves);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t nb = Npad / EROWS;
    hipLaunchKernelGGL(k_gemm_epi, dim3((unsigned)nb), dim3(256), 0, s, ypart, nsplit, Npad, Nnm, in, out, a,
                       partial);
    if (nparts) *nparts = nb;
    return hipGetLastError();
}

// the split-K slices, then (three-piece mode) P's BF16 pieces: 3 x 2 bytes
// per element, 1.5 floats, rounded up to 2
int64_t gemm_ypart_floats(int64_t Npad) { return (int64_t)(GSPLIT_MAX + 2) * Npad * EW; }

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
                        const uint8_t* done, const uint8_t* live, const uint8_t* kind, float* alpha) {
    hipLaunchKernelGGL(k_alpha, dim3((unsigned)kpad), dim3(256), 0, s, kpad, partial, nparts, rz, done,
                       live, kind, alpha);
    return hipGetLastError();
}

hipError_t launch_update(hipStream_t s, int kpad, int64_t Nnm, int64_t Npad, const float* alpha,
                         const uint8_t* done, const uint8_t* live, const uint8_t* kind, const float* minv,
                         const int32_t* sys, float* X, float* R, float* Z, const float* P, const float* AP,
                         double* partial, int64_t* nparts) {
    (void)Nnm;
    const int64_t nb = Npad / VROWS;
    hipLaunchKernelGGL(k_update, dim3((unsigned)nb), dim3(256), 0, s, kpad, Npad, alpha, done, live, kind, minv,
                       sys, X, R, Z, P, AP, partial);
    if (nparts) *nparts = nb;
    return hipGetLastError();
}

hipError_t launch_beta(hipStream_t s, int kpad, const double* partial, int64_t nparts, double* rz,
                       double* sumr2, const int32_t* maxiter, const double* tol, int32_t* iter,
                       uint8_t* done, const uint8_t* live, const uint8_t* kind, float* beta) {
    hipLaunchKernelGGL(k_beta, dim3((unsigned)kpad), dim3(256), 0, s, kpad, partial, nparts, rz, sumr2,
                       maxiter, tol, iter, done, live, kind, beta);
    return hipGetLastError();
}

hipError_t launch_pupdate(hipStream_t s, int kpad, int64_t Nnm, int64_t Npad, const float* beta,
                          const uint8_t* done, const uint8_t* live, const uint8_t* kind, const float* Z, float* P) {
    (void)Nnm;
    hipLaunchKernelGGL(k_pupdate, dim3(nblk(Npad * kpad, 256)), dim3(256), 0, s, kpad, Npad, beta, done,
                       live, kind, Z, P);
    return hipGetLastError();
}

hipError_t launch_cols_init(hipStream_t s, const float* B, int k, const int32_t* slots, const int32_t* sys,
                            int kpad, int64_t Nnm, int64_t Npad, const float* minv, const uint8_t* mask, float* X,
                            float* R, float* Z, float* P, double* partial, int64_t* nparts, int product) {
    const int64_t nb = Npad / VROWS;
    hipLaunchKernelGGL(k_cols_init, dim3((unsigned)nb, (unsigned)k), dim3(256), 0, s, B, slots, sys, kpad, Nnm, Npad,
                       minv, mask, X, R, Z, P, partial, product);
    if (nparts) *nparts = nb;
    return hipGetLastError();
}

hipError_t launch_cols_init_finish(hipStream_t s, int k, const int32_t* slots, int kpad, const double* partial,
                                   int64_t nparts, double* rz, double* sumr2, const int32_t* maxiter,
                                   const double* tol, int32_t* iter, uint8_t* done, int product) {
    hipLaunchKernelGGL(k_cols_init_finish, dim3((unsigned)k), dim3(256), 0, s, slots, kpad, partial, nparts, rz,
                       sumr2, maxiter, tol, iter, done, product);
    return hipGetLastError();
}

hipError_t launch_cols_gather(hipStream_t s, const float* src, int k, const int32_t* slots, int kpad, int64_t Nnm,
                              float* dst) {
    hipLaunchKernelGGL(k_cols_gather, dim3(nblk(Nnm, 256), (unsigned)k), dim3(256), 0, s, src, slots, kpad, Nnm, dst);
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

hipError_t launch_rademacher_rows(hipStream_t s, float* dst, int64_t ld, int64_t N, const int64_t* rows, int64_t n,
                                  uint64_t seed, int64_t i0, int k) {
    if (k < 1 || n < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_rademacher_rows, dim3(nblk(n, 256), (unsigned)k), dim3(256), 0, s, dst, ld, N, rows, n,
                       seed, i0);
    return hipGetLastError();
}

hipError_t launch_rademacher(hipStream_t s, float* dst, int64_t ld, int64_t N, uint64_t seed, int64_t i0, int k) {
    if (k <= 0 || N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rademacher, dim3(nblk(N, 256), (unsigned)k), dim3(256), 0, s, dst, ld, N, seed, i0);
    return hipGetLastError();
}

int col_dots_max() { return DOT_MAX; }
int64_t col_dots_partial_doubles(int64_t Npad, int k) { return (Npad / DOT_ROWS + 1) * (int64_t)k * DOT_MAX; }

hipError_t launch_cols_gather_to(hipStream_t s, const float* src, int k, const int32_t* slots, const int32_t* cols,
                                 int kpad, int64_t Nnm, int64_t ld, float* dst) {
    hipLaunchKernelGGL(k_cols_gather_to, dim3(nblk(Nnm, 256), (unsigned)k), dim3(256), 0, s, src, slots, cols, kpad,
                       Nnm, ld, dst);
    return hipGetLastError();
}

hipError_t launch_col_dots(hipStream_t s, const float* X, int64_t ldx, int kpad, const int32_t* slots, int k,
                           const float* P1, const float* P2, int64_t ldp, const int32_t* pcols, const float* Q,
                           int m, int64_t ldq, int64_t Nnm, double* partial, double* out) {
    const int nd = (P1 ? 1 : 0) + (P2 ? 1 : 0) + m;
    if (k <= 0 || nd <= 0 || nd > DOT_MAX) return hipErrorInvalidValue;
    const int64_t nb = (Nnm + DOT_ROWS - 1) / DOT_ROWS;
    hipLaunchKernelGGL(k_col_dots, dim3((unsigned)nb, (unsigned)k), dim3(256), 0, s, X, ldx, kpad, slots, k, P1, P2,
                       ldp, pcols, Q, m, ldq, Nnm, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_col_dots_finish, dim3(nblk((int64_t)k * nd, 64)), dim3(64), 0, s, partial, nb, k, nd, out);
    return hipGetLastError();
}

hipError_t launch_fill_hash(hipStream_t s, float* p, int64_t n, uint32_t seed) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_hash, dim3(nblk(n, 256)), dim3(256), 0, s, p, n, seed);
    return hipGetLastError();
}

}  // namespace saige
