// GRM SYRK on CDNA4 matrix cores: S (+)= Zp Zp^T for one marker panel.
//
// Zp is the standardised panel, [Npad][ldz] row-major (sample rows, marker
// columns), so both MFMA operands are K-contiguous rows of the same matrix.
// Each workgroup computes one 128x128 tile of the lower triangle (tile row
// >= tile column) and writes it and its mirror into the full row-major S.
// Tile order: the host lists the triangle's tiles grouped into 8x8-tile
// super-blocks, and the launch index is remapped so that each XCD (launch
// order round-robins over the 8 XCDs) walks a contiguous run of that list.
// The ~100 tiles an XCD holds at once then span one or two super-blocks,
// i.e. 16-24 distinct 128-row panels of Zp, whose current K chunk fits in
// that XCD's 4 MB L2.
//
// Operand staging: global -> registers (prefetch of the next K chunk while
// the current one is multiplied) -> LDS -> MFMA fragments.
#include <algorithm>

#include "saige_internal.hpp"

namespace saige {

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;   // native vector: stays in VGPRs
typedef __attribute__((ext_vector_type(16))) float f32x16;

constexpr int BT = 128;  // tile edge


// Launch index -> triangle tile index, grouping consecutive tiles per XCD.
__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nblk) {
    const int64_t nx = 8;
    const int64_t x = bid % nx, loc = bid / nx;
    const int64_t q = nblk / nx, rem = nblk % nx;
    return (x < rem) ? x * (q + 1) + loc : rem * (q + 1) + (x - rem) * q + loc;
}

// Accumulator layout of a 32x32 MFMA block: lane l, register r holds
// row 8*(r/4) + 4*(l/32) + (r%4), column l%32.
__device__ __forceinline__ void store_tile(float* S, int64_t ldS, int64_t bi, int64_t bj, int wr,
                                           int wc, int lane, const f32x16 (&acc)[2][2],
                                           int accumulate) {
    const bool mirror = bi != bj;
#pragma unroll
    for (int mi = 0; mi < 2; mi++) {
#pragma unroll
        for (int ni = 0; ni < 2; ni++) {
            const int64_t col = bj * BT + wc * 64 + ni * 32 + (lane & 31);
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int64_t row0 = bi * BT + wr * 64 + mi * 32 + 8 * g + 4 * (lane >> 5);
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; e++) v[e] = acc[mi][ni][4 * g + e];
                if (accumulate) {
#pragma unroll
                    for (int e = 0; e < 4; e++) v[e] += S[(row0 + e) * ldS + col];
                }
#pragma unroll
                for (int e = 0; e < 4; e++) S[(row0 + e) * ldS + col] = v[e];
                if (mirror) {
                    float4* dst = reinterpret_cast<float4*>(S + col * ldS + row0);
                    *dst = make_float4(v[0], v[1], v[2], v[3]);
                }
            }
        }
    }
}

// ------------------------------------------------------------------ BF16
constexpr int BK16 = 64;            // markers per K chunk
constexpr int P16 = BK16 + 8;       // LDS row pitch (bf16 elements): 144 B, conflict-free b128 reads

__global__ __launch_bounds__(256) void k_syrk_bf16(const __hip_bfloat16* __restrict__ Zp, int64_t ldz,
                                                   int64_t kdim, const int2* __restrict__ tiles, int64_t nblk,
                                                   float* __restrict__ S, int64_t ldS, int accumulate) {
    __shared__ __attribute__((aligned(16))) __hip_bfloat16 sA[BT * P16];
    __shared__ __attribute__((aligned(16))) __hip_bfloat16 sB[BT * P16];
    const int2 t = tiles[xcd_remap(blockIdx.x, nblk)];
    const int64_t bi = t.x, bj = t.y;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;

    // global chunk mapping: 16-byte chunk c = tid + 256 q : row c/8, k (c%8)*8
    const int lrow = tid >> 3, lk = (tid & 7) * 8;
    const __hip_bfloat16* gA = Zp + (bi * BT + lrow) * ldz + lk;
    const __hip_bfloat16* gB = Zp + (bj * BT + lrow) * ldz + lk;
    u32x4 ra[4], rb[4];
    const int64_t nk = kdim / BK16;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        ra[q] = *reinterpret_cast<const u32x4*>(gA + (int64_t)(32 * q) * ldz);
        rb[q] = *reinterpret_cast<const u32x4*>(gB + (int64_t)(32 * q) * ldz);
    }
    f32x16 acc[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[mi][ni][r] = 0.0f;

    for (int64_t kc = 0; kc < nk; kc++) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; q++) {
            *reinterpret_cast<u32x4*>(sA + (lrow + 32 * q) * P16 + lk) = ra[q];
            *reinterpret_cast<u32x4*>(sB + (lrow + 32 * q) * P16 + lk) = rb[q];
        }
        __syncthreads();
        {
            // prefetch the next chunk (the last iteration re-reads its own chunk:
            // unconditional loads keep ra/rb in registers)
            const int64_t off = ((kc + 1 < nk) ? kc + 1 : kc) * BK16;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                ra[q] = *reinterpret_cast<const u32x4*>(gA + (int64_t)(32 * q) * ldz + off);
                rb[q] = *reinterpret_cast<const u32x4*>(gB + (int64_t)(32 * q) * ldz + off);
            }
        }
#pragma unroll
        for (int ks = 0; ks < BK16 / 16; ks++) {
            const int kk = ks * 16 + (lane >> 5) * 8;
            bf16x8 fa[2], fb[2];
#pragma unroll
            for (int mi = 0; mi < 2; mi++)
                fa[mi] = *reinterpret_cast<const bf16x8*>(sA + (wr * 64 + mi * 32 + (lane & 31)) * P16 + kk);
#pragma unroll
            for (int ni = 0; ni < 2; ni++)
                fb[ni] = *reinterpret_cast<const bf16x8*>(sB + (wc * 64 + ni * 32 + (lane & 31)) * P16 + kk);
#pragma unroll
            for (int mi = 0; mi < 2; mi++)
#pragma unroll
                for (int ni = 0; ni < 2; ni++)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[mi], fb[ni], acc[mi][ni], 0, 0, 0);
        }
    }
    store_tile(S, ldS, bi, bj, wr, wc, lane, acc, accumulate);
}

// ------------------------------------------------------------------ FP32
constexpr int BK32 = 32;            // markers per K chunk
constexpr int P32 = BK32 + 4;       // LDS row pitch (floats)
// LDS element (row, k) lives at row*P32 + (k&1)*16 + (k>>1), so lane l's
// k = 2s + l/32 sequence over s is contiguous (four b128 reads per chunk).

__global__ __launch_bounds__(256) void k_syrk_f32(const float* __restrict__ Zp, int64_t ldz, int64_t kdim,
                                                  const int2* __restrict__ tiles, int64_t nblk,
                                                  float* __restrict__ S, int64_t ldS, int accumulate) {
    __shared__ __attribute__((aligned(16))) float sA[BT * P32];
    __shared__ __attribute__((aligned(16))) float sB[BT * P32];
    const int2 t = tiles[xcd_remap(blockIdx.x, nblk)];
    const int64_t bi = t.x, bj = t.y;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;

    // 16-byte chunk c = tid + 256 q: row c/8, k (c%8)*4
    const int lrow = tid >> 3, lk = (tid & 7) * 4;
    const float* gA = Zp + (bi * BT + lrow) * ldz + lk;
    const float* gB = Zp + (bj * BT + lrow) * ldz + lk;
    float4 ra[4], rb[4];
    const int64_t nk = kdim / BK32;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        ra[q] = *reinterpret_cast<const float4*>(gA + (int64_t)(32 * q) * ldz);
        rb[q] = *reinterpret_cast<const float4*>(gB + (int64_t)(32 * q) * ldz);
    }
    f32x16 acc[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[mi][ni][r] = 0.0f;

    const int hk = lk >> 1;  // k/2 of the chunk's first element
    for (int64_t kc = 0; kc < nk; kc++) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float* a = sA + (lrow + 32 * q) * P32;
            float* b = sB + (lrow + 32 * q) * P32;
            *reinterpret_cast<float2*>(a + hk) = make_float2(ra[q].x, ra[q].z);
            *reinterpret_cast<float2*>(a + 16 + hk) = make_float2(ra[q].y, ra[q].w);
            *reinterpret_cast<float2*>(b + hk) = make_float2(rb[q].x, rb[q].z);
            *reinterpret_cast<float2*>(b + 16 + hk) = make_float2(rb[q].y, rb[q].w);
        }
        __syncthreads();
        {
            const int64_t off = ((kc + 1 < nk) ? kc + 1 : kc) * BK32;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                ra[q] = *reinterpret_cast<const float4*>(gA + (int64_t)(32 * q) * ldz + off);
                rb[q] = *reinterpret_cast<const float4*>(gB + (int64_t)(32 * q) * ldz + off);
            }
        }
        const int h = (lane >> 5) * 16;
#pragma unroll
        for (int sq = 0; sq < 4; sq++) {
            float4 fa[2], fb[2];
#pragma unroll
            for (int mi = 0; mi < 2; mi++)
                fa[mi] = *reinterpret_cast<const float4*>(sA + (wr * 64 + mi * 32 + (lane & 31)) * P32 + h + 4 * sq);
#pragma unroll
            for (int ni = 0; ni < 2; ni++)
                fb[ni] = *reinterpret_cast<const float4*>(sB + (wc * 64 + ni * 32 + (lane & 31)) * P32 + h + 4 * sq);
#pragma unroll
            for (int e = 0; e < 4; e++) {
#pragma unroll
                for (int mi = 0; mi < 2; mi++)
#pragma unroll
                    for (int ni = 0; ni < 2; ni++)
                        acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[mi][e], fb[ni][e], acc[mi][ni], 0, 0, 0);
            }
        }
    }
    store_tile(S, ldS, bi, bj, wr, wc, lane, acc, accumulate);
}

}  // namespace

int syrk_kchunk_f32() { return BK32; }
int syrk_kchunk_bf16() { return BK16; }

int syrk_tile() { return BT; }

// Lower-triangle tiles (bi >= bj) of an nt x nt tile grid, grouped into
// g x g super-blocks (super-block row >= column, each walked row by row).
std::vector<int2> syrk_tile_order(int64_t nt, int64_t g) {
    std::vector<int2> out;
    out.reserve((size_t)(nt * (nt + 1) / 2));
    const int64_t ns = (nt + g - 1) / g;
    for (int64_t SI = 0; SI < ns; SI++)
        for (int64_t SJ = 0; SJ <= SI; SJ++)
            for (int64_t bi = SI * g; bi < std::min(nt, (SI + 1) * g); bi++)
                for (int64_t bj = SJ * g; bj < std::min(bi + 1, (SJ + 1) * g); bj++)
                    out.push_back(make_int2((int)bi, (int)bj));
    return out;
}

hipError_t launch_syrk_bf16(hipStream_t s, const __hip_bfloat16* Zp, int64_t Npad, int64_t ldz,
                            int64_t kdim, const int2* tiles, float* S, int accumulate) {
    const int64_t nt = Npad / BT;
    const int64_t nblk = nt * (nt + 1) / 2;
    hipLaunchKernelGGL(k_syrk_bf16, dim3((unsigned)nblk), dim3(256), 0, s, Zp, ldz, kdim, tiles, nblk, S, Npad,
                       accumulate);
    return hipGetLastError();
}

hipError_t launch_syrk_f32(hipStream_t s, const float* Zp, int64_t Npad, int64_t ldz, int64_t kdim,
                           const int2* tiles, float* S, int accumulate) {
    const int64_t nt = Npad / BT;
    const int64_t nblk = nt * (nt + 1) / 2;
    hipLaunchKernelGGL(k_syrk_f32, dim3((unsigned)nblk), dim3(256), 0, s, Zp, ldz, kdim, tiles, nblk, S, Npad,
                       accumulate);
    return hipGetLastError();
}

}  // namespace saige
