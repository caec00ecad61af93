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
                if (accumulate > 0) {        // S += tile
#pragma unroll
                    for (int e = 0; e < 4; e++) v[e] = S[(row0 + e) * ldS + col] + v[e];
                } else if (accumulate < 0) { // S -= tile (LOCO: remove one chromosome's markers)
#pragma unroll
                    for (int e = 0; e < 4; e++) v[e] = S[(row0 + e) * ldS + col] - v[e];
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

// ------------------------------------------------------- BF16, 2-bit in
// Same tile and MFMA loop as k_syrk_bf16, but the operands arrive as 2-bit
// genotypes and are standardised in registers: 8x fewer bytes per element
// than a BF16 panel, so the K loop is no longer bound by L2-miss traffic.
//
// Input layout (built by k_pack_tiles): for each 128-sample panel p and
// 64-marker chunk c, a 2 KB block [b][m]: byte b (samples 4b..4b+3 of the
// panel, 2 bits each, low bits first) of marker m.  Code 0b01 marks padding
// (samples >= Nnm, markers >= Mqc); the internal codes are 00/10/11.
// lut64[j]: four BF16 values of marker j indexed by the internal code:
// 00 -> L[2], 01 -> 0, 10 -> L[1], 11 -> L[0].
//
// Thread t decodes byte b = t/8 of markers 8*(t%8)..+7: four sample rows of
// eight consecutive markers, i.e. four 16-byte LDS rows of the operand.
__device__ __forceinline__ u32x4 decode_row(const uint64_t (&lut)[8], const uint8_t (&byte)[8], int j) {
    u32x4 r;
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const unsigned c0 = (byte[2 * h] >> (2 * j)) & 3u;
        const unsigned c1 = (byte[2 * h + 1] >> (2 * j)) & 3u;
        const unsigned v0 = (unsigned)(lut[2 * h] >> (16 * c0)) & 0xFFFFu;
        const unsigned v1 = (unsigned)(lut[2 * h + 1] >> (16 * c1)) & 0xFFFFu;
        r[h] = v0 | (v1 << 16);
    }
    return r;
}

// The same row as decode_row, by byte permutes: lo[e] / hi[e] hold the low /
// high bytes of marker e's four BF16 table entries (entry c in byte c), so
// the 8-byte pair table {lo[2k+1], lo[2k]} is indexed by c_even and
// 4 + c_odd.  (byte >> 2j) & 3 of all eight markers at once gives the
// selectors; four permutes pick the low and high bytes of two marker pairs
// and two more interleave them into packed BF16 pairs.
__device__ __forceinline__ u32x4 decode_row_perm(const uint32_t (&lo)[8], const uint32_t (&hi)[8], uint64_t bytes,
                                                 int j) {
    const uint64_t d = (bytes >> (2 * j)) & 0x0303030303030303ull;
    u32x4 r;
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const uint32_t sel = (uint32_t)(d >> (32 * q)) + 0x04000400u;   // [c0, c1+4, c2, c3+4]
        const uint32_t a01 = __builtin_amdgcn_perm(lo[4 * q + 1], lo[4 * q], sel);
        const uint32_t b01 = __builtin_amdgcn_perm(hi[4 * q + 1], hi[4 * q], sel);
        const uint32_t a23 = __builtin_amdgcn_perm(lo[4 * q + 3], lo[4 * q + 2], sel);
        const uint32_t b23 = __builtin_amdgcn_perm(hi[4 * q + 3], hi[4 * q + 2], sel);
        r[2 * q] = __builtin_amdgcn_perm(b01, a01, 0x05010400u);        // markers 4q, 4q+1
        r[2 * q + 1] = __builtin_amdgcn_perm(b23, a23, 0x07030602u);    // markers 4q+2, 4q+3
    }
    return r;
}

__global__ __launch_bounds__(256) void k_syrk_pk(const uint8_t* __restrict__ tp, int64_t nchunk, int64_t c0,
                                                 int64_t c1, const uint64_t* __restrict__ lut64,
                                                 const int2* __restrict__ tiles, int64_t nblk,
                                                 float* __restrict__ S, int64_t ldS, int accumulate) {
    __shared__ __attribute__((aligned(16))) __hip_bfloat16 sA[BT * P16];
    __shared__ __attribute__((aligned(16))) __hip_bfloat16 sB[BT * P16];
    const int2 t = tiles[xcd_remap(blockIdx.x, nblk)];
    const int64_t bi = t.x, bj = t.y;
    const bool diag = bi == bj;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int o = tid & 7, b = tid >> 3;          // marker octet, byte within the panel
    const uint64_t* gA = reinterpret_cast<const uint64_t*>(tp + (bi * nchunk * 32 + b) * 64 + 8 * o);
    const uint64_t* gB = reinterpret_cast<const uint64_t*>(tp + (bj * nchunk * 32 + b) * 64 + 8 * o);
    // one chunk = 2048 bytes = 256 uint64; this launch covers chunks [c0, c1)
    uint64_t ra = gA[c0 * 256], rb = gB[c0 * 256];
    const uint64_t* gl = lut64 + 8 * o;
    uint64_t la[8];
#pragma unroll
    for (int e = 0; e < 8; e++) la[e] = gl[c0 * 64 + e];

    f32x16 acc[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[mi][ni][r] = 0.0f;

    for (int64_t kc = c0; kc < c1; kc++) {
        uint8_t ba[8], bb[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            ba[e] = (uint8_t)(ra >> (8 * e));
            bb[e] = (uint8_t)(rb >> (8 * e));
        }
        uint64_t lc[8];
#pragma unroll
        for (int e = 0; e < 8; e++) lc[e] = la[e];
        // prefetch the next chunk's bytes and LUT entries (re-read the last one at the end)
        const int64_t kn = (kc + 1 < c1) ? kc + 1 : kc;
        ra = gA[kn * 256];
        rb = gB[kn * 256];
#pragma unroll
        for (int e = 0; e < 8; e++) la[e] = gl[kn * 64 + e];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; j++) {
            *reinterpret_cast<u32x4*>(sA + (4 * b + j) * P16 + 8 * o) = decode_row(lc, ba, j);
            if (!diag) *reinterpret_cast<u32x4*>(sB + (4 * b + j) * P16 + 8 * o) = decode_row(lc, bb, j);
        }
        __syncthreads();
        const __hip_bfloat16* pB = diag ? sA : sB;
#pragma unroll
        for (int ks = 0; ks < BK16 / 16; ks++) {
            const int kk = ks * 16 + (lane >> 5) * 8;
            bf16x8 fa[2], fb[2];
#pragma unroll
            for (int mi = 0; mi < 2; mi++)
                fa[mi] = *reinterpret_cast<const bf16x8*>(sA + (wr * 64 + mi * 32 + (lane & 31)) * P16 + kk);
#pragma unroll
            for (int ni = 0; ni < 2; ni++)
                fb[ni] = *reinterpret_cast<const bf16x8*>(pB + (wc * 64 + ni * 32 + (lane & 31)) * P16 + kk);
#pragma unroll
            for (int mi = 0; mi < 2; mi++)
#pragma unroll
                for (int ni = 0; ni < 2; ni++)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[mi], fb[ni], acc[mi][ni], 0, 0, 0);
        }
    }
    store_tile(S, ldS, bi, bj, wr, wc, lane, acc, accumulate);
}

// 256x256 tiles for the 2-bit input: 512 threads (8 waves as 4 x 2), each
// wave a 64 x 128 block of the tile (2 x 4 MFMA 32x32 blocks, 128
// accumulator registers).  Per 64-marker chunk a thread decodes the same 64
// values as in k_syrk_pk, but each decoded value feeds twice the MFMA work,
// which halves the VALU decode per MFMA (the bound of the 128 tile).  The
// per-element accumulation order (chunk by chunk, 16 markers per MFMA) is
// that of k_syrk_pk, so both give bitwise equal S.
constexpr int BT2 = 256;

// x rounded to nearest (ties to even) at 16 significant bits: the FP32
// value with the low 8 of its 23 fraction bits cleared.  Such a value is
// exactly the sum of its two BF16 pieces by truncation (the upper half of
// its bits, then the upper half of the remainder's), which the block GEMM
// over the lower triangle relies on.  Finite x only.
__device__ __forceinline__ float round16(float x) {
    const uint32_t b = __float_as_uint(x);
    return __uint_as_float((b + 0x7Fu + ((b >> 8) & 1u)) & 0xFFFFFF00u);
}

// Store a 256x256 tile from waves of (32*MI) x 128 (wave (wr, wc) owns rows
// wr*32*MI + [0, 32*MI) and columns wc*128 + [0, 128)), mirrored for
// off-diagonal tiles; accumulate: +1 add to S, -1 subtract from S, 0 store.
// R16: the stored values (after the add or subtract) rounded by round16.
template <int MI, bool R16>
__device__ __forceinline__ void store_tile256(float* S, int64_t ldS, int64_t bi, int64_t bj, int wr, int wc,
                                              int lane, const f32x16 (&acc)[MI][4], int accumulate) {
    const bool mirror = bi != bj;
#pragma unroll
    for (int mi = 0; mi < MI; mi++) {
#pragma unroll
        for (int ni = 0; ni < 4; ni++) {
            const int64_t col = bj * BT2 + wc * 128 + ni * 32 + (lane & 31);
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int64_t row0 = bi * BT2 + wr * 32 * MI + mi * 32 + 8 * g + 4 * (lane >> 5);
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; e++) v[e] = acc[mi][ni][4 * g + e];
                if (accumulate > 0) {
#pragma unroll
                    for (int e = 0; e < 4; e++) v[e] = S[(row0 + e) * ldS + col] + v[e];
                } else if (accumulate < 0) {
#pragma unroll
                    for (int e = 0; e < 4; e++) v[e] = S[(row0 + e) * ldS + col] - v[e];
                }
                if (R16) {
#pragma unroll
                    for (int e = 0; e < 4; e++) v[e] = round16(v[e]);
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

// wait for this wave's LDS traffic, then the workgroup barrier; outstanding
// global loads (the next chunk's bytes and tables) stay in flight
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Double-buffered: while the MFMAs run on chunk k from one LDS buffer, the
// same waves decode chunk k+1 into the other, so the matrix pipe and the
// VALU decode overlap (one workgroup per CU); one LDS-only barrier per chunk.
// The two buffers are separate arrays and the loop is unrolled by two, so
// the compiler sees that the decode's stores and the MFMA operand loads
// never alias and can interleave them; the loop body is one basic block (the
// last step decodes a clamped copy of the last chunk that nothing reads, and
// diagonal tiles, both panels the same, decode their B copy too).
__global__ __launch_bounds__(512) void k_syrk_pk256(const uint8_t* __restrict__ tp, int64_t nchunk, int64_t c0,
                                                    int64_t c1, const uint64_t* __restrict__ lut64,
                                                    const int2* __restrict__ tiles, int64_t nblk,
                                                    float* __restrict__ S, int64_t ldS, int accumulate) {
    __shared__ __attribute__((aligned(16))) __hip_bfloat16 sA0[BT2 * P16];
    __shared__ __attribute__((aligned(16))) __hip_bfloat16 sB0[BT2 * P16];
    __shared__ __attribute__((aligned(16))) __hip_bfloat16 sA1[BT2 * P16];
    __shared__ __attribute__((aligned(16))) __hip_bfloat16 sB1[BT2 * P16];
    const int2 t = tiles[xcd_remap(blockIdx.x, nblk)];
    const int64_t bi = t.x, bj = t.y;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    // thread -> marker octet o, byte row b of the 256-sample panel (two
    // 128-sample blocks of the tiled layout)
    const int o = tid & 7, b = tid >> 3;
    const int64_t pa = 2 * bi + (b >> 5), pb = 2 * bj + (b >> 5);
    const uint64_t* gA = reinterpret_cast<const uint64_t*>(tp + (pa * nchunk * 32 + (b & 31)) * 64 + 8 * o);
    const uint64_t* gB = reinterpret_cast<const uint64_t*>(tp + (pb * nchunk * 32 + (b & 31)) * 64 + 8 * o);
    const uint64_t* gl = lut64 + 8 * o;

    f32x16 acc[2][4];
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int ni = 0; ni < 4; ni++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[mi][ni][r] = 0.0f;

    uint64_t ra, rb, la[8];
    auto fetch = [&](int64_t kc) {
        ra = gA[kc * 256];
        rb = gB[kc * 256];
#pragma unroll
        for (int e = 0; e < 8; e++) la[e] = gl[kc * 64 + e];
    };
    // a table's low and high bytes of its BF16 entries
    auto split_bytes = [&](const uint64_t (&t)[8], uint32_t (&lo)[8], uint32_t (&hi)[8]) {
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const uint32_t l = (uint32_t)t[e], h = (uint32_t)(t[e] >> 32);
            lo[e] = __builtin_amdgcn_perm(h, l, 0x06040200u);
            hi[e] = __builtin_amdgcn_perm(h, l, 0x07050301u);
        }
    };
    // decode the chunk held in ra/rb/la into (dA, dB), fetching chunk kn
    auto decode_next = [&](__hip_bfloat16* dA, __hip_bfloat16* dB, int64_t kn) {
        const uint64_t ca = ra, cb = rb;
        uint32_t lo[8], hi[8];
        split_bytes(la, lo, hi);
        fetch(kn);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            *reinterpret_cast<u32x4*>(dA + (4 * b + j) * P16 + 8 * o) = decode_row_perm(lo, hi, ca, j);
            *reinterpret_cast<u32x4*>(dB + (4 * b + j) * P16 + 8 * o) = decode_row_perm(lo, hi, cb, j);
        }
    };
    auto multiply = [&](const __hip_bfloat16* pA, const __hip_bfloat16* pB) {
#pragma unroll
        for (int ks = 0; ks < BK16 / 16; ks++) {
            const int kk = ks * 16 + (lane >> 5) * 8;
            bf16x8 fa[2], fb[4];
#pragma unroll
            for (int mi = 0; mi < 2; mi++)
                fa[mi] = *reinterpret_cast<const bf16x8*>(pA + (wr * 64 + mi * 32 + (lane & 31)) * P16 + kk);
#pragma unroll
            for (int ni = 0; ni < 4; ni++)
                fb[ni] = *reinterpret_cast<const bf16x8*>(pB + (wc * 128 + ni * 32 + (lane & 31)) * P16 + kk);
#pragma unroll
            for (int mi = 0; mi < 2; mi++)
#pragma unroll
                for (int ni = 0; ni < 4; ni++)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[mi], fb[ni], acc[mi][ni], 0, 0, 0);
        }
    };
    // step: multiply chunk kc from (cA, cB) while decoding chunk kc+1 into (nA, nB)
    auto step = [&](const __hip_bfloat16* cA, const __hip_bfloat16* cB, __hip_bfloat16* nA, __hip_bfloat16* nB,
                    int64_t kc) {
        decode_next(nA, nB, kc + 2 < c1 ? kc + 2 : c1 - 1);
        multiply(cA, cB);
        // one MFMA, then about five decode instructions, 32 times
#pragma unroll
        for (int i = 0; i < 32; i++) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);
        }
        lds_barrier();
    };
    fetch(c0);
    decode_next(sA0, sB0, c0 + 1 < c1 ? c0 + 1 : c0);
    lds_barrier();
    int64_t kc = c0;
    for (; kc + 2 <= c1; kc += 2) {
        step(sA0, sB0, sA1, sB1, kc);
        step(sA1, sB1, sA0, sB0, kc + 1);
    }
    if (kc < c1) multiply(sA0, sB0);
    store_tile256<2, false>(S, ldS, bi, bj, wr, wc, lane, acc, accumulate);
}

// The split GRM in one pass: per 32-marker chunk each thread decodes its
// panel's bytes through both tables, hi and lo, and every wave adds
// hi.hi, hi.lo and lo.hi of its 64 x 128 block in that order (three MFMAs
// per operand pair into one accumulator).  Against three separate passes
// this reads the genotypes and the tables once, keeps a single set of
// accumulators and stores S once.  LDS: two buffers of the four decoded
// operands (A hi, A lo, B hi, B lo; 256 samples x 32 markers each, pitch 40
// for conflict-free 16-byte reads) = 160 KiB.  Threads 0-255 decode the row
// panel, 256-511 the column panel: thread (o, b) holds marker octet o of the
// chunk and byte row b (four samples).  S is stored rounded to 16 significant
// bits (round16, after the add or subtract of an accumulating pass), which
// moves each value by at most 2^-16 of itself; the block GEMM then needs two
// BF16 pieces of S instead of three.
constexpr int BKS = 32;            // markers per chunk
constexpr int PS = BKS + 8;        // LDS row pitch (bf16 elements)

__global__ __launch_bounds__(512) void k_syrk_pk256_split(const uint8_t* __restrict__ tp, int64_t nchunk,
                                                          int64_t k0, int64_t k1,
                                                          const uint64_t* __restrict__ lhi,
                                                          const uint64_t* __restrict__ llo,
                                                          const int2* __restrict__ tiles, int64_t nblk,
                                                          float* __restrict__ S, int64_t ldS, int accumulate) {
    __shared__ __attribute__((aligned(16))) __hip_bfloat16 s0[4 * BT2 * PS];   // [A hi | A lo | B hi | B lo]
    __shared__ __attribute__((aligned(16))) __hip_bfloat16 s1[4 * BT2 * PS];
    const int2 t = tiles[xcd_remap(blockIdx.x, nblk)];
    const int64_t bi = t.x, bj = t.y;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int r = tid >> 8, o = tid & 3, b = (tid & 255) >> 2;   // panel (0 rows, 1 columns), octet, byte row
    const int64_t pn = 2 * (r ? bj : bi) + (b >> 5);
    // uint64 units: 64-marker chunk c is 256 words on from chunk c-1, its
    // second 32 markers 4 words into a byte row
    const uint64_t* g = reinterpret_cast<const uint64_t*>(tp + (pn * nchunk * 32 + (b & 31)) * 64 + 8 * o);
    const uint64_t* gh = lhi + 8 * o;
    const uint64_t* gl = llo + 8 * o;
    const int dst = (2 * r) * BT2 * PS + (4 * b) * PS + 8 * o;   // this thread's hi rows; lo rows BT2*PS on

    f32x16 acc[2][4];
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int ni = 0; ni < 4; ni++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[mi][ni][q] = 0.0f;

    uint64_t raw, th[8], tl[8];
    auto fetch = [&](int64_t kc) {
        raw = g[(kc >> 1) * 256 + (kc & 1) * 4];
#pragma unroll
        for (int e = 0; e < 8; e++) th[e] = gh[kc * 32 + e];
#pragma unroll
        for (int e = 0; e < 8; e++) tl[e] = gl[kc * 32 + e];
    };
    auto split_bytes = [&](const uint64_t (&tb)[8], uint32_t (&lo)[8], uint32_t (&hi)[8]) {
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const uint32_t l = (uint32_t)tb[e], h = (uint32_t)(tb[e] >> 32);
            lo[e] = __builtin_amdgcn_perm(h, l, 0x06040200u);
            hi[e] = __builtin_amdgcn_perm(h, l, 0x07050301u);
        }
    };
    // decode the chunk held in raw/th/tl into buffer d, fetching chunk kn
    auto decode_next = [&](__hip_bfloat16* d, int64_t kn) {
        const uint64_t c = raw;
        uint32_t hlo[8], hhi[8], llo8[8], lhi8[8];
        split_bytes(th, hlo, hhi);
        split_bytes(tl, llo8, lhi8);
        fetch(kn);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            *reinterpret_cast<u32x4*>(d + dst + j * PS) = decode_row_perm(hlo, hhi, c, j);
            *reinterpret_cast<u32x4*>(d + dst + BT2 * PS + j * PS) = decode_row_perm(llo8, lhi8, c, j);
        }
    };
    auto multiply = [&](const __hip_bfloat16* d) {
        const __hip_bfloat16* Ah = d;
        const __hip_bfloat16* Al = d + BT2 * PS;
        const __hip_bfloat16* Bh = d + 2 * BT2 * PS;
        const __hip_bfloat16* Bl = d + 3 * BT2 * PS;
#pragma unroll
        for (int ks = 0; ks < BKS / 16; ks++) {
            const int kk = ks * 16 + (lane >> 5) * 8;
            bf16x8 fah[2], fal[2];
#pragma unroll
            for (int mi = 0; mi < 2; mi++) {
                const int row = (wr * 64 + mi * 32 + (lane & 31)) * PS + kk;
                fah[mi] = *reinterpret_cast<const bf16x8*>(Ah + row);
                fal[mi] = *reinterpret_cast<const bf16x8*>(Al + row);
            }
#pragma unroll
            for (int ni = 0; ni < 4; ni++) {
                const int col = (wc * 128 + ni * 32 + (lane & 31)) * PS + kk;
                const bf16x8 fbh = *reinterpret_cast<const bf16x8*>(Bh + col);
                const bf16x8 fbl = *reinterpret_cast<const bf16x8*>(Bl + col);
#pragma unroll
                for (int mi = 0; mi < 2; mi++) {
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fah[mi], fbh, acc[mi][ni], 0, 0, 0);
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fah[mi], fbl, acc[mi][ni], 0, 0, 0);
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fal[mi], fbh, acc[mi][ni], 0, 0, 0);
                }
            }
        }
    };
    // multiply chunk kc from cur while decoding chunk kc+1 into nxt
    auto step = [&](const __hip_bfloat16* cur, __hip_bfloat16* nxt, int64_t kc) {
        decode_next(nxt, kc + 2 < k1 ? kc + 2 : k1 - 1);
        multiply(cur);
#pragma unroll
        for (int i = 0; i < 48; i++) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);
        }
        lds_barrier();
    };
    fetch(k0);
    decode_next(s0, k0 + 1 < k1 ? k0 + 1 : k0);
    lds_barrier();
    int64_t kc = k0;
    for (; kc + 2 <= k1; kc += 2) {
        step(s0, s1, kc);
        step(s1, s0, kc + 1);
    }
    if (kc < k1) multiply(s0);
    store_tile256<2, true>(S, ldS, bi, bj, wr, wc, lane, acc, accumulate);
}

// Re-tile the packed genotypes for k_syrk_pk: block (p, c) byte (b, m) =
// byte 32p + b of QC'd marker 64c + m, with padding codes 0b01.
__global__ __launch_bounds__(256) void k_pack_tiles(const uint8_t* __restrict__ packed, int64_t Mqc, int64_t Nnm,
                                                    int64_t bpm_nm, int64_t nchunk, uint8_t* __restrict__ tp) {
    const int64_t p = blockIdx.x, c = blockIdx.y;
    uint8_t* blk = tp + (p * nchunk + c) * 2048;
    for (int i = threadIdx.x; i < 2048; i += 256) {
        const int b = i >> 6, m = i & 63;
        const int64_t j = c * 64 + m;
        const int64_t byte_idx = p * 32 + b;
        const int64_t s0 = byte_idx * 4;
        unsigned v = 0x55u;
        if (j < Mqc && byte_idx < bpm_nm) {
            v = packed[j * bpm_nm + byte_idx];
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (s0 + q >= Nnm) v = (v & ~(3u << (2 * q))) | (1u << (2 * q));
        }
        blk[i] = (uint8_t)v;
    }
}

// lut64[j] for j < nchunk*64 (zero for padding markers).  piece 0: each
// value rounded to BF16 (hi); piece 1: the remainder x - hi, exact in FP32,
// rounded to BF16 (lo), so that x - hi - lo is within 2^-16 |x| (BF16 keeps
// 8 significant bits: each rounding is within 2^-8 relative).
__device__ __forceinline__ uint64_t bf16_piece(float x, int piece) {
    const __hip_bfloat16 hi = __float2bfloat16(x);
    if (piece == 0) return __bfloat16_as_ushort(hi);
    return __bfloat16_as_ushort(__float2bfloat16(__fsub_rn(x, __bfloat162float(hi))));
}

__global__ void k_lut64(const float4* __restrict__ lut, int64_t j0, int64_t j1, int64_t n, int piece,
                        uint64_t* __restrict__ out) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint64_t v = 0;
    if (j >= j0 && j < j1) {
        const float4 L = lut[j];
        const uint64_t l0 = bf16_piece(L.x, piece);
        const uint64_t l1 = bf16_piece(L.y, piece);
        const uint64_t l2 = bf16_piece(L.z, piece);
        v = l2 | (0ull << 16) | (l1 << 32) | (l0 << 48);   // codes 00, 01, 10, 11
    }
    out[j] = v;
}

// Copy each diagonal tile's lower half onto its upper half: S[c][r] =
// S[r][c] for r > c inside tile t (rows and columns t*T + [0, T)).  The split
// GRM's cross terms add hi.lo and lo.hi to the two mirror entries of a
// diagonal tile in opposite orders; this keeps S exactly symmetric, with the
// lower triangle's values, as off-diagonal tiles are mirrored.  Block (t, r):
// row r of tile t; its threads write column r of the upper part.
__global__ __launch_bounds__(256) void k_mirror_diag(float* __restrict__ S, int64_t ld, int T) {
    const int64_t t = blockIdx.x, r = blockIdx.y;
    const int64_t base = t * T;
    for (int64_t c = threadIdx.x; c < r; c += blockDim.x)
        S[(base + c) * ld + base + r] = S[(base + r) * ld + base + c];
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

hipError_t launch_pack_tiles(hipStream_t s, const uint8_t* packed, const float4* lut, int64_t Mqc, int64_t Nnm,
                             int64_t Npad, int64_t bpm_nm, uint8_t* tp, uint64_t* lut64) {
    const int64_t nchunk = (Mqc + 63) / 64;
    hipLaunchKernelGGL(k_pack_tiles, dim3((unsigned)(Npad / BT), (unsigned)nchunk), dim3(256), 0, s, packed, Mqc,
                       Nnm, bpm_nm, nchunk, tp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_lut64, dim3((unsigned)((nchunk * 64 + 255) / 256)), dim3(256), 0, s, lut, (int64_t)0, Mqc,
                       nchunk * 64, 0, lut64);
    return hipGetLastError();
}

hipError_t launch_lut64_range(hipStream_t s, const float4* lut, int64_t Mqc, int64_t j0, int64_t j1, int piece,
                              uint64_t* lut64) {
    const int64_t nchunk = (Mqc + 63) / 64;
    hipLaunchKernelGGL(k_lut64, dim3((unsigned)((nchunk * 64 + 255) / 256)), dim3(256), 0, s, lut, j0, j1,
                       nchunk * 64, piece, lut64);
    return hipGetLastError();
}

hipError_t launch_syrk_pk(hipStream_t s, const uint8_t* tp, const uint64_t* lut64, int64_t Mqc, int64_t Npad,
                          const int2* tiles, float* S, int accumulate, int64_t j0, int64_t j1, int tile) {
    if ((tile != BT && tile != BT2) || Npad % tile != 0) return hipErrorInvalidValue;
    const int64_t nt = Npad / tile;
    const int64_t nblk = nt * (nt + 1) / 2;
    const int64_t nchunk = (Mqc + 63) / 64;
    if (j1 < 0) j1 = Mqc;
    const int64_t c0 = j0 / 64, c1 = (j1 + 63) / 64;
    if (c0 >= c1) return hipSuccess;
    if (tile == BT2)
        hipLaunchKernelGGL(k_syrk_pk256, dim3((unsigned)nblk), dim3(512), 0, s, tp, nchunk, c0, c1, lut64, tiles,
                           nblk, S, Npad, accumulate);
    else
        hipLaunchKernelGGL(k_syrk_pk, dim3((unsigned)nblk), dim3(256), 0, s, tp, nchunk, c0, c1, lut64, tiles,
                           nblk, S, Npad, accumulate);
    return hipGetLastError();
}

hipError_t launch_syrk_pk_split(hipStream_t s, const uint8_t* tp, const uint64_t* lhi, const uint64_t* llo,
                                int64_t Mqc, int64_t Npad, const int2* tiles, float* S, int accumulate, int64_t j0,
                                int64_t j1) {
    if (Npad % BT2 != 0 || !lhi || !llo) return hipErrorInvalidValue;
    const int64_t nt = Npad / BT2;
    const int64_t nblk = nt * (nt + 1) / 2;
    const int64_t nchunk = (Mqc + 63) / 64;
    if (j1 < 0) j1 = Mqc;
    const int64_t k0 = j0 / BKS, k1 = (j1 + BKS - 1) / BKS;
    if (k0 >= k1) return hipSuccess;
    hipLaunchKernelGGL(k_syrk_pk256_split, dim3((unsigned)nblk), dim3(512), 0, s, tp, nchunk, k0, k1, lhi, llo,
                       tiles, nblk, S, Npad, accumulate);
    return hipGetLastError();
}

hipError_t launch_mirror_diag(hipStream_t s, float* S, int64_t Npad, int tile) {
    if (tile <= 0 || Npad % tile != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mirror_diag, dim3((unsigned)(Npad / tile), (unsigned)tile), dim3(256), 0, s, S, Npad, tile);
    return hipGetLastError();
}

// Tile edge of the 2-bit-input SYRK: 256 unless SAIGE_SYRK_TILE=128 (the
// tests compare the two, which are bitwise equal).
int syrk_pk_tile() {
    if (const char* env = std::getenv("SAIGE_SYRK_TILE"))
        if (std::atoi(env) == BT) return BT;
    return BT2;
}

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
