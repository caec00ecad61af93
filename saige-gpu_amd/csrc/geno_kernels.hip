// Genotype kernels: PLINK-1 2-bit decode with allele / missing counts and
// QC, imputation + repacking, the standardisation lookup table, the FP32
// diagonal sum_j z_ij^2, standardised marker panels for the SYRK, and a
// counter-based synthetic genotype generator.
//
// Integer work is bit-exact with the reference by construction.  The FP32
// scalar formulas use explicitly rounded intrinsics (__fadd_rn ...) so no
// contraction into FMA changes a rounding step.
#include "saige_internal.hpp"

#pragma clang fp contract(off)

namespace saige {

namespace {

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// PLINK-1 code -> A1 allele count, 3 = missing (FG.cpp:343-357: b = bit0,
// a = bit1; (0,0)->2, (1,0)->missing, (0,1)->1, (1,1)->0).
__device__ __forceinline__ int code_geno(unsigned c) {
    return c == 0u ? 2 : (c == 1u ? 3 : (c == 2u ? 1 : 0));
}

// internal code after repacking (FG.cpp:41-44, 565-571)
__device__ __forceinline__ unsigned internal_code(int g) {
    return g == 0 ? 3u : (g == 1 ? 2u : 0u);
}

// Counts of one 32-bit word of PLINK codes (16 samples) under a code mask
// (bit 2s set: sample s of the word is counted): with l, h the low and high
// bit of a code, missing = l & ~h (01), and a counted non-missing code adds
// 2 - l - h A1 alleles, where l = 1 only together with h = 1 (code 11).
// miss += missing codes, nonm += counted non-missing codes, t += sum of
// l + h over them; alleleCount = 2 nonm - t.
__device__ __forceinline__ void count_word(uint32_t w, uint32_t msk, uint32_t& miss, uint32_t& nonm,
                                           uint32_t& t) {
    const uint32_t L = w & msk, H = (w >> 1) & msk;
    const uint32_t Mi = L & ~H;
    const uint32_t nmiss = (uint32_t)__popc(Mi);
    miss += nmiss;
    nonm += (uint32_t)__popc(msk) - nmiss;
    t += (uint32_t)__popc(L & H) + (uint32_t)__popc(H);
}

// The reference's FP32 per-marker formulas in order (FG.cpp:424-547,
// 914-921) from the integer counts over the phenotyped samples.
__device__ __forceinline__ MarkerStat marker_stat(int alleleCount, int numMissing, const DecodeParams& p,
                                                  int inr) {
    const uint64_t nnm = (uint64_t)p.Nnm;
    // FG.cpp:424 and 442
    float altFreq = div_rn((float)alleleCount, (float)((nnm - (uint64_t)numMissing) * 2u));
    float missingRate = div_rn((float)numMissing, (float)nnm);
    // FG.cpp:447; NaN (all calls missing) maps to 0
    float twice = __fmul_rn(2.0f, altFreq);
    int fill = (twice != twice) ? 0 : (int)roundf(twice);
    if (numMissing > 0) alleleCount = alleleCount + fill * numMissing;   // FG.cpp:454
    altFreq = div_rn((float)alleleCount, (float)(nnm * 2u));          // FG.cpp:484
    float one_m = __fsub_rn(1.0f, altFreq);
    float maf = (one_m < altFreq) ? one_m : altFreq;                     // std::min
    int twoN = (int)p.Nnm * 2;
    int mac = (alleleCount < twoN - alleleCount) ? alleleCount : twoN - alleleCount;
    int passQC = (maf >= p.min_maf && missingRate <= p.max_missing) ? 1 : 0;
    int passVR = 0;
    if (p.vr_mode == 1) {
        if ((float)mac >= p.min_mac_vr && (float)mac < p.max_mac_vr) passVR = 1;
        else if ((float)mac >= p.max_mac_vr) passVR = inr;
    } else if (p.vr_mode == 2) {
        if ((float)mac >= p.min_mac_vr) passVR = inr;
    }
    if (passVR) passQC = 0;
    // FG.cpp:914-921
    float Std = sqrt_rn(__fmul_rn(__fmul_rn(2.0f, altFreq), one_m));
    float invStd = (Std == 0.0f) ? 0.0f : div_rn(1.0f, Std);
    MarkerStat st;
    st.allele_count = alleleCount;
    st.num_missing = numMissing;
    st.mac = mac;
    st.fill = fill;
    st.freq = altFreq;
    st.missing_rate = missingRate;
    st.inv_std = invStd;
    st.pass_qc = (uint8_t)passQC;
    st.pass_vr = (uint8_t)passVR;
    st.pad[0] = st.pad[1] = 0;
    return st;
}

// One wave per marker, four markers per workgroup, no workgroup barrier.
// A row of bpm bytes is read as the 16-byte chunks of the aligned span that
// covers it (a chunk always holds at least one byte of the row, so it never
// leaves the row's page), two chunks per lane in flight.  Bytes outside the
// row and the pad samples of its last byte are masked out: without a sample
// indicator (every .fam sample phenotyped) from the byte positions; with
// one (MASKED) from imask, 16 copies of the row's code mask (2 bits per
// sample, 0b11 when indicatorGenoSamplesWithPheno is set), copy r shifted by
// r bytes with stride ld, so that copy (row address mod 16) lines up with
// the aligned chunks.  Lane 0 evaluates the per-marker formulas.
template <bool MASKED>
__global__ __launch_bounds__(256) void k_decode_stats(const uint8_t* __restrict__ bed,
                                                      const uint8_t* __restrict__ imask, int64_t ld,
                                                      const uint8_t* __restrict__ randset, DecodeParams p,
                                                      MarkerStat* __restrict__ stats) {
    const int lane = threadIdx.x & 63;
    const int64_t m = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (m >= p.M) return;
    const uint8_t* row = bed + m * p.bpm;
    const int64_t r = (int64_t)((uintptr_t)row & 15u);
    const uint8_t* A = row - r;
    const int64_t span = r + p.bpm;                   // bytes from A to the row's end
    const int64_t nchunk = (span + 15) / 16;
    const uint8_t* mk = MASKED ? imask + r * ld : nullptr;
    // codes of the last byte that are samples (the rest is padding)
    const int last_codes = (int)(p.N - 4 * (p.bpm - 1));
    const uint32_t last_mask = last_codes >= 4 ? 0xFFu : ((1u << (2 * last_codes)) - 1u);
    uint32_t miss = 0, nonm = 0, t = 0;
    auto chunk = [&](int64_t q, const uint4& v) {
        const int64_t x0 = 16 * q;
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
        uint32_t mw[4];
        if (MASKED) {
            const uint4 mv = *reinterpret_cast<const uint4*>(mk + x0);
            mw[0] = mv.x; mw[1] = mv.y; mw[2] = mv.z; mw[3] = mv.w;
        } else if (x0 - r >= 0 && x0 - r + 16 <= p.bpm - 1) {
            mw[0] = mw[1] = mw[2] = mw[3] = 0xFFFFFFFFu;    // interior: every byte a full row byte
        } else {
#pragma unroll
            for (int d = 0; d < 4; d++) {
                uint32_t mm = 0;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const int64_t pos = x0 + 4 * d + b - r;
                    const uint32_t bm = (pos < 0 || pos >= p.bpm) ? 0u : (pos == p.bpm - 1 ? last_mask : 0xFFu);
                    mm |= bm << (8 * b);
                }
                mw[d] = mm;
            }
        }
#pragma unroll
        for (int d = 0; d < 4; d++) count_word(w[d], mw[d] & 0x55555555u, miss, nonm, t);
    };
    int64_t q = lane;
    for (; q + 64 < nchunk; q += 128) {
        const uint4 v0 = *reinterpret_cast<const uint4*>(A + 16 * q);
        const uint4 v1 = *reinterpret_cast<const uint4*>(A + 16 * (q + 64));
        chunk(q, v0);
        chunk(q + 64, v1);
    }
    if (q < nchunk) chunk(q, *reinterpret_cast<const uint4*>(A + 16 * q));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        miss += __shfl_xor(miss, o, 64);
        nonm += __shfl_xor(nonm, o, 64);
        t += __shfl_xor(t, o, 64);
    }
    if (lane == 0) {
        const int alleleCount = (int)(2u * nonm - t);
        stats[m] = marker_stat(alleleCount, (int)miss, p, randset ? (int)randset[m] : 0);
    }
}

// Repack listed markers into subSampleInGeno order with imputation
// (FG.cpp:551-576).  One workgroup per marker; each thread builds whole
// output bytes.  subidx == nullptr means the identity order with Nnm == N:
// then every output byte is a function of its input byte and the marker's
// fill genotype alone, taken from a 256-entry table per fill (0, 1, 2) built
// once per workgroup in LDS, four consecutive bytes per thread; the last
// byte keeps only the bits of samples below Nnm.
__global__ __launch_bounds__(256) void k_repack(const uint8_t* __restrict__ bed, int64_t bpm,
                                                const int32_t* __restrict__ markers, int64_t nmark,
                                                const MarkerStat* __restrict__ stats,
                                                const int32_t* __restrict__ subidx, int64_t Nnm,
                                                int64_t bpm_nm, uint8_t* __restrict__ out) {
    if (subidx == nullptr) {
        __shared__ uint8_t tab[3][256];
        for (int e = threadIdx.x; e < 3 * 256; e += blockDim.x) {
            const int fill = e / 256, in = e % 256;
            unsigned byte = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int g = code_geno((unsigned)(in >> (2 * j)) & 3u);
                if (g == 3) g = fill;
                byte |= internal_code(g) << (2 * j);
            }
            tab[fill][in] = (uint8_t)byte;
        }
        __syncthreads();
        const int tail = (int)(Nnm % 4);                      // samples in a partial last byte
        const unsigned tail_mask = tail ? (1u << (2 * tail)) - 1u : 0xFFu;
        for (int64_t mi = blockIdx.x; mi < nmark; mi += gridDim.x) {
            const int64_t m = markers[mi];
            const uint8_t* row = bed + m * bpm;
            const uint8_t* t = tab[min(max(stats[m].fill, 0), 2)];   // fill = round(2p) is 0, 1 or 2
            uint8_t* orow = out + mi * bpm_nm;
            for (int64_t o = 4 * (int64_t)threadIdx.x; o < bpm_nm; o += 4 * (int64_t)blockDim.x) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int64_t oj = o + j;
                    if (oj < bpm_nm) {
                        unsigned v = t[row[oj]];
                        if (oj == bpm_nm - 1) v &= tail_mask;
                        orow[oj] = (uint8_t)v;
                    }
                }
            }
        }
        return;
    }
    for (int64_t mi = blockIdx.x; mi < nmark; mi += gridDim.x) {
        const int64_t m = markers[mi];
        const uint8_t* row = bed + m * bpm;
        const int fill = stats[m].fill;
        uint8_t* orow = out + mi * bpm_nm;
        for (int64_t o = threadIdx.x; o < bpm_nm; o += blockDim.x) {
            unsigned byte = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int64_t indx = o * 4 + j;
                if (indx >= Nnm) break;
                const int64_t src = (int64_t)subidx[indx] - 1;
                int g = code_geno((row[src >> 2] >> ((src & 3) * 2)) & 3u);
                if (g == 3) g = fill;
                byte |= internal_code(g) << (2 * j);
            }
            orow[o] = (uint8_t)byte;
        }
    }
}

// setStdGenoLookUpArr (FG.cpp:152-159)
__global__ void k_lut(const int32_t* __restrict__ markers, int64_t nmark,
                      const MarkerStat* __restrict__ stats, float4* __restrict__ lut) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nmark) return;
    const MarkerStat st = stats[markers[i]];
    float m2 = __fmul_rn(2.0f, st.freq);
    float4 L;
    L.x = __fmul_rn(__fsub_rn(0.0f, m2), st.inv_std);
    L.y = __fmul_rn(__fsub_rn(1.0f, m2), st.inv_std);
    L.z = __fmul_rn(__fsub_rn(2.0f, m2), st.inv_std);
    L.w = 0.0f;
    lut[i] = L;
}

// internal code -> 2-(a+b) (FG.cpp:596-600)
__device__ __forceinline__ int internal_geno(unsigned c) { return 2 - (int)((c & 1u) + (c >> 1)); }

__device__ __forceinline__ float lut_pick(const float4 L, int g) {
    return g == 0 ? L.x : (g == 1 ? L.y : L.z);
}

// Get_Diagof_StdGeno (FG.cpp:665-704): one thread per sample, markers summed
// strictly in order in FP32 (z*z rounded, then added), with DU marker loads
// in flight per iteration (32 ran no faster: a wave's load covers 16 bytes of
// a marker row, and that request rate, not the latency, bounds it).
constexpr int DU = 8;
__global__ __launch_bounds__(256) void k_diag_std(const uint8_t* __restrict__ packed,
                                                  const float4* __restrict__ lut, int64_t Mqc,
                                                  int64_t Nnm, int64_t bpm_nm,
                                                  float* __restrict__ diag) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= Nnm) return;
    const int64_t byte_i = i >> 2;
    const int sh = (int)(i & 3) * 2;
    float d = 0.0f;
    int64_t j = 0;
    for (; j + DU <= Mqc; j += DU) {
        unsigned b[DU];
#pragma unroll
        for (int u = 0; u < DU; u++) b[u] = packed[(j + u) * bpm_nm + byte_i];
#pragma unroll
        for (int u = 0; u < DU; u++) {
            float z = lut_pick(lut[j + u], internal_geno((b[u] >> sh) & 3u));
            d = __fadd_rn(d, __fmul_rn(z, z));
        }
    }
    for (; j < Mqc; j++) {
        unsigned bb = packed[j * bpm_nm + byte_i];
        float z = lut_pick(lut[j], internal_geno((bb >> sh) & 3u));
        d = __fadd_rn(d, __fmul_rn(z, z));
    }
    diag[i] = d;
}

__global__ void k_get_geno(const uint8_t* __restrict__ row, int64_t Nnm, int32_t* __restrict__ out) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= Nnm) return;
    out[i] = internal_geno((row[i >> 2] >> ((i & 3) * 2)) & 3u);
}

__global__ void k_std_geno(const uint8_t* __restrict__ row, const float4* __restrict__ lut,
                           int64_t Nnm, float* __restrict__ out) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= Nnm) return;
    out[i] = lut_pick(lut[0], internal_geno((row[i >> 2] >> ((i & 3) * 2)) & 3u));
}

// Standardised panel: Zp[r][c] = z of panel marker c (global j0 + c) at
// sample r, for a 128-sample x 64-marker tile per workgroup, transposed
// through LDS so both the packed reads and the row writes are contiguous.
template <typename T>
__device__ __forceinline__ T to_elem(float v);
template <>
__device__ __forceinline__ float to_elem<float>(float v) { return v; }
template <>
__device__ __forceinline__ __hip_bfloat16 to_elem<__hip_bfloat16>(float v) { return __float2bfloat16(v); }

template <typename T>
__global__ __launch_bounds__(256) void k_std_panel(const uint8_t* __restrict__ packed,
                                                   const float4* __restrict__ lut, int64_t j0,
                                                   int64_t np, int64_t Nnm, int64_t bpm_nm,
                                                   int64_t ldz, T* __restrict__ Zp) {
    __shared__ float tile[128][65];
    const int tid = threadIdx.x;
    const int64_t s0 = (int64_t)blockIdx.x * 128;   // first sample of the tile
    const int64_t c0 = (int64_t)blockIdx.y * 64;    // first panel column of the tile
    {
        const int mk = tid >> 2;        // marker within tile 0..63
        const int part = tid & 3;       // 32 samples each
        const int64_t c = c0 + mk;
        const bool valid_m = c < np;
        float4 L = valid_m ? lut[j0 + c] : make_float4(0.f, 0.f, 0.f, 0.f);
        const uint8_t* row = packed + (j0 + c) * bpm_nm;
#pragma unroll
        for (int bq = 0; bq < 8; bq++) {
            const int64_t sb = s0 + part * 32 + bq * 4;   // first sample of this byte
            unsigned byte = 0;
            if (valid_m && sb < Nnm) byte = row[sb >> 2];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int64_t s = sb + j;
                float z = 0.0f;
                if (valid_m && s < Nnm) z = lut_pick(L, internal_geno((byte >> (2 * j)) & 3u));
                tile[part * 32 + bq * 4 + j][mk] = z;
            }
        }
    }
    __syncthreads();
    const int lane = tid & 63, wv = tid >> 6;
    const int64_t c = c0 + lane;
    if (c < ldz) {
        for (int r = wv; r < 128; r += 4) {
            Zp[(s0 + r) * ldz + c] = to_elem<T>(tile[r][lane]);
        }
    }
}

// Counter-based uniform in [0, 1) from a 64-bit mix (splitmix64 finaliser).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float unif(uint64_t seed, uint64_t a, uint64_t b, uint64_t k) {
    uint64_t h = mix64(seed ^ mix64(a * 0x9E3779B97F4A7C15ull + mix64(b * 0xC2B2AE3D27D4EB4Full + k)));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ int synth_geno(uint64_t seed, int64_t m, int64_t s, float pj) {
    return (unif(seed, m, s, 1) < pj ? 1 : 0) + (unif(seed, m, s, 2) < pj ? 1 : 0);
}

__device__ __forceinline__ void synth_byte(uint8_t* __restrict__ bed, int64_t t, int64_t N, int64_t bpm,
                                           float maf_lo, float maf_hi, float miss, int64_t npairs,
                                           uint64_t seed) {
    const int64_t m = t / bpm, b = t - m * bpm;
    const float pj = maf_lo + (maf_hi - maf_lo) * unif(seed, m, 0x7FFFFFFF, 0);
    unsigned byte = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int64_t s = b * 4 + j;
        if (s >= N) break;   // padding bits stay 00, as PLINK writes them
        int g = synth_geno(seed, m, s, pj);
        if (s < 2 * npairs && (s & 1) && unif(seed, m, s, 4) < 0.5f) g = synth_geno(seed, m, s - 1, pj);
        unsigned code;
        if (unif(seed, m, s, 3) < miss) code = 1u;
        else code = (g == 2) ? 0u : (g == 1 ? 2u : 3u);
        byte |= code << (2 * j);
    }
    bed[t] = (uint8_t)byte;
}

// Grid-stride over the M * bpm bytes: at N = 200k, M = 500k that is 2.5e10
// bytes, more work-items than one launch can address (2^32).
__global__ void k_synth_bed(uint8_t* __restrict__ bed, int64_t N, int64_t M, int64_t bpm,
                            float maf_lo, float maf_hi, float miss, int64_t npairs, uint64_t seed) {
    const int64_t total = M * bpm;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += stride)
        synth_byte(bed, t, N, bpm, maf_lo, maf_hi, miss, npairs, seed);
}


// out = a - b, FP32 (diag(K) minus one chromosome's part, FG.cpp:4951)
__global__ void k_sub(const float* __restrict__ a, const float* __restrict__ b, int64_t n, float* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = __fsub_rn(a[i], b[i]);
}

inline unsigned grid_for(int64_t n, int64_t cap) { return (unsigned)(n < cap ? (n > 0 ? n : 1) : cap); }

}  // namespace

hipError_t launch_decode_stats(hipStream_t s, const uint8_t* bed, const uint8_t* imask, int64_t ld,
                               const uint8_t* randset, DecodeParams p, MarkerStat* stats) {
    const dim3 grid((unsigned)((p.M + 3) / 4));
    if (imask)
        hipLaunchKernelGGL(k_decode_stats<true>, grid, dim3(256), 0, s, bed, imask, ld, randset, p, stats);
    else
        hipLaunchKernelGGL(k_decode_stats<false>, grid, dim3(256), 0, s, bed, imask, ld, randset, p, stats);
    return hipGetLastError();
}

hipError_t launch_repack(hipStream_t s, const uint8_t* bed, int64_t bpm, const int32_t* markers,
                         int64_t nmark, const MarkerStat* stats, const int32_t* subidx, int64_t Nnm,
                         int64_t bpm_nm, uint8_t* out) {
    if (nmark == 0) return hipSuccess;
    hipLaunchKernelGGL(k_repack, dim3(grid_for(nmark, 16384)), dim3(256), 0, s, bed, bpm, markers,
                       nmark, stats, subidx, Nnm, bpm_nm, out);
    return hipGetLastError();
}

hipError_t launch_lut(hipStream_t s, const int32_t* markers, int64_t nmark, const MarkerStat* stats,
                      float4* lut) {
    if (nmark == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lut, dim3((unsigned)((nmark + 255) / 256)), dim3(256), 0, s, markers, nmark,
                       stats, lut);
    return hipGetLastError();
}

hipError_t launch_diag_std(hipStream_t s, const uint8_t* packed, const float4* lut, int64_t Mqc,
                           int64_t Nnm, int64_t bpm_nm, float* diag) {
    hipLaunchKernelGGL(k_diag_std, dim3((unsigned)((Nnm + 255) / 256)), dim3(256), 0, s, packed, lut,
                       Mqc, Nnm, bpm_nm, diag);
    return hipGetLastError();
}

hipError_t launch_sub(hipStream_t s, const float* a, const float* b, int64_t n, float* out) {
    hipLaunchKernelGGL(k_sub, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, n, out);
    return hipGetLastError();
}

hipError_t launch_get_geno(hipStream_t s, const uint8_t* row, int64_t Nnm, int32_t* out) {
    hipLaunchKernelGGL(k_get_geno, dim3((unsigned)((Nnm + 255) / 256)), dim3(256), 0, s, row, Nnm, out);
    return hipGetLastError();
}

hipError_t launch_std_geno(hipStream_t s, const uint8_t* row, const float4* lut_row, int64_t Nnm,
                           float* out) {
    hipLaunchKernelGGL(k_std_geno, dim3((unsigned)((Nnm + 255) / 256)), dim3(256), 0, s, row, lut_row,
                       Nnm, out);
    return hipGetLastError();
}

hipError_t launch_synth_bed(hipStream_t s, uint8_t* bed, int64_t N, int64_t M, float maf_lo,
                            float maf_hi, float miss, int64_t npairs, uint64_t seed) {
    const int64_t bpm = (N + 3) / 4;
    const int64_t n = M * bpm;
    hipLaunchKernelGGL(k_synth_bed, dim3(grid_for((n + 255) / 256, 1 << 16)), dim3(256), 0, s, bed, N, M, bpm,
                       maf_lo, maf_hi, miss, npairs, seed);
    return hipGetLastError();
}

hipError_t launch_std_panel_f32(hipStream_t s, const uint8_t* packed, const float4* lut, int64_t j0,
                                int64_t np, int64_t Nnm, int64_t Npad, int64_t bpm_nm, int64_t ldz,
                                float* Zp) {
    dim3 grid((unsigned)(Npad / 128), (unsigned)((ldz + 63) / 64));
    hipLaunchKernelGGL(k_std_panel<float>, grid, dim3(256), 0, s, packed, lut, j0, np, Nnm, bpm_nm, ldz,
                       Zp);
    return hipGetLastError();
}

hipError_t launch_std_panel_bf16(hipStream_t s, const uint8_t* packed, const float4* lut, int64_t j0,
                                 int64_t np, int64_t Nnm, int64_t Npad, int64_t bpm_nm, int64_t ldz,
                                 __hip_bfloat16* Zp) {
    dim3 grid((unsigned)(Npad / 128), (unsigned)((ldz + 63) / 64));
    hipLaunchKernelGGL(k_std_panel<__hip_bfloat16>, grid, dim3(256), 0, s, packed, lut, j0, np, Nnm,
                       bpm_nm, ldz, Zp);
    return hipGetLastError();
}

}  // namespace saige
