// k_mfma.hip -- FP16-MFMA fused fixed-grid beamformer for gfx950.
//
// Per frequency f the contraction y[t][b] = sum_a W[b][f][a] * X[t][f][a]
// (reference src/cpu_beamformer.cpp:56-78, no conjugate) is a real GEMM
//   [xr0 xi0 xr1 xi1 ...] (1 x 2A)  *  B (2A x 2B),
// where output column 2b holds Re(y_b) (weights (wr, -wi) per antenna) and
// column 2b+1 holds Im(y_b) (weights (wi, wr)).  int4 samples are exact in
// FP16, so the only rounding in the products comes from the weights: the
// SPLIT variant stores each weight as hi = fp16(w) plus lo = fp16(w - hi)
// and runs the K loop twice over the same A operand (products are exact in
// the FP32 accumulator), recovering ~22 significant bits per weight.
//
// Instruction: v_mfma_f32_16x16x32_f16.  A wave owns one frequency and
// NT N-tiles (8 beams each, re/im interleaved); its weights live in VGPRs
// for the whole time range it processes.  Each lane loads A/4 contiguous
// packed bytes of one time row per 16-row tile and converts them to FP16
// pairs with a byte permute, an xor and one packed subtraction.
//
// Output layout of a 16x16 MFMA: lane l holds column l%16 for rows
// 4*(l/16)+r, r = 0..3.  Tile rows are mapped to time so that row group
// G = l/16 of the super-tile covers GROUP_ROWS consecutive spectra: each
// lane then accumulates whole integration windows in registers.
//   MODE 0 (no integration): super-tile = 1 tile, row 4G+r -> t = base+4G+r
//   MODE 1 (S == 10): 5 tiles, G covers 20 spectra = 2 windows
//   MODE 2 (S % 4 == 0): S/4 tiles, G covers one window of S spectra
// Re and im of a beam sit in adjacent lanes; one quad-permute DPP adds the
// two squared sums.
#include "bf_kernels.hpp"

namespace bfhip {
namespace {

constexpr int kWaveSize = 64;
constexpr int kXcds = 8;

__device__ __forceinline__ float dpp_swap_pairs(float v) {
    // quad_perm [1,0,3,2]: lane i <- lane i^1.
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true));
}

template <int A>
struct RowLoad;
template <>
struct RowLoad<32> {  // 8 bytes per lane
    static __device__ __forceinline__ void load(const uint8_t* p, bool ok, unsigned d[2]) {
        uint2 v = ok ? *reinterpret_cast<const uint2*>(p) : make_uint2(0u, 0u);
        d[0] = v.x;
        d[1] = v.y;
    }
};
template <>
struct RowLoad<64> {  // 16 bytes per lane
    static __device__ __forceinline__ void load(const uint8_t* p, bool ok, unsigned d[4]) {
        uint4 v = ok ? *reinterpret_cast<const uint4*>(p) : make_uint4(0u, 0u, 0u, 0u);
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    }
};
template <>
struct RowLoad<128> {  // 32 bytes per lane
    static __device__ __forceinline__ void load(const uint8_t* p, bool ok, unsigned d[8]) {
        const uint4* q = reinterpret_cast<const uint4*>(p);
        uint4 v0 = ok ? q[0] : make_uint4(0u, 0u, 0u, 0u);
        uint4 v1 = ok ? q[1] : make_uint4(0u, 0u, 0u, 0u);
        d[0] = v0.x; d[1] = v0.y; d[2] = v0.z; d[3] = v0.w;
        d[4] = v1.x; d[5] = v1.y; d[6] = v1.z; d[7] = v1.w;
    }
};
template <>
struct RowLoad<256> {  // 64 bytes per lane
    static __device__ __forceinline__ void load(const uint8_t* p, bool ok, unsigned d[16]) {
        const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint4 v = ok ? q[i] : make_uint4(0u, 0u, 0u, 0u);
            d[4 * i + 0] = v.x;
            d[4 * i + 1] = v.y;
            d[4 * i + 2] = v.z;
            d[4 * i + 3] = v.w;
        }
    }
};

template <int MODE>
struct TileGeom;
template <>
struct TileGeom<0> {
    static constexpr int kTiles = 1;
    static __device__ __forceinline__ int group_rows(int) { return 4; }
};
template <>
struct TileGeom<1> {
    static constexpr int kTiles = 5;
    static __device__ __forceinline__ int group_rows(int) { return 20; }
};
template <>
struct TileGeom<2> {
    static constexpr int kTiles = 0;  // S / 4, at run time
    static __device__ __forceinline__ int group_rows(int S) { return S; }
};

// Loads the B operand (weights) of one wave into registers.
template <int A, int NT, bool SPLIT>
__device__ __forceinline__ void load_weights(const float2* __restrict__ w, int F, int B,
                                             int f, int beam0, int lane,
                                             bf_half8 (&bhi)[NT][A / 16],
                                             bf_half8 (&blo)[NT][A / 16]) {
    constexpr int KS = A / 16;
    const int c = lane & 15;
    const int g = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int b = beam0 + nt * 8 + (c >> 1);
        const bool comp_im = (c & 1) != 0;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            bf_half8 h, l;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int a = g * (A / 4) + 4 * s + j;
                float2 v = make_float2(0.0f, 0.0f);
                if (b < B) v = w[((size_t)b * F + f) * A + a];
                // column re: (wr, -wi); column im: (wi, wr)
                const float e0 = comp_im ? v.y : v.x;
                const float e1 = comp_im ? v.x : -v.y;
                const _Float16 h0 = (_Float16)e0;
                const _Float16 h1 = (_Float16)e1;
                h[2 * j] = h0;
                h[2 * j + 1] = h1;
                if (SPLIT) {
                    l[2 * j] = (_Float16)(e0 - (float)h0);
                    l[2 * j + 1] = (_Float16)(e1 - (float)h1);
                }
            }
            bhi[nt][s] = h;
            if (SPLIT) blo[nt][s] = l;
        }
    }
}

template <int A, int NT, bool SPLIT, int MODE>
__global__ __launch_bounds__(256) void mfma_fused_kernel(
    const uint8_t* __restrict__ x, const float2* __restrict__ w, float* __restrict__ out,
    int T, int F, int B, int S, int W, int n_supertiles, int st_per_chunk, int n_items,
    int per_xcd) {
    constexpr int KS = A / 16;         // k-steps of 32 (16 antennas x re/im)
    constexpr int BYTES = A / 4;       // packed bytes per lane per row
    const int lane = threadIdx.x & (kWaveSize - 1);
    const int wave = threadIdx.x / kWaveSize;
    // XCD-aware order: consecutive work items (adjacent frequencies of one
    // time chunk, which share 128-byte lines) run on the same XCD / L2.
    const int item = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
    if (item >= n_items) return;
    const int f = item % F;
    const int chunk = item / F;
    const int beam0 = wave * NT * 8;
    if (beam0 >= B) return;

    bf_half8 bhi[NT][KS];
    bf_half8 blo[NT][KS];
    load_weights<A, NT, SPLIT>(w, F, B, f, beam0, lane, bhi, blo);

    const int rho = lane & 15;  // A row / D column
    const int g = lane >> 4;    // k group of the A operand / row group of D
    const int grows = TileGeom<MODE>::group_rows(S);
    const int st_rows = 4 * grows;
    const int tiles = MODE == 2 ? S / 4 : TileGeom<MODE>::kTiles;
    const int a_row_off = (rho >> 2) * grows + (rho & 3);
    const uint8_t* xf = x + (size_t)f * A + g * BYTES;
    const size_t row_stride = (size_t)F * A;

    const int st_begin = chunk * st_per_chunk;
    int st_end = st_begin + st_per_chunk;
    st_end = st_end < n_supertiles ? st_end : n_supertiles;

    for (int st = st_begin; st < st_end; ++st) {
        const int base = st * st_rows;
        float win[NT][2];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            win[nt][0] = 0.0f;
            win[nt][1] = 0.0f;
        }
#pragma unroll(MODE == 1 ? 5 : 1)
        for (int i = 0; i < tiles; ++i) {
            const int t = base + a_row_off + 4 * i;
            unsigned d[BYTES / 4];
            RowLoad<A>::load(xf + (size_t)t * row_stride, t < T, d);
            bf_half8 av[KS];
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                unsigned h2[4];
                bf_int4x4_to_half2x4(d[s], h2);
                bf_uint4v u = {h2[0], h2[1], h2[2], h2[3]};
                av[s] = __builtin_bit_cast(bf_half8, u);
            }
            bf_float4 acc[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[nt] = bf_float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < KS; ++s) {
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av[s], bhi[nt][s], acc[nt],
                                                                     0, 0, 0);
                if (SPLIT) {
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av[s], blo[nt][s],
                                                                         acc[nt], 0, 0, 0);
                }
            }
            if (MODE == 0) {
                // One output per row: t = base + 4*g + r.
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const int b = beam0 + nt * 8 + (rho >> 1);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float sq = acc[nt][r] * acc[nt][r];
                        const float p = sq + dpp_swap_pairs(sq);
                        const int tt = base + 4 * g + r;
                        if ((rho & 1) == 0 && b < B && tt < T)
                            out[((size_t)tt * F + f) * B + b] = p;
                    }
                }
            } else {
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float v = acc[nt][r];
                        // MODE 1: local spectrum 4i+r of the group's 20 -> window (4i+r)/10.
                        if (MODE == 1 && 4 * i + r >= 10)
                            win[nt][1] = fmaf(v, v, win[nt][1]);
                        else
                            win[nt][0] = fmaf(v, v, win[nt][0]);
                    }
                }
            }
        }
        if (MODE != 0) {
            constexpr int kWinPerGroup = MODE == 1 ? 2 : 1;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int b = beam0 + nt * 8 + (rho >> 1);
#pragma unroll
                for (int q = 0; q < kWinPerGroup; ++q) {
                    const float p = win[nt][q] + dpp_swap_pairs(win[nt][q]);
                    const int wi = base / S + g * kWinPerGroup + q;
                    if ((rho & 1) == 0 && b < B && wi < W)
                        out[((size_t)wi * F + f) * B + b] = p;
                }
            }
        }
    }
}

template <int A, int NT, bool SPLIT>
void launch_mode(hipStream_t stream, const uint8_t* x, const float* w, float* out, int T,
                 int F, int B, int S, int mode) {
    const int st_rows = mode == 0 ? 16 : (mode == 1 ? 80 : 4 * S);
    const int W = mode == 0 ? T : (int)ceil_div((size_t)T, (size_t)S);
    const int n_st = (int)ceil_div((size_t)T, (size_t)st_rows);
    // ~1280 spectra per wave keeps the weight load amortised and gives
    // several waves per SIMD at the benchmark shapes.
    int st_per_chunk = (int)ceil_div(1280, (size_t)st_rows);
    if (st_per_chunk < 1) st_per_chunk = 1;
    const int n_chunks = (int)ceil_div((size_t)n_st, (size_t)st_per_chunk);
    const int n_items = n_chunks * F;
    const int per_xcd = (int)ceil_div((size_t)n_items, kXcds);
    const int waves = (int)ceil_div((size_t)B, (size_t)(NT * 8));
    const dim3 grid((unsigned)(per_xcd * kXcds));
    const dim3 block((unsigned)(waves * kWaveSize));
    const float2* w2 = reinterpret_cast<const float2*>(w);
    switch (mode) {
    case 0:
        hipLaunchKernelGGL((mfma_fused_kernel<A, NT, SPLIT, 0>), grid, block, 0, stream, x, w2,
                           out, T, F, B, S, W, n_st, st_per_chunk, n_items, per_xcd);
        break;
    case 1:
        hipLaunchKernelGGL((mfma_fused_kernel<A, NT, SPLIT, 1>), grid, block, 0, stream, x, w2,
                           out, T, F, B, S, W, n_st, st_per_chunk, n_items, per_xcd);
        break;
    default:
        hipLaunchKernelGGL((mfma_fused_kernel<A, NT, SPLIT, 2>), grid, block, 0, stream, x, w2,
                           out, T, F, B, S, W, n_st, st_per_chunk, n_items, per_xcd);
        break;
    }
}

template <bool SPLIT>
void launch_split(hipStream_t stream, const uint8_t* x, const float* w, float* out, int T,
                  int F, int A, int B, int S, int mode) {
    // Beams per wave chosen so the weights take 128 VGPRs (split) per lane.
    switch (A) {
    case 32: launch_mode<32, 8, SPLIT>(stream, x, w, out, T, F, B, S, mode); break;
    case 64: launch_mode<64, 4, SPLIT>(stream, x, w, out, T, F, B, S, mode); break;
    case 128: launch_mode<128, 2, SPLIT>(stream, x, w, out, T, F, B, S, mode); break;
    case 256: launch_mode<256, 1, SPLIT>(stream, x, w, out, T, F, B, S, mode); break;
    default: invalid("n_ant must be 32, 64, 128, or 256");
    }
}

}  // namespace

bool mfma_supported(size_t A, size_t B, size_t S) {
    (void)B;
    const bool a_ok = A == 32 || A == 64 || A == 128 || A == 256;
    const bool s_ok = S == 0 || S == 10 || (S % 4 == 0);
    return a_ok && s_ok;
}

void launch_mfma_fused(hipStream_t stream, bool split, const uint8_t* x, const float* w,
                       float* out, size_t T, size_t F, size_t A, size_t B, size_t S) {
    if (!mfma_supported(A, B, S))
        invalid("the MFMA beamformer supports integration of 10 spectra or a multiple of 4");
    if (T > (size_t)INT32_MAX / 2 || T * F > ((size_t)1 << 40))
        fail(BF_ERR_OVERFLOW, "MFMA beamformer grid exceeds the supported size");
    const int mode = S == 0 ? 0 : (S == 10 ? 1 : 2);
    if (split)
        launch_split<true>(stream, x, w, out, (int)T, (int)F, (int)A, (int)B, (int)S, mode);
    else
        launch_split<false>(stream, x, w, out, (int)T, (int)F, (int)A, (int)B, (int)S, mode);
    hip_check(hipGetLastError(), "mfma_fused_kernel launch");
}

}  // namespace bfhip
