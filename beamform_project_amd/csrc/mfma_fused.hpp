// mfma_fused.hpp -- FP16-MFMA fused fixed-grid beamformer for gfx950
// (kernel template; instantiated per antenna count in k_mfma_a*.hip).
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
// Instruction: v_mfma_f32_16x16x32_f16.  A workgroup of WV waves owns one
// frequency, one chunk of time and WV*NT*8 beams; each wave keeps the
// weights of NT N-tiles (8 beams each, re/im interleaved) in VGPRs for the
// whole chunk.  The packed int4 rows of a stage (ST tiles of 16 spectra)
// are loaded once per workgroup with coalesced dword loads, converted to
// FP16 pairs (byte permute, xor, one packed subtraction) and written to LDS
// in the MFMA A-operand layout, so each lane reads its operand with one
// 16-byte LDS load per k-step.  The next stage's global loads are in flight
// while the current stage is computed.  With DB the LDS holds two stages:
// the next stage is written while other waves still compute the current
// one, so a stage costs one barrier instead of two.
//
// Output layout of a 16x16 MFMA: lane l holds column l%16 for rows
// 4*(l/16)+r, r = 0..3.  Tile rows are mapped to time so that row group
// G = l/16 of a super-tile covers consecutive spectra: each lane then
// accumulates whole integration windows in registers.
//   MODE 0 (no integration): tile i, row 4G+r -> t = 16i + 4G + r
//   MODE 1 (S == 10): super-tile of 5 tiles, G covers 20 spectra (2 windows)
//   MODE 2 (S % 4 == 0): super-tile of S/4 tiles, G covers one window
// Re and im of a beam sit in adjacent lanes; a quad-permute DPP adds the
// two squared sums.
#pragma once

#include "bf_kernels.hpp"

namespace bfhip {

struct MfmaLaunch {
    const uint8_t* x = nullptr;
    const float* w = nullptr;
    float* out = nullptr;
    int T = 0, F = 0, B = 0, S = 0;
    bool split = true;
};

// Launch the kernel of `cfg` for one antenna count; false when `cfg` is
// not an instantiated shape (k_mfma_a32/64/128/256.hip).
bool launch_mfma_a32(hipStream_t stream, const MfmaLaunch& L, const MfmaConfig& cfg);
bool launch_mfma_a64(hipStream_t stream, const MfmaLaunch& L, const MfmaConfig& cfg);
bool launch_mfma_a128(hipStream_t stream, const MfmaLaunch& L, const MfmaConfig& cfg);
bool launch_mfma_a256(hipStream_t stream, const MfmaLaunch& L, const MfmaConfig& cfg);

namespace mfma_detail {

constexpr int kWaveSize = 64;
constexpr int kXcds = 8;

__device__ __forceinline__ float dpp_swap_pairs(float v) {
    // quad_perm [1,0,3,2]: lane i <- lane i^1.
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true));
}

// Time of row (G, r) of global tile i.
template <int MODE>
__device__ __forceinline__ int tile_time(int i, int G, int r, int S) {
    if (MODE == 0) return 16 * i + 4 * G + r;
    if (MODE == 1) return 80 * (i / 5) + 20 * G + 4 * (i % 5) + r;
    const int tps = S >> 2;
    return 4 * S * (i / tps) + S * G + 4 * (i % tps) + r;
}

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

template <int A, int NT, int WV, bool SPLIT, int MODE, int ST, bool DB>
__global__ __launch_bounds__(WV * kWaveSize) void mfma_fused_kernel(
    const uint8_t* __restrict__ x, const float2* __restrict__ w, float* __restrict__ out,
    int T, int F, int B, int S, int W, int tiles_total, int tiles_per_chunk, int n_items,
    int per_xcd) {
    constexpr int kThreads = WV * kWaveSize;
    constexpr int KS = A / 16;                      // k-steps of 32 (16 antennas x re/im)
    constexpr int TILE_BYTES = KS * kWaveSize * 16;  // LDS bytes per 16-row tile
    constexpr int STAGE_BYTES = ST * TILE_BYTES;
    constexpr int DW_ROW = A / 4;                    // packed dwords per row
    constexpr int DW_TILE = 16 * DW_ROW;
    constexpr int DW_STAGE = ST * DW_TILE;
    constexpr int DW_THREAD = (DW_STAGE + kThreads - 1) / kThreads;
    __shared__ __attribute__((aligned(16))) unsigned char lds[(DB ? 2 : 1) * STAGE_BYTES];

    const int tid = threadIdx.x;
    const int lane = tid & (kWaveSize - 1);
    const int wave = tid / kWaveSize;
    // XCD-aware order: workgroups are dispatched round-robin over the 8
    // XCDs, so blockIdx.x % 8 names the XCD.  Each XCD takes a contiguous
    // range of frequency-major items: all time chunks of a channel run on
    // one XCD back to back (its weights stay in that L2) and the
    // neighbouring channel, which shares the 128-byte lines of every packed
    // row, follows on the same XCD.
    const int item = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
    if (item >= n_items) return;  // whole workgroup
    const int n_chunks = n_items / F;
    const int f = item / n_chunks;
    const int chunk = item - f * n_chunks;
    const int beam0 = (blockIdx.y * WV + wave) * NT * 8;
    const bool has_beams = beam0 < B;

    bf_half8 bhi[NT][KS];
    bf_half8 blo[NT][KS];
    if (has_beams) load_weights<A, NT, SPLIT>(w, F, B, f, beam0, lane, bhi, blo);

    const int tile_begin = chunk * tiles_per_chunk;
    int tile_end = tile_begin + tiles_per_chunk;
    tile_end = tile_end < tiles_total ? tile_end : tiles_total;
    const size_t row_stride = (size_t)F * A;
    const uint8_t* xf = x + (size_t)f * A;

    // Stage loader: dword q of the stage -> (tile, row, dword in row).
    unsigned pre[DW_THREAD];
    auto load_stage = [&](int stage_tile0) {
#pragma unroll
        for (int k = 0; k < DW_THREAD; ++k) {
            const int q = tid + k * kThreads;
            unsigned v = 0u;
            if (q < DW_STAGE) {
                const int tl = q / DW_TILE;
                const int rem = q - tl * DW_TILE;
                const int rho = rem & 15;  // rows fastest: 16 lanes fill
                const int j = rem >> 4;    // 16 consecutive LDS slots
                const int t = tile_time<MODE>(stage_tile0 + tl, rho >> 2, rho & 3, S);
                if (t < T)
                    v = *reinterpret_cast<const unsigned*>(xf + (size_t)t * row_stride + 4 * j);
            }
            pre[k] = v;
        }
    };
    auto store_stage = [&](unsigned char* base) {
#pragma unroll
        for (int k = 0; k < DW_THREAD; ++k) {
            const int q = tid + k * kThreads;
            if (q < DW_STAGE) {
                const int tl = q / DW_TILE;
                const int rem = q - tl * DW_TILE;
                const int rho = rem & 15;  // rows fastest: 16 lanes fill
                const int j = rem >> 4;    // 16 consecutive LDS slots
                const int g = j / (A / 16);
                const int s = j - g * (A / 16);
                unsigned h2[4];
                bf_int4x4_to_half2x4(pre[k], h2);
                bf_uint4v u = {h2[0], h2[1], h2[2], h2[3]};
                *reinterpret_cast<bf_uint4v*>(base + tl * TILE_BYTES +
                                              (s * kWaveSize + rho + 16 * g) * 16) = u;
            }
        }
    };

    float win[NT][2];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) win[nt][0] = win[nt][1] = 0.0f;
    const int col = lane & 15;
    const int G = lane >> 4;
    const int tps = MODE == 2 ? (S >> 2) : 5;

    // MFMAs and epilogue of one stage held in LDS at `base`.
    auto compute_stage = [&](const unsigned char* base, int st0) {
        // One tile ahead: the A operand of tile tl+1 is read from LDS while
        // the MFMAs of tile tl run.
        bf_half8 av[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s)
            av[s] = *reinterpret_cast<const bf_half8*>(base + (s * kWaveSize + lane) * 16);
#pragma unroll 1
        for (int tl = 0; tl < ST; ++tl) {
            const int i = st0 + tl;
            bf_half8 an[KS];
            if (tl + 1 < ST) {
#pragma unroll
                for (int s = 0; s < KS; ++s)
                    an[s] = *reinterpret_cast<const bf_half8*>(
                        base + (tl + 1) * TILE_BYTES + (s * kWaveSize + lane) * 16);
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
#pragma unroll
            for (int s = 0; s < KS; ++s) av[s] = an[s];
            if (MODE == 0) {
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const int b = beam0 + nt * 8 + (col >> 1);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float sq = acc[nt][r] * acc[nt][r];
                        const float p = sq + dpp_swap_pairs(sq);
                        const int t = tile_time<0>(i, G, r, S);
                        if ((col & 1) == 0 && b < B && t < T)
                            out[((size_t)t * F + f) * B + b] = p;
                    }
                }
                continue;
            }
            const int ii = MODE == 1 ? tl : (i % tps);  // tile within super-tile
            if (MODE == 1 && tl == 2) {
                // spectra 8, 9 close window 0; 10, 11 open window 1
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    win[nt][0] = fmaf(acc[nt][0], acc[nt][0], win[nt][0]);
                    win[nt][0] = fmaf(acc[nt][1], acc[nt][1], win[nt][0]);
                    win[nt][1] = fmaf(acc[nt][2], acc[nt][2], win[nt][1]);
                    win[nt][1] = fmaf(acc[nt][3], acc[nt][3], win[nt][1]);
                }
            } else {
                const int q = (MODE == 1 && tl > 2) ? 1 : 0;  // wave-uniform
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    float a = 0.0f;
#pragma unroll
                    for (int r = 0; r < 4; ++r) a = fmaf(acc[nt][r], acc[nt][r], a);
                    // both updates, no dynamic index (x + 0 == x for these sums)
                    win[nt][0] += q ? 0.0f : a;
                    win[nt][1] += q ? a : 0.0f;
                }
            }
            if (ii == tps - 1) {
                constexpr int kWinPerGroup = MODE == 1 ? 2 : 1;
                const int super = i / tps;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const int b = beam0 + nt * 8 + (col >> 1);
#pragma unroll
                    for (int q = 0; q < kWinPerGroup; ++q) {
                        const float p = win[nt][q] + dpp_swap_pairs(win[nt][q]);
                        const int wi = super * 4 * kWinPerGroup + G * kWinPerGroup + q;
                        if ((col & 1) == 0 && b < B && wi < W)
                            out[((size_t)wi * F + f) * B + b] = p;
                        win[nt][q] = 0.0f;
                    }
                }
            }
        }
    };

    if (tile_begin < tile_end) load_stage(tile_begin);
    if (DB) {
        if (tile_begin < tile_end) store_stage(lds);
        __syncthreads();
        if (tile_begin + ST < tile_end) load_stage(tile_begin + ST);
        int buf = 0;
        for (int st0 = tile_begin; st0 < tile_end; st0 += ST) {
            if (has_beams) compute_stage(lds + buf * STAGE_BYTES, st0);
            // The other buffer was last read before the previous barrier.
            if (st0 + ST < tile_end) store_stage(lds + (buf ^ 1) * STAGE_BYTES);
            __syncthreads();
            if (st0 + 2 * ST < tile_end) load_stage(st0 + 2 * ST);
            buf ^= 1;
        }
    } else {
        for (int st0 = tile_begin; st0 < tile_end; st0 += ST) {
            __syncthreads();  // previous stage fully consumed
            store_stage(lds);
            __syncthreads();  // stage visible to all waves
            if (st0 + ST < tile_end) load_stage(st0 + ST);  // in flight during compute
            if (has_beams) compute_stage(lds, st0);
        }
    }
}

struct Plan {
    int mode, st, tiles_total, tiles_per_chunk, n_items, per_xcd, beam_blocks, W;
};

inline Plan make_plan(int T, int F, int B, int S, const MfmaConfig& cfg) {
    Plan p{};
    p.mode = S == 0 ? 0 : (S == 10 ? 1 : 2);
    int tiles_per_super = 1;
    if (p.mode == 0) {
        p.st = 4;
        p.W = T;
    } else if (p.mode == 1) {
        p.st = 5;
        tiles_per_super = 5;
        p.W = (int)ceil_div((size_t)T, 10);
    } else {
        tiles_per_super = S / 4;
        p.st = tiles_per_super % 4 == 0 ? 4 : 1;
        p.W = (int)ceil_div((size_t)T, (size_t)S);
    }
    const int rows_super = 16 * tiles_per_super;
    const int n_super = (int)ceil_div((size_t)T, (size_t)rows_super);
    p.tiles_total = n_super * tiles_per_super;
    if (p.mode == 0) p.tiles_total = (int)ceil_div((size_t)p.tiles_total, 4) * 4;
    p.beam_blocks = (int)ceil_div((size_t)B, (size_t)(cfg.waves * cfg.nt * 8));
    // Chunk length: a whole number of super-tiles and stages, chosen to
    // minimise (rounds of resident workgroups) x (tiles per workgroup + a
    // fixed start-up cost for loading the weights).
    int unit = tiles_per_super;
    while (unit % p.st) unit += tiles_per_super;
    const long resident = cfg.resident, startup_tiles = 12;
    long best_cost = -1;
    int per_chunk = unit;
    for (int c = unit; c <= 640 && c <= p.tiles_total + unit; c += unit) {
        const long chunks = (long)ceil_div((size_t)p.tiles_total, (size_t)c);
        const long rounds =
            (long)ceil_div((size_t)(chunks * F * p.beam_blocks), (size_t)resident);
        const long cost = rounds * (c + startup_tiles);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            per_chunk = c;
        }
    }
    p.tiles_per_chunk = per_chunk;
    const int n_chunks = (int)ceil_div((size_t)p.tiles_total, (size_t)per_chunk);
    p.n_items = n_chunks * F;
    p.per_xcd = (int)ceil_div((size_t)p.n_items, kXcds);
    return p;
}

template <int A, int NT, int WV, bool SPLIT, bool DB>
void launch_shape(hipStream_t stream, const MfmaLaunch& L, const MfmaConfig& cfg) {
    const Plan p = make_plan(L.T, L.F, L.B, L.S, cfg);
    const dim3 grid((unsigned)(p.per_xcd * kXcds), (unsigned)p.beam_blocks);
    const dim3 block(WV * kWaveSize);
    const float2* w2 = reinterpret_cast<const float2*>(L.w);
#define BF_LAUNCH(MODE_, ST_)                                                                 \
    hipLaunchKernelGGL((mfma_fused_kernel<A, NT, WV, SPLIT, MODE_, ST_, DB>), grid, block, 0, \
                       stream, L.x, w2, L.out, L.T, L.F, L.B, L.S, p.W, p.tiles_total,         \
                       p.tiles_per_chunk, p.n_items, p.per_xcd)
    if (p.mode == 0) {
        BF_LAUNCH(0, 4);
    } else if (p.mode == 1) {
        BF_LAUNCH(1, 5);
    } else if (p.st == 4) {
        BF_LAUNCH(2, 4);
    } else {
        BF_LAUNCH(2, 1);
    }
#undef BF_LAUNCH
}

template <int A, int NT, int WV>
void launch_split_db(hipStream_t stream, const MfmaLaunch& L, const MfmaConfig& cfg) {
    if (L.split) {
        if (cfg.db)
            launch_shape<A, NT, WV, true, true>(stream, L, cfg);
        else
            launch_shape<A, NT, WV, true, false>(stream, L, cfg);
    } else {
        if (cfg.db)
            launch_shape<A, NT, WV, false, true>(stream, L, cfg);
        else
            launch_shape<A, NT, WV, false, false>(stream, L, cfg);
    }
}

}  // namespace mfma_detail

// Instantiated shapes per antenna count: (NT, waves) pairs whose split
// weights fit in at most 128 VGPRs per lane and that compile without
// register spills for the common modes.
#define BF_MFMA_SHAPES_32_64(A_)                                            \
    BF_SHAPE(A_, 4, 4) BF_SHAPE(A_, 2, 8) BF_SHAPE(A_, 2, 4) BF_SHAPE(A_, 1, 8) \
    BF_SHAPE(A_, 1, 4)
#define BF_MFMA_SHAPES_128 BF_SHAPE(128, 2, 4) BF_SHAPE(128, 1, 8) BF_SHAPE(128, 1, 4)
#define BF_MFMA_SHAPES_256 BF_SHAPE(256, 1, 4)

#define BF_DEFINE_MFMA_FOR_A(A_, SHAPES)                                                   \
    bool launch_mfma_a##A_(hipStream_t stream, const MfmaLaunch& L, const MfmaConfig& cfg) { \
        SHAPES                                                                             \
        return false;                                                                      \
    }
#define BF_SHAPE(A_, NT_, WV_)                                               \
    if (cfg.nt == NT_ && cfg.waves == WV_) {                                 \
        mfma_detail::launch_split_db<A_, NT_, WV_>(stream, L, cfg);          \
        return true;                                                         \
    }

}  // namespace bfhip
