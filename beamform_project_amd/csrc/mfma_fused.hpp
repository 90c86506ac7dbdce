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
// 16-byte LDS load per k-step.  Every staging address is a per-thread
// constant plus a wave-uniform (scalar) stage offset.  The next stage's
// global loads are in flight while the current stage is computed.  With
// DB the LDS holds two stages and the conversion and stores of the next
// stage, and the loads of the one after, are spread over the tiles of the
// current stage, inside the same (fully unrolled) block as its MFMAs, so
// this VALU work issues between MFMAs; a stage costs one barrier.  The
// window bookkeeping of a tile is static in the unrolled stage, and the
// two lanes of a re/im pair split each window's stores.
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

#include <type_traits>

#include "bf_kernels.hpp"
#include "mfma_stage.hpp"

namespace bfhip {

struct MfmaLaunch {
    const uint8_t* x = nullptr;
    const float* w = nullptr;
    float* out = nullptr;
    const uint8_t* mask = nullptr;  // [T][F] frame validity (0 = lost), or null
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

// At least two waves per SIMD (at most 256 registers per lane): the
// unrolled stage otherwise keeps several tiles' accumulators in flight.
template <int A, int NT, int WV, bool SPLIT, int MODE, int ST, bool DB, bool MASKED = false>
__global__ __launch_bounds__(WV * kWaveSize) __attribute__((amdgpu_waves_per_eu(2)))
void mfma_fused_kernel(
    const uint8_t* __restrict__ x, const float2* __restrict__ w, float* __restrict__ out,
    const uint8_t* __restrict__ mask, int T, int F, int B, int S, int W, int tiles_total,
    int tiles_per_chunk, int n_items, int per_xcd) {
    constexpr int kThreads = WV * kWaveSize;
    constexpr int KS = A / 16;                      // k-steps of 32 (16 antennas x re/im)
    using Map = StageMap<A, kThreads>;
    constexpr int TILE_BYTES = Map::TILE_BYTES;     // LDS bytes per 16-row tile
    constexpr int STAGE_BYTES = ST * TILE_BYTES;
    constexpr int DW_TILE = Map::DW_TILE;           // packed dwords per tile
    constexpr int DW_STAGE = ST * DW_TILE;
    constexpr int DW_THREAD = (DW_STAGE + kThreads - 1) / kThreads;
    constexpr bool ALL_VALID = DW_STAGE % kThreads == 0;
    // MODE 1 stages hold whole super-tiles of 5 tiles
    static_assert(MODE != 1 || ST % 5 == 0, "S = 10 stages are whole super-tiles");
    static_assert(kThreads % DW_TILE == 0 || DW_TILE % kThreads == 0, "power-of-two staging");
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
    const int tps = MODE == 2 ? (S >> 2) : 5;           // tiles per super-tile
    const int pitch = MODE == 0 ? 4 : (MODE == 1 ? 20 : S);  // spectra between row groups

    // Per-thread staging constants (StageMap: rows fastest within 16 lanes).
    const int tl0 = Map::tile0(wave);  // uniform per wave
    const int rem0 = Map::rem0(tid);
    const int rho = rem0 & 15;
    const int j0 = rem0 >> 4;
    // spectrum of this thread's row within its tile's row group pattern
    const int prow = (rho >> 2) * pitch + (rho & 3);
    const uint8_t* xs = x + (size_t)f * A + (size_t)prow * row_stride + 4 * j0;
    const int lds_mine = Map::lds_offset(tl0, rem0);

    // First spectrum of the stage starting at tile st0, and offset of tile
    // tl of a stage from it (row group 0, row 0).
    auto stage_time = [&](int st0) -> int {
        if (MODE == 0 || MODE == 1) return 16 * st0;  // MODE 1: 80 per super-tile
        return 4 * S * (st0 / tps) + 4 * (st0 % tps);
    };
    auto tile_rows = [&](int tl) -> int {
        if (MODE == 0) return 16 * tl;
        if (MODE == 1) return ST == 5 ? 4 * tl : 80 * (tl / 5) + 4 * (tl % 5);
        return 4 * tl;
    };
    // Whether every row of the stage at st0 lies before T.
    auto stage_full = [&](int st0) -> bool {
        return stage_time(st0) + tile_rows(ST - 1) + 3 * pitch + 3 < T;
    };
    auto dword_valid = [&](int k) -> bool {
        return ALL_VALID || tl0 + Map::tld(k) < ST;
    };
    unsigned pre[DW_THREAD];
    // Global load of dword k of the stage whose first spectrum is t_stage;
    // GUARD: rows at or after t_lim, and with MASKED rows of lost frames
    // (mask byte 0), read as zero.  MASKED kernels always take this path.
    auto load_dword = [&](auto guard, int k, int t_stage, int t_lim) {
        constexpr bool GUARD = decltype(guard)::value;
        const int rel = t_stage + tile_rows(tl0 + Map::tld(k));
        const uint8_t* p = xs + (size_t)rel * row_stride + Map::byte_delta(k);
        unsigned v = 0u;
        if (GUARD) {
            const int t = prow + rel;
            if (dword_valid(k) && t < t_lim && (!MASKED || mask[(size_t)t * F + f] != 0))
                v = *reinterpret_cast<const unsigned*>(p);
        } else if (dword_valid(k)) {
            v = *reinterpret_cast<const unsigned*>(p);
        }
        pre[k] = v;
    };
    auto store_dword = [&](int k, unsigned char* buf) {
        if (!dword_valid(k)) return;
        unsigned h2[4];
        bf_int4x4_to_half2x4(pre[k], h2);
        bf_uint4v u = {h2[0], h2[1], h2[2], h2[3]};
        *reinterpret_cast<bf_uint4v*>(buf + lds_mine + Map::lds_delta(k)) = u;
    };
    auto load_stage = [&](int st0) {
        const int ts = stage_time(st0);
        if (!MASKED && stage_full(st0)) {
#pragma unroll
            for (int k = 0; k < DW_THREAD; ++k) load_dword(std::false_type{}, k, ts, T);
        } else {
#pragma unroll
            for (int k = 0; k < DW_THREAD; ++k) load_dword(std::true_type{}, k, ts, T);
        }
    };

    float win[NT][2];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) win[nt][0] = win[nt][1] = 0.0f;
    const int col = lane & 15;
    const int G = lane >> 4;
    const int odd = col & 1;
    // Output pointer of this lane at window (or spectrum) group 0.  Lanes
    // 2c and 2c+1 hold the same power after the pair swap, so they split
    // the stores: MODE 1 by window of the pair (q = odd), MODE 2 by N-tile
    // (nt + odd), MODE 0 by row (r = odd + 2m).
    const int bl = beam0 + (col >> 1);
    float* const out_lane =
        MODE == 1 ? out + ((size_t)(2 * G + odd) * F + f) * B + bl
        : MODE == 2 ? out + ((size_t)G * F + f) * B + bl + (NT >= 2 ? 8 * odd : 0)
                    : out + ((size_t)(4 * G + odd) * F + f) * B + bl;
    const size_t group_stride = (size_t)F * B;  // floats between consecutive windows

    // Powers and window sums of tile tl (global tile i) from acc.
    auto epilogue = [&](int tl, int i, const bf_float4 (&acc)[NT]) {
        if (MODE == 0) {
            float* const o = out_lane + (size_t)(16 * i) * group_stride;
            const int t0 = 16 * i + 4 * G + odd;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                float p[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float sq = acc[nt][r] * acc[nt][r];
                    p[r] = sq + dpp_swap_pairs(sq);
                }
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                    const float v = odd ? p[2 * m + 1] : p[2 * m];
                    if (bl + nt * 8 < B && t0 + 2 * m < T)
                        o[(size_t)(2 * m) * group_stride + nt * 8] = v;
                }
            }
            return;
        }
        if (MODE == 1 && tl % 5 == 2) {
            // spectra 8, 9 close window 0; 10, 11 open window 1
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                win[nt][0] = fmaf(acc[nt][0], acc[nt][0], win[nt][0]);
                win[nt][0] = fmaf(acc[nt][1], acc[nt][1], win[nt][0]);
                win[nt][1] = fmaf(acc[nt][2], acc[nt][2], win[nt][1]);
                win[nt][1] = fmaf(acc[nt][3], acc[nt][3], win[nt][1]);
            }
            return;
        }
        const int q = (MODE == 1 && tl % 5 > 2) ? 1 : 0;  // static once the tile loop is unrolled
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            float a = acc[nt][0] * acc[nt][0];
#pragma unroll
            for (int r = 1; r < 4; ++r) a = fmaf(acc[nt][r], acc[nt][r], a);
            if (q == 0)
                win[nt][0] += a;
            else
                win[nt][1] += a;
        }
    };
    // Stores and resets the windows of the super-tile (MODE 1) or stage
    // (MODE 2) that starts at tile i0.
    auto close_windows = [&](int i0) {
        if (MODE == 1) {
            const int super = i0 / 5;
            float* const o = out_lane + (size_t)(8 * super) * group_stride;
            const bool ok = 8 * super + 2 * G + odd < W;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const float p0 = win[nt][0] + dpp_swap_pairs(win[nt][0]);
                const float p1 = win[nt][1] + dpp_swap_pairs(win[nt][1]);
                if (ok && bl + nt * 8 < B) o[nt * 8] = odd ? p1 : p0;
                win[nt][0] = win[nt][1] = 0.0f;
            }
        } else if (MODE == 2) {
            const int super = i0 / tps;
            float* const o = out_lane + (size_t)(4 * super) * group_stride;
            const bool ok = 4 * super + G < W;
            float p[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                p[nt] = win[nt][0] + dpp_swap_pairs(win[nt][0]);
                win[nt][0] = 0.0f;
            }
            if (NT == 1) {
                if (ok && !odd && bl < B) o[0] = p[0];
            } else {
#pragma unroll
                for (int nt = 0; nt < NT; nt += 2)
                    if (ok && bl + (nt + odd) * 8 < B) o[nt * 8] = odd ? p[nt + 1] : p[nt];
            }
        }
    };
    // MODE 2: whether the stage at st0 ends a super-tile (MODE 1 closes
    // its windows inside the stage, after every fifth tile).
    auto closes_window = [&](int st0) -> bool {
        return MODE == 2 && (st0 % tps) + ST == tps;
    };

    // MFMAs and epilogues of the stage held at `cur`; with DB, the dwords
    // of the next stage (in pre) are converted into `nxt` and the loads of
    // the stage after it are issued, spread over the tiles so that this
    // VALU and memory work interleaves with the MFMAs.
    auto stage = [&](auto mf, auto guard, const unsigned char* cur, unsigned char* nxt, int st0,
                     int t_load, int t_lim) {
        constexpr bool MF = decltype(mf)::value;
        // Up to 64 antennas the next tile's A operands are read while this
        // tile's MFMAs run; wider tiles read them per k-step (a second
        // tile of operands would not fit in registers).
        constexpr bool PREFETCH = KS <= 4;
        bf_half8 av[KS];
        if (MF && PREFETCH) {
#pragma unroll
            for (int s = 0; s < KS; ++s)
                av[s] = *reinterpret_cast<const bf_half8*>(cur + (s * kWaveSize + lane) * 16);
        }
#pragma unroll
        for (int tl = 0; tl < ST; ++tl) {
            bf_half8 an[KS];
            bf_float4 acc[NT];
            if (MF) {
                if (PREFETCH && tl + 1 < ST) {
#pragma unroll
                    for (int s = 0; s < KS; ++s)
                        an[s] = *reinterpret_cast<const bf_half8*>(
                            cur + (tl + 1) * TILE_BYTES + (s * kWaveSize + lane) * 16);
                }
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[nt] = bf_float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const bf_half8 a_s =
                        PREFETCH ? av[s]
                                 : *reinterpret_cast<const bf_half8*>(
                                       cur + tl * TILE_BYTES + (s * kWaveSize + lane) * 16);
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_s, bhi[nt][s],
                                                                         acc[nt], 0, 0, 0);
                    if (SPLIT) {
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt)
                            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_s, blo[nt][s],
                                                                             acc[nt], 0, 0, 0);
                    }
                }
            }
            if (DB) {
#pragma unroll
                for (int k = tl * DW_THREAD / ST; k < (tl + 1) * DW_THREAD / ST; ++k) {
                    store_dword(k, nxt);
                    load_dword(guard, k, t_load, t_lim);
                }
            }
            if (MF) {
                epilogue(tl, st0 + tl, acc);
                if (MODE == 1 && tl % 5 == 4) close_windows(st0 + tl - 4);
                if (PREFETCH && tl + 1 < ST) {
#pragma unroll
                    for (int s = 0; s < KS; ++s) av[s] = an[s];
                }
            }
            // Wide tiles: keep the scheduler from hoisting the next tile's
            // operand reads into this one (registers).
            if (!PREFETCH) __builtin_amdgcn_sched_barrier(0);
        }
    };

    if (DB) {
        if (tile_begin < tile_end) {
            load_stage(tile_begin);
#pragma unroll
            for (int k = 0; k < DW_THREAD; ++k) store_dword(k, lds);
        }
        if (tile_begin + ST < tile_end) load_stage(tile_begin + ST);
        __syncthreads();
        int buf = 0;
        for (int st0 = tile_begin; st0 < tile_end; st0 += ST) {
            const unsigned char* cur = lds + buf * STAGE_BYTES;
            unsigned char* nxt = lds + (buf ^ 1) * STAGE_BYTES;
            // The other buffer was last read before the previous barrier.
            // Loads of the stage after next: unguarded when it exists and
            // lies before T; otherwise guarded (t_lim 0 when it does not exist).
            const int st2 = st0 + 2 * ST;
            const int t_load = stage_time(st2);
            if (!MASKED && st2 < tile_end && stage_full(st2)) {
                if (has_beams)
                    stage(std::true_type{}, std::false_type{}, cur, nxt, st0, t_load, T);
                else
                    stage(std::false_type{}, std::false_type{}, cur, nxt, st0, t_load, T);
            } else {
                const int t_lim = st2 < tile_end ? T : 0;
                if (has_beams)
                    stage(std::true_type{}, std::true_type{}, cur, nxt, st0, t_load, t_lim);
                else
                    stage(std::false_type{}, std::true_type{}, cur, nxt, st0, t_load, t_lim);
            }
            if (has_beams && closes_window(st0)) close_windows(st0);
            __syncthreads();
            buf ^= 1;
        }
    } else {
        if (tile_begin < tile_end) load_stage(tile_begin);
        for (int st0 = tile_begin; st0 < tile_end; st0 += ST) {
            __syncthreads();  // previous stage fully consumed
#pragma unroll
            for (int k = 0; k < DW_THREAD; ++k) store_dword(k, lds);
            if (st0 + ST < tile_end) load_stage(st0 + ST);  // in flight during compute
            __syncthreads();  // stage visible to all waves
            if (has_beams) {
                stage(std::true_type{}, std::false_type{}, lds, lds, st0, 0, 0);
                if (closes_window(st0)) close_windows(st0);
            }
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

template <int A, int NT, int WV, bool SPLIT, bool DB, bool MASKED = false>
void launch_shape(hipStream_t stream, const MfmaLaunch& L, const MfmaConfig& cfg) {
    const Plan p = make_plan(L.T, L.F, L.B, L.S, cfg);
    const dim3 grid((unsigned)(p.per_xcd * kXcds), (unsigned)p.beam_blocks);
    const dim3 block(WV * kWaveSize);
    const float2* w2 = reinterpret_cast<const float2*>(L.w);
#define BF_LAUNCH(MODE_, ST_)                                                                 \
    hipLaunchKernelGGL((mfma_fused_kernel<A, NT, WV, SPLIT, MODE_, ST_, DB, MASKED>), grid,   \
                       block, 0,                                                               \
                       stream, L.x, w2, L.out, L.mask, L.T, L.F, L.B, L.S, p.W, p.tiles_total, \
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

// A launch with a loss mask runs the one masked instantiation per antenna
// count: 8 beams per wave, 4 waves, one LDS stage (any beam count).
#define BF_DEFINE_MFMA_FOR_A(A_, SHAPES)                                                   \
    bool launch_mfma_a##A_(hipStream_t stream, const MfmaLaunch& L, const MfmaConfig& cfg) { \
        if (L.mask != nullptr) {                                                           \
            MfmaConfig m = cfg;                                                            \
            m.nt = 1;                                                                      \
            m.waves = 4;                                                                   \
            m.db = 0;                                                                      \
            if (L.split)                                                                   \
                mfma_detail::launch_shape<A_, 1, 4, true, false, true>(stream, L, m);      \
            else                                                                           \
                mfma_detail::launch_shape<A_, 1, 4, false, false, true>(stream, L, m);     \
            return true;                                                                   \
        }                                                                                  \
        SHAPES                                                                             \
        return false;                                                                      \
    }
#define BF_SHAPE(A_, NT_, WV_)                                               \
    if (cfg.nt == NT_ && cfg.waves == WV_) {                                 \
        mfma_detail::launch_split_db<A_, NT_, WV_>(stream, L, cfg);          \
        return true;                                                         \
    }

}  // namespace bfhip
