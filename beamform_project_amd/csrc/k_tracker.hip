// k_tracker.hip -- dynamic beam tracker kernels for gfx950.
//
// The tracker re-points every beam once per integration window: window w
// uses the direction of its first spectrum (reference
// src/beam_tracker.cpp:166-190) and the steering weights
// w[b][f][a] = (float)cos(phi), (float)sin(phi) with the phase
// phi = (2 pi f / c) * (p_a . d) evaluated in FP64 (src/weights.cpp:37-55),
// then contracts like the fixed grid (src/cpu_beamformer.cpp:56-78).
// Weights are generated inside the kernels, per window, never stored.  The
// FP32 kernel evaluates cos/sin in FP64 as the reference does.  The MFMA
// kernels form the same FP64 phase, reduce it to [-pi, pi] in FP64 and
// evaluate it in FP32 with a first-order correction (about one float ulp
// from (float)cos(phi), below the FP16 split's own rounding, at a fraction
// of an FP64 sincos); for windows shorter than 64 spectra (where weight
// generation would dominate) they do so for every 16th window and rotate
// it by the small phase change to the windows in between.
//
// This translation unit is compiled without FMA contraction so that the
// direction (float) and phase (double) arithmetic rounds exactly like the
// reference's ISO C++ build; the accumulations use explicit fmaf.
//
// Direction sources:
//   DIR_TRAJECTORY  per-beam linear drift (l0, m0, rate_l, rate_m):
//                   l = l0 + (float)t * rate_l, n = sqrt(1 - l^2 - m^2)
//                   (src/beam_tracker.cpp:166-179, src/geometry.cpp:97-106)
//   DIR_TABLE       a device table of unit vectors [window][beam][3]
//
// MFMA kernel: the fixed-grid kernel's scheme (v_mfma_f32_16x16x32_f16,
// LDS-shared FP16 A operands, weights in VGPRs, split hi+lo) with time
// tiles aligned to windows: a window of S spectra covers ceil(S/16) tiles
// of 16 consecutive spectra (rows past the window are zero), so one B
// operand serves a whole tile.  Integrated output folds the four row groups
// of a lane column with two cross-lane adds at the window's last tile.
#include <algorithm>

#include "bf_kernels.hpp"
#include "bf_serve.hpp"
#include "mfma_stage.hpp"
#include "bf_tracker.hpp"

namespace bfhip {
namespace {

constexpr int kWaveSize = 64;
constexpr int kWaves = 4;
constexpr int kThreads = kWaves * kWaveSize;
constexpr int kXcds = 8;

__device__ __forceinline__ float dpp_swap_pairs(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true));
}

struct DirArgs {
    const float* table;  // DIR_TABLE: [n_windows_table][B][3]
    const float* traj;   // DIR_TRAJECTORY: [B][4] = l0, m0, rate_l, rate_m
    long window_offset;  // absolute index of local window 0
    int S;
    int B;
};

// Direction of beam b in local window w.
template <int DIRSRC>
__device__ __forceinline__ void beam_direction(const DirArgs& da, int w, int b, float d[3]) {
    if (DIRSRC == kDirTable) {
        const float* p = da.table + ((size_t)w * da.B + b) * 3;
        d[0] = p[0];
        d[1] = p[1];
        d[2] = p[2];
    } else {
        const float* q = da.traj + (size_t)b * 4;
        const float tf = (float)((da.window_offset + w) * (long)da.S);
        const float l = q[0] + tf * q[2];
        const float m = q[1] + tf * q[3];
        const float t2 = l * l + m * m;
        d[0] = l;
        d[1] = m;
        d[2] = __builtin_sqrtf(1.0f - t2);
    }
}

// (float)cos(phi), (float)sin(phi) of phi = k * (p . d) in FP64.
__device__ __forceinline__ float2 steering_weight(const float p[3], const float d[3], double k) {
    const double delay = (double)p[0] * d[0] + (double)p[1] * d[1] + (double)p[2] * d[2];
    const double phase = k * delay;
    double s, c;
    sincos(phase, &s, &c);
    return make_float2((float)c, (float)s);
}

__device__ __forceinline__ double wave_number(float f_hz) {
    return kTwoPi * (double)f_hz / kSpeedOfLight;
}

// Antenna position: a device table, or the reference's regular grid
// (src/geometry.cpp:33-67: x = column * spacing, y = row * spacing).
__device__ __forceinline__ void antenna_position(const Geom& geo, int a, float p[3]) {
    if (geo.pos != nullptr) {
        p[0] = geo.pos[3 * a];
        p[1] = geo.pos[3 * a + 1];
        p[2] = geo.pos[3 * a + 2];
    } else {
        p[0] = (float)(a % geo.cols) * geo.spacing;
        p[1] = (float)(a / geo.cols) * geo.spacing;
        p[2] = 0.0f;
    }
}

// Positions of all antennas, [A][3], copied into LDS by every workgroup
// that generates weights (read many times per window).
template <int A>
__device__ __forceinline__ void stage_positions(const Geom& geo, float* s_pos) {
    for (int a = threadIdx.x; a < A; a += blockDim.x) {
        float p[3];
        antenna_position(geo, a, p);
        s_pos[3 * a] = p[0];
        s_pos[3 * a + 1] = p[1];
        s_pos[3 * a + 2] = p[2];
    }
}

// Channel frequency: a device table, or src/geometry.cpp:79-95.
__device__ __forceinline__ float channel_frequency(const Geom& geo, int f) {
    if (geo.freqs != nullptr) return geo.freqs[f];
    const float step = (float)f * geo.f_step;
    return geo.f_start + step;
}

// cos/sin of phi = k * (p . d) for the MFMA kernels: the phase is formed and
// reduced to [-pi, pi] in FP64, then evaluated in FP32 with a first-order
// correction for the rounding of the reduced argument, so the weights
// differ from (float)cos(phi), (float)sin(phi) by about one float ulp --
// below the FP16 split's own rounding -- at a fraction of an FP64 sincos.
__device__ __forceinline__ float2 steering_weight_reduced(const float p[3], const float d[3],
                                                         double k) {
    const double delay = (double)p[0] * d[0] + (double)p[1] * d[1] + (double)p[2] * d[2];
    const double phase = k * delay;
    const double turns = rint(phase * (1.0 / kTwoPi));
    const double r = phase - turns * kTwoPi;
    const float rh = (float)r;
    const float rl = (float)(r - (double)rh);  // |rl| <= 2^-23
    float sn, cs;
    sincosf(rh, &sn, &cs);
    // sin(rh + rl) = sin rh + rl cos rh, cos(rh + rl) = cos rh - rl sin rh
    return make_float2(fmaf(-rl, sn, cs), fmaf(rl, cs, sn));
}

// Windows of at least this many spectra evaluate every window's weights
// from its own phase (steering_weight_reduced).  Shorter windows, where
// weight generation would dominate, use anchored weights: every kAnchorWindows-th window (by
// window index) is an anchor evaluated with steering_weight_reduced, and
// window w rotates the anchor's phasor by the phase change
// delta = k * (p . (d_w - d_anchor)), formed in FP32 from the small
// direction difference.  Each window's weights depend only on w and its
// anchor, never on how windows are grouped into launches.
constexpr int kDirectWeightMinSpectra = 64;
constexpr int kAnchorWindows = 16;

// Anchor phasors of this lane's own weights and the anchor's directions.
template <int A, int NT>
struct WeightAnchor {
    float2 own[NT][A / 16][2];
    float dir[NT][3];
    int window;  // -1: none yet
};

// (cos, sin)(phi + delta) from (cos, sin)(phi) = z; |delta| < 2^-10 uses
// cos delta = 1 - delta^2/2, sin delta = delta - delta^3/6 (the next terms
// are below 2^-40 relative).
__device__ __forceinline__ float2 rotate_phasor(float2 z, float delta) {
    float sd, cd;
    if (fabsf(delta) < 0x1p-10f) {
        const float d2 = delta * delta;
        cd = fmaf(d2, -0.5f, 1.0f);
        sd = fmaf(delta * d2, -1.0f / 6.0f, delta);
    } else {
        sincosf(delta, &sd, &cd);
    }
    return make_float2(fmaf(z.x, cd, -(z.y * sd)), fmaf(z.y, cd, z.x * sd));
}

// hi/lo FP16 pairs of one complex weight, packed (re in the low half):
// ph = (f16(wr), f16(wi)), pl = (f16(wr - hi), f16(wi - hi)).
__device__ __forceinline__ void split_pack(float2 v, bool split, unsigned& ph, unsigned& pl) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const _Float16 hr = (_Float16)v.x;
    const _Float16 hi = (_Float16)v.y;
    ph = __builtin_bit_cast(unsigned, h2{hr, hi});
    pl = 0u;
    if (split) {
        const _Float16 lr = (_Float16)(v.x - (float)hr);
        const _Float16 li = (_Float16)(v.y - (float)hi);
        pl = __builtin_bit_cast(unsigned, h2{lr, li});
    }
}

__device__ __forceinline__ unsigned dpp_swap_pairs_u(unsigned v) {
    return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);
}

// B operand of one wave for window w.  Lane c holds column c of each
// 8-beam N-tile (Re of beam c/2 for even c, Im for odd c) over antennas
// g*(A/4) + 4s + j.  Lanes c and c^1 need the same complex weights, so each
// evaluates two of every four (j = 2*comp_im + jj), splits them into hi/lo
// FP16 pairs and takes the partner's two pairs by DPP; a Re column then
// holds (w_re, -w_im) (sign flip of the high half) and an Im column
// (w_im, w_re) (halves swapped).  Without ANCHORED (or with `direct`) every
// weight comes from steering_weight_reduced; with ANCHORED the anchor window's
// phasors are kept in `anc` and rotated to window w.  Positions come from
// the LDS copy s_pos.
template <int A, int NT, bool SPLIT, int DIRSRC, bool ANCHORED>
__device__ __forceinline__ void window_operands(const DirArgs& da, const float* s_pos,
                                                double k, int w, int beam0, int lane,
                                                bool direct, WeightAnchor<A, NT>& anc,
                                                bf_half8 (&bhi)[NT][A / 16],
                                                bf_half8 (&blo)[NT][A / 16]) {
    constexpr int KS = A / 16;
    const int c = lane & 15;
    const int g = lane >> 4;
    const int comp_im = c & 1;
    const bool use_anchor = ANCHORED && !direct;
    const int wa = w - w % kAnchorWindows;
    if (use_anchor && anc.window != wa) {
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int b = beam0 + nt * 8 + (c >> 1);
            const bool live = b < da.B;
            float d[3] = {0.0f, 0.0f, 1.0f};
            if (live) beam_direction<DIRSRC>(da, wa, b, d);
            anc.dir[nt][0] = d[0];
            anc.dir[nt][1] = d[1];
            anc.dir[nt][2] = d[2];
#pragma unroll
            for (int s = 0; s < KS; ++s) {
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) {
                    anc.own[nt][s][jj] = make_float2(0.0f, 0.0f);
                    if (live) {
                        const float* p = s_pos + 3 * (g * (A / 4) + 4 * s + 2 * comp_im + jj);
                        anc.own[nt][s][jj] = steering_weight_reduced(p, d, k);
                    }
                }
            }
        }
        anc.window = wa;
    }
    const float kf = (float)k;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int b = beam0 + nt * 8 + (c >> 1);
        const bool live = b < da.B;
        float d[3] = {0.0f, 0.0f, 1.0f};
        if (live) beam_direction<DIRSRC>(da, w, b, d);
        const bool rotate = use_anchor && live && w != wa;
        float dd[3] = {0.0f, 0.0f, 0.0f};
        if (rotate) {
            dd[0] = d[0] - anc.dir[nt][0];
            dd[1] = d[1] - anc.dir[nt][1];
            dd[2] = d[2] - anc.dir[nt][2];
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            unsigned ph[2], pl[2];
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                const float* p = s_pos + 3 * (g * (A / 4) + 4 * s + 2 * comp_im + jj);
                float2 v;
                if (use_anchor) {
                    v = anc.own[nt][s][jj];
                    if (rotate) {
                        float t = p[0] * dd[0];
                        t = fmaf(p[1], dd[1], t);
                        t = fmaf(p[2], dd[2], t);
                        v = rotate_phasor(v, kf * t);
                    }
                } else {
                    v = make_float2(0.0f, 0.0f);
                    if (live) v = steering_weight_reduced(p, d, k);
                }
                split_pack(v, SPLIT, ph[jj], pl[jj]);
            }
            unsigned qh[2], ql[2];  // partner's pairs
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                qh[jj] = dpp_swap_pairs_u(ph[jj]);
                if (SPLIT) ql[jj] = dpp_swap_pairs_u(pl[jj]);
            }
            // Re column: (re, -im) = sign flip of the high half; Im column:
            // (im, re) = halves rotated.
            const unsigned rot = comp_im ? 16u : 0u;
            const unsigned flip = comp_im ? 0u : 0x80000000u;
            bf_uint4v uh, ul;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool mine = (j >> 1) == comp_im;
                const unsigned xh = mine ? ph[j & 1] : qh[j & 1];
                uh[j] = __builtin_amdgcn_alignbit(xh, xh, rot) ^ flip;
                if (SPLIT) {
                    const unsigned xl = mine ? pl[j & 1] : ql[j & 1];
                    ul[j] = __builtin_amdgcn_alignbit(xl, xl, rot) ^ flip;
                }
            }
            bhi[nt][s] = __builtin_bit_cast(bf_half8, uh);
            if (SPLIT) blo[nt][s] = __builtin_bit_cast(bf_half8, ul);
        }
    }
}

// Time rows of one channel.  Spectrum t (absolute) lives at
// x_f + (t & ring_mask) * row_stride: a linear [T][F][A] array uses
// ring_mask = -1, a ring of 2^k spectra uses 2^k - 1.  Window w covers
// spectra [t_base + w*S, t_base + (w+1)*S) clipped to t_limit.
// With a loss mask, frame (t, f) is valid when valid_f[(t & ring_mask) * F]
// is non-zero; a lost frame's row reads as zero.
struct RowSource {
    const uint8_t* x_f;  // x + f * A
    size_t row_stride;   // F * A bytes
    long ring_mask;
    long t_base;
    long t_limit;
    const uint8_t* valid_f = nullptr;  // mask + f, [t][F] bytes (MASKED only)
    size_t mask_stride = 0;             // F
};

// Output of one channel: integrated [slot][F][B] with slot = w mod
// out_windows (0: slot = w), or per-spectrum [t - t_base][F][B].
struct OutTarget {
    float* out;
    long out_windows;
    int F;
    int f;
};

// A tile of a tracker call: its first spectrum (absolute) and how many of
// its 16 rows hold spectra of the window (0 past the call's tiles).
struct TileRef {
    long t;
    int rows;
};

// Windows [w_begin, w_end) of one channel for the beams wg_beam0 ..
// wg_beam0 + kWaves*NT*8 - 1.  Entered by the whole workgroup (kThreads
// threads): stages of ST window-aligned tiles of 16 spectra go through LDS
// in the A-operand layout (two stage buffers, one barrier per stage).  The
// packed rows of stage s + 2 are loaded at the start of stage s and those
// of stage s + 1, loaded a stage earlier, are converted into the other
// buffer at its end, so every load has two stages of compute to land.
// Each thread's rows sit at a per-thread offset from a wave-uniform tile
// base; a stage whose tiles are all full and do not wrap the ring loads
// without guards.  The B operand (a window's weights) is regenerated at
// each window's first tile.  Each wave owns one block of NT*8 beams; when a
// window fits one tile and fewer than kWaves blocks hold beams, the idle
// waves share the blocks and the tiles of each stage are dealt out among a
// block's waves, so every wave generates weights.
template <int A, int NT, bool SPLIT, bool INTEGRATE, int DIRSRC, int ST, bool ANCHORED,
          bool MASKED = false>
__device__ __forceinline__ void tracker_windows(const RowSource& rs, const float* s_pos,
                                                const DirArgs& da, const OutTarget& ot, double k,
                                                int w_begin, int w_end, int tpw, int wg_beam0) {
    constexpr int KS = A / 16;
    using Map = StageMap<A, kThreads>;
    constexpr int TILE_BYTES = Map::TILE_BYTES;
    constexpr int STAGE_BYTES = ST * TILE_BYTES;
    constexpr int DW_STAGE = ST * Map::DW_TILE;
    constexpr int DW_THREAD = (DW_STAGE + kThreads - 1) / kThreads;
    constexpr bool ALL_VALID = DW_STAGE % kThreads == 0;
    __shared__ __attribute__((aligned(16))) unsigned char lds[2 * STAGE_BYTES];

    const int S = da.S;
    const int B = da.B;
    const int F = ot.F;
    const int f = ot.f;
    const int tid = threadIdx.x;
    const int lane = tid & (kWaveSize - 1);
    const int wave = tid / kWaveSize;
    int blocks = (B - wg_beam0 + NT * 8 - 1) / (NT * 8);  // blocks holding beams
    blocks = blocks < 1 ? 1 : (blocks > kWaves ? kWaves : blocks);
    const int per_block = tpw == 1 ? kWaves / blocks : 1;  // waves sharing a block
    const int beam0 = wg_beam0 + (wave / per_block) * NT * 8;
    const int tile_first = wave % per_block;
    const bool has_beams = beam0 < B;
    const int q_end = (w_end - w_begin) * tpw;  // tiles of this call

    // Staging constants of this thread (StageMap).
    const int tl0 = Map::tile0(wave);
    const int rem0 = Map::rem0(tid);
    const int rho = rem0 & 15;
    const int lds_mine = Map::lds_offset(tl0, rem0);
    const size_t row_stride = rs.row_stride;
    const long mask = rs.ring_mask;
    const uint8_t* const x_thread = rs.x_f + 4 * (rem0 >> 4);
    const uint8_t* const x_row = x_thread + (size_t)rho * row_stride;  // + tile base

    // Tile q = (wl, jt): window w_begin + wl of the call, tile jt of it.
    auto tile_ref = [&](int q, int wl, int jt) -> TileRef {
        TileRef r;
        r.t = rs.t_base + (long)(w_begin + wl) * S + 16 * jt;
        long rows = S - 16 * jt;
        rows = rows < 16 ? rows : 16;
        const long room = rs.t_limit - r.t;
        rows = rows < room ? rows : room;
        r.rows = q < q_end && rows > 0 ? (int)rows : 0;
        return r;
    };
    auto next_tile = [&](int& wl, int& jt) {
        if (++jt == tpw) {
            jt = 0;
            ++wl;
        }
    };
    // Stage loads: the cursor (lw, lj) names the stage's first tile q0;
    // dword kk of this thread belongs to tile q0 + tl0 + tld(kk).
    int lw = 0, lj = 0;
    auto ref_at = [&](int q0, int i) -> TileRef {
        int wl = lw, jt = lj;
        for (int m = 0; m < i; ++m) next_tile(wl, jt);
        return tile_ref(q0 + i, wl, jt);
    };
    auto dword_valid = [&](int kk) -> bool { return ALL_VALID || tl0 + Map::tld(kk) < ST; };
    // Loads the stage at q0 into dst and moves the cursor on by ST tiles.
    // Common case: the stage lies inside one window, all its tiles are
    // full and it does not wrap the ring -- tile i then starts 16 i rows
    // after the first, and no row needs a guard.
    auto load_stage = [&](int q0, unsigned (&dst)[DW_THREAD]) {
        const long t0 = rs.t_base + (long)(w_begin + lw) * S + 16 * lj;
        const bool fast = !MASKED && lj + ST <= tpw && 16 * (lj + ST) <= S && q0 + ST <= q_end &&
                          t0 + 16 * ST <= rs.t_limit &&
                          (mask < 0 || (t0 & mask) <= mask + 1 - 16 * ST);
        if (fast) {
            const uint8_t* const x0 = x_row + (size_t)(t0 & mask) * row_stride;
#pragma unroll
            for (int kk = 0; kk < DW_THREAD; ++kk) {
                const size_t off = (size_t)(16 * (tl0 + Map::tld(kk))) * row_stride +
                                   Map::byte_delta(kk);
                dst[kk] = dword_valid(kk) ? *reinterpret_cast<const unsigned*>(x0 + off) : 0u;
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < DW_THREAD; ++kk) {
                const TileRef r = ref_at(q0, tl0 + Map::tld(kk));
                unsigned v = 0u;
                if (dword_valid(kk) && rho < r.rows) {
                    const size_t row = (size_t)((r.t + rho) & mask);
                    if (!MASKED || rs.valid_f[row * rs.mask_stride] != 0)
                        v = *reinterpret_cast<const unsigned*>(x_thread + row * row_stride +
                                                               Map::byte_delta(kk));
                }
                dst[kk] = v;
            }
        }
        // move the cursor on by ST tiles
        lj += ST;
        if (lj >= tpw) {
            lw += lj / tpw;
            lj %= tpw;
        }
    };
    auto store_stage = [&](const unsigned (&src)[DW_THREAD], unsigned char* buf) {
#pragma unroll
        for (int kk = 0; kk < DW_THREAD; ++kk) {
            if (!dword_valid(kk)) continue;
            unsigned h2[4];
            bf_int4x4_to_half2x4(src[kk], h2);
            bf_uint4v u = {h2[0], h2[1], h2[2], h2[3]};
            *reinterpret_cast<bf_uint4v*>(buf + lds_mine + Map::lds_delta(kk)) = u;
        }
    };

    bf_half8 bhi[NT][KS];
    bf_half8 blo[NT][KS];
    WeightAnchor<A, NT> anc;
    anc.window = -1;
    float win[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) win[nt] = 0.0f;
    const int col = lane & 15;
    const int G = lane >> 4;
    const bool direct = S >= kDirectWeightMinSpectra;

    unsigned pre_a[DW_THREAD];  // stage s + 1
    unsigned pre_b[DW_THREAD];  // stage s + 2
    if (q_end <= 0) return;
    load_stage(0, pre_a);
    store_stage(pre_a, lds);
    load_stage(ST, pre_a);
    __syncthreads();
    int cw = 0, cj = 0;  // compute cursor: tile q0 + tl
    int buf = 0;
    for (int q0 = 0; q0 < q_end; q0 += ST) {
        const unsigned char* cur = lds + buf * STAGE_BYTES;
        load_stage(q0 + 2 * ST, pre_b);
        // The stage in segments of consecutive tiles of one window: the
        // weights at a window's first tile, a check-free tile loop, the
        // window's sum at its last tile.  When waves share a block
        // (one-tile windows), a segment is one tile, taken by one wave.
        int tl = 0;
        while (tl < ST && q0 + tl < q_end) {
            const int w = w_begin + cw;
            const int jt0 = cj;
            int seg = tpw - jt0;
            seg = seg < ST - tl ? seg : ST - tl;
            seg = seg < q_end - q0 - tl ? seg : q_end - q0 - tl;
            if (has_beams && tl % per_block == tile_first) {
                if (jt0 == 0)
                    window_operands<A, NT, SPLIT, DIRSRC, ANCHORED>(da, s_pos, k, w, beam0, lane,
                                                                    direct, anc, bhi, blo);
#pragma unroll 1
                for (int i = 0; i < seg; ++i) {
                    const unsigned char* tile = cur + (tl + i) * TILE_BYTES;
                    bf_half8 av[KS];
#pragma unroll
                    for (int s = 0; s < KS; ++s)
                        av[s] = *reinterpret_cast<const bf_half8*>(tile +
                                                                   (s * kWaveSize + lane) * 16);
                    bf_float4 acc[NT];
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) acc[nt] = bf_float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int s = 0; s < KS; ++s) {
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt)
                            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                                av[s], bhi[nt][s], acc[nt], 0, 0, 0);
                        if (SPLIT) {
#pragma unroll
                            for (int nt = 0; nt < NT; ++nt)
                                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                                    av[s], blo[nt][s], acc[nt], 0, 0, 0);
                        }
                    }
                    if (!INTEGRATE) {
                        const int jt = jt0 + i;
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt) {
                            const int b = beam0 + nt * 8 + (col >> 1);
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const float sq = acc[nt][r] * acc[nt][r];
                                const float p = sq + dpp_swap_pairs(sq);
                                const int in_win = 16 * jt + 4 * G + r;
                                const long t = (long)w * S + in_win;
                                if ((col & 1) == 0 && b < B && in_win < S &&
                                    rs.t_base + t < rs.t_limit)
                                    ot.out[((size_t)t * F + f) * B + b] = p;
                            }
                        }
                    } else {
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt) {
                            float a = win[nt];
#pragma unroll
                            for (int r = 0; r < 4; ++r) a = fmaf(acc[nt][r], acc[nt][r], a);
                            win[nt] = a;
                        }
                    }
                }
                if (INTEGRATE && jt0 + seg == tpw) {
                    const long slot = ot.out_windows > 0 ? (long)w % ot.out_windows : (long)w;
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) {
                        float v = win[nt];
                        v += __shfl_xor(v, 16);
                        v += __shfl_xor(v, 32);
                        v += dpp_swap_pairs(v);
                        const int b = beam0 + nt * 8 + (col >> 1);
                        if (G == 0 && (col & 1) == 0 && b < B)
                            ot.out[((size_t)slot * F + f) * B + b] = v;
                        win[nt] = 0.0f;
                    }
                }
            }
            tl += seg;
            cj += seg;
            if (cj == tpw) {
                cj = 0;
                ++cw;
            }
        }
        // the cursor stays on the stage grid when the call ends mid-stage
        for (; tl < ST; ++tl) next_tile(cw, cj);
        // Stage s + 1 into the other buffer (last read before the previous
        // barrier); stage s + 2's loads move up.
        store_stage(pre_a, lds + (buf ^ 1) * STAGE_BYTES);
#pragma unroll
        for (int kk = 0; kk < DW_THREAD; ++kk) pre_a[kk] = pre_b[kk];
        __syncthreads();
        buf ^= 1;
    }
}

template <int A, int NT, bool SPLIT, bool INTEGRATE, int DIRSRC, int ST, bool ANCHORED>
__global__ __launch_bounds__(kThreads) void tracker_mfma_kernel(
    const uint8_t* __restrict__ x, Geom geo, DirArgs da, float* __restrict__ out, int T, int F,
    int tpw, int n_windows, int windows_per_chunk, int n_items, int per_xcd) {
    __shared__ float s_pos[3 * A];
    const int item = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
    if (item >= n_items) return;
    stage_positions<A>(geo, s_pos);  // visible after tracker_windows' first barrier
    // Items run XCD by XCD in this order: channel pairs (f, f+1), which share
    // the 128-byte lines of every packed row, alternate within a pair's time
    // chunks so both halves of a line are used while it sits in that L2.
    const int n_chunks = n_items / F;
    const int paired = n_chunks * (F & ~1);
    int f, chunk;
    if (item < paired) {
        const int pair = item / (2 * n_chunks);
        const int r = item - pair * 2 * n_chunks;
        f = 2 * pair + (r & 1);
        chunk = r >> 1;
    } else {
        f = F - 1;
        chunk = item - paired;
    }
    const int wg_beam0 = blockIdx.y * kWaves * NT * 8;
    const double k = wave_number(channel_frequency(geo, f));
    const int w_begin = chunk * windows_per_chunk;
    int w_end = w_begin + windows_per_chunk;
    w_end = w_end < n_windows ? w_end : n_windows;
    const RowSource rs{x + (size_t)f * A, (size_t)F * A, -1L, 0L, (long)T};
    const OutTarget ot{out, 0L, F, f};
    tracker_windows<A, NT, SPLIT, INTEGRATE, DIRSRC, ST, ANCHORED>(rs, s_pos, da, ot, k, w_begin,
                                                                   w_end, tpw, wg_beam0);
}

// ------------------------------------------------------------ FP32 tracker
// One workgroup per (window, channel, block of 16 beams): the block's
// A x 16 weights are generated into LDS (FP64 phase, one sincos each), the
// 256 threads stride over the window's spectra doing 4 FMAs per complex MAC,
// and integrated output is folded with an LDS tree reduction.
template <int A, int BB, bool INTEGRATE, int DIRSRC>
__global__ __launch_bounds__(256) void tracker_fp32_kernel(
    const uint8_t* __restrict__ x, Geom geo, DirArgs da, float* __restrict__ out, int T, int F) {
    __shared__ float2 ws[A * BB];
    __shared__ float red[INTEGRATE ? 256 * BB : 1];
    const int S = da.S;
    const int B = da.B;
    const int w = blockIdx.x;
    const int f = blockIdx.y;
    const int b0 = blockIdx.z * BB;
    const double k = wave_number(channel_frequency(geo, f));
    for (int i = threadIdx.x; i < A * BB; i += 256) {
        const int a = i / BB;
        const int j = i % BB;
        float2 v = make_float2(0.0f, 0.0f);
        if (b0 + j < B) {
            float d[3];
            beam_direction<DIRSRC>(da, w, b0 + j, d);
            float p[3];
            antenna_position(geo, a, p);
            v = steering_weight(p, d, k);
        }
        ws[i] = v;
    }
    __syncthreads();
    float acc[BB];
#pragma unroll
    for (int j = 0; j < BB; ++j) acc[j] = 0.0f;
    for (int r = threadIdx.x; r < S; r += 256) {
        const long t = (long)w * S + r;
        if (t >= T) break;
        const uint8_t* row = x + ((size_t)t * F + f) * A;
        float sr[BB], si[BB];
#pragma unroll
        for (int j = 0; j < BB; ++j) sr[j] = si[j] = 0.0f;
        for (int a0 = 0; a0 < A; a0 += 16) {
            const uint4 v = *reinterpret_cast<const uint4*>(row + a0);
            const unsigned words[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const unsigned wd = words[q / 4];
                const unsigned sh = (q % 4) * 8;
                const float xr = (float)bf_nibble(wd, sh);
                const float xi = (float)bf_nibble(wd, sh + 4);
                const float2* wa = ws + (a0 + q) * BB;
#pragma unroll
                for (int j = 0; j < BB; ++j) {
                    const float2 c = wa[j];
                    sr[j] = fmaf(c.x, xr, sr[j]);
                    sr[j] = fmaf(-c.y, xi, sr[j]);
                    si[j] = fmaf(c.x, xi, si[j]);
                    si[j] = fmaf(c.y, xr, si[j]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < BB; ++j) {
            const float p = fmaf(sr[j], sr[j], si[j] * si[j]);
            if (INTEGRATE) {
                acc[j] += p;
            } else if (b0 + j < B) {
                out[((size_t)t * F + f) * B + b0 + j] = p;
            }
        }
    }
    if (INTEGRATE) {
#pragma unroll
        for (int j = 0; j < BB; ++j) red[j * 256 + threadIdx.x] = acc[j];
        __syncthreads();
        for (int off = 128; off >= 1; off >>= 1) {
            if (threadIdx.x < off) {
#pragma unroll
                for (int j = 0; j < BB; ++j)
                    red[j * 256 + threadIdx.x] += red[j * 256 + threadIdx.x + off];
            }
            __syncthreads();
        }
        if (threadIdx.x < BB && b0 + threadIdx.x < B)
            out[((size_t)w * F + f) * B + b0 + threadIdx.x] = red[threadIdx.x * 256];
    }
}

// ------------------------------------------------------------ serving loop
// Plan: one thread per request slot reads the step's head, finds the
// windows from the slot's cursor whose last spectrum is at or before the
// head (a slot admitted with a start behind the previous head gets all its
// windows held by the ring at once) and the work they make; block scans
// turn the counts into offsets.  Each thread advances only its own slot's
// cursor; the step's head is recorded as the processed head.
__global__ __launch_bounds__(kServeMaxSlots) void serve_plan_kernel(
    ServeSlot* __restrict__ slots, ServeCtrl* __restrict__ ctrl,
    ServeWork* __restrict__ work, const long long* heads, int F) {
    __shared__ int scan_l[kServeMaxSlots];
    __shared__ int scan_s[kServeMaxSlots];
    const int s = threadIdx.x;
    const long long step = ctrl->step;
    // system-scope load: the host wrote this entry just before the launch
    const long long H = __hip_atomic_load(heads + (step & (kServeHeadRing - 1)),
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    ServeSlot& sl = slots[s];
    int n_win = 0;
    int w_lo = 0;
    int n_long = 0;
    int n_short = 0;
    if (sl.active) {
        const long long S = sl.S;
        const long long t0 = sl.t0;
        const long long te = sl.t_end;
        const long long W = (te - t0 + S - 1) / S;
        // windows from the slot's cursor up to the last one ending at or
        // before H (the last window may be partial and end at te)
        const long long lo = sl.w_next;
        const long long h = H >= te ? W : (H < t0 ? 0 : (H - t0) / S);
        const long long hi = h > lo ? h : lo;
        sl.w_next = hi;
        if (hi > lo) {
            n_win = (int)(hi - lo);
            w_lo = (int)lo;
            if (S <= 16)
                n_short = F * ((sl.B + 15) / 16) * ((n_win + kServeShortRun - 1) / kServeShortRun);
            else
                n_long = F * ((n_win + sl.wpi - 1) / sl.wpi);
        }
    }
    scan_l[s] = n_long;
    scan_s[s] = n_short;
    __syncthreads();
    for (int d = 1; d < kServeMaxSlots; d <<= 1) {
        const int vl = s >= d ? scan_l[s - d] : 0;
        const int vs = s >= d ? scan_s[s - d] : 0;
        __syncthreads();
        scan_l[s] += vl;
        scan_s[s] += vs;
        __syncthreads();
    }
    work->off_long[s] = scan_l[s] - n_long;
    work->off_short[s] = scan_s[s] - n_short;
    work->w_lo[s] = w_lo;
    work->n_win[s] = n_win;
    if (s == kServeMaxSlots - 1) {
        work->off_long[kServeMaxSlots] = scan_l[s];
        work->off_short[kServeMaxSlots] = scan_s[s];
        ctrl->processed = H;
        ctrl->step = step + 1;
    }
}

// Last slot s with off[s] <= i (off non-decreasing, off[0] = 0 <= i < off[n]).
__device__ __forceinline__ int owner_slot(const int* off, int i) {
    int lo = 0;
    int hi = kServeMaxSlots;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// One wave task of a one-tile-window slot: windows [w_begin, w_end) of
// channel f for the 16 beams from beam0.  The wave works alone: each lane
// loads its A-operand dwords (row lane % 16, antenna group lane / 16)
// straight from the ring, one window ahead, and the arithmetic -- anchored
// weights, MFMA order, window sum and cross-lane fold -- is that of
// tracker_windows for one-tile windows, so results are bit-identical to the
// multi-beam tracker.
template <int A, bool SPLIT, bool MASKED>
__device__ __forceinline__ void serve_short_task(const uint8_t* __restrict__ ring,
                                                 const uint8_t* __restrict__ valid,
                                                 long long ring_mask, int F,
                                                 const float* s_pos, const ServeSlot* sl,
                                                 int f, int beam0, int w_begin, int w_end,
                                                 double k) {
    constexpr int KS = A / 16;
    constexpr int NT = 2;
    const int lane = threadIdx.x & (kWaveSize - 1);
    const int col = lane & 15;
    const int G = lane >> 4;
    const int S = sl->S;
    const int B = sl->B;
    const long t0 = (long)sl->t0;
    const long te = (long)sl->t_end;
    const long out_windows = (long)sl->out_windows;
    float* out = sl->out;
    const DirArgs da{nullptr, &sl->traj[0][0], 0L, S, B};
    const size_t row_stride = (size_t)F * A;
    const uint8_t* base = ring + (size_t)f * A + (size_t)G * (A / 4);
    const int row = col;  // A operand: lane holds row lane % 16, group lane / 16

    WeightAnchor<A, NT> anc;
    anc.window = -1;
    // A-operand dwords of window w for this lane (zero past the window)
    unsigned cur[KS], nxt[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) cur[s] = nxt[s] = 0u;
    // frame (t, f) of the ring is lost when its validity byte is 0
    auto frame_ok = [&](long t) {
        return !MASKED || valid[(size_t)(t & ring_mask) * F + f] != 0;
    };
    {
        const long t = t0 + (long)w_begin * S + row;
        if (row < S && t < te && frame_ok(t)) {
            const uint8_t* p = base + (size_t)(t & ring_mask) * row_stride;
#pragma unroll
            for (int s = 0; s < KS; ++s) cur[s] = *reinterpret_cast<const unsigned*>(p + 4 * s);
        }
    }
#pragma unroll 1
    for (int w = w_begin; w < w_end; ++w) {
        {
            const long t = t0 + (long)(w + 1) * S + row;
            const bool ok = w + 1 < w_end && row < S && t < te && frame_ok(t);
            const uint8_t* p = base + (size_t)(t & ring_mask) * row_stride;
#pragma unroll
            for (int s = 0; s < KS; ++s)
                nxt[s] = ok ? *reinterpret_cast<const unsigned*>(p + 4 * s) : 0u;
        }
        bf_half8 bhi[NT][KS];
        bf_half8 blo[NT][KS];
        window_operands<A, NT, SPLIT, kDirTrajectory, true>(da, s_pos, k, w, beam0, lane,
                                                            false, anc, bhi, blo);
        bf_half8 av[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            unsigned h2[4];
            bf_int4x4_to_half2x4(cur[s], h2);
            av[s] = __builtin_bit_cast(bf_half8, bf_uint4v{h2[0], h2[1], h2[2], h2[3]});
        }
        bf_float4 acc[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[nt] = bf_float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av[s], bhi[nt][s], acc[nt], 0, 0,
                                                                 0);
            if (SPLIT) {
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av[s], blo[nt][s], acc[nt],
                                                                     0, 0, 0);
            }
        }
        const long slot = out_windows > 0 ? (long)w % out_windows : (long)w;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            float v = 0.0f;
#pragma unroll
            for (int r = 0; r < 4; ++r) v = fmaf(acc[nt][r], acc[nt][r], v);
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            v += dpp_swap_pairs(v);
            const int b = beam0 + nt * 8 + (col >> 1);
            if (G == 0 && (col & 1) == 0 && b < B) out[((size_t)slot * F + f) * B + b] = v;
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) cur[s] = nxt[s];
    }
}

// Items: persistent workgroups first share the workgroup items of
// multi-tile-window slots (four waves cover up to 64 beams), then each wave
// takes wave tasks of one-tile-window slots on its own.
// Two workgroups per CU (the persistent grid) need at most 256 registers.
template <int A, bool SPLIT, bool MASKED>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2)))
void serve_items_kernel(
    const uint8_t* __restrict__ ring, const uint8_t* __restrict__ valid, long long ring_mask,
    int F, Geom geo, const ServeSlot* __restrict__ slots, const ServeWork* __restrict__ work) {
    __shared__ int off_l[kServeMaxSlots + 1];
    __shared__ int off_s[kServeMaxSlots + 1];
    __shared__ float s_pos[3 * A];
    for (int i = threadIdx.x; i <= kServeMaxSlots; i += kThreads) {
        off_l[i] = work->off_long[i];
        off_s[i] = work->off_short[i];
    }
    stage_positions<A>(geo, s_pos);
    __syncthreads();
    const int n_long = off_l[kServeMaxSlots];
    const int n_short = off_s[kServeMaxSlots];
    // XCD-aware order: blockIdx.x % 8 names the XCD; consecutive items
    // (neighbouring channels) go to one XCD.
    const int lin = gridDim.x % kXcds == 0
                        ? (int)(blockIdx.x % kXcds) * (int)(gridDim.x / kXcds) + (int)(blockIdx.x / kXcds)
                        : (int)blockIdx.x;
    for (int item = lin; item < n_long; item += gridDim.x) {
        const int sidx = owner_slot(off_l, item);
        const ServeSlot* sl = slots + sidx;
        const int local = item - off_l[sidx];
        const int run = local / F;
        const int f = local - run * F;
        const int S = sl->S;
        const int w_first = work->w_lo[sidx];
        const int w_last = w_first + work->n_win[sidx];
        const int w_begin = w_first + run * sl->wpi;
        const int w_end = w_begin + sl->wpi < w_last ? w_begin + sl->wpi : w_last;
        const DirArgs da{nullptr, &sl->traj[0][0], 0L, S, sl->B};
        const RowSource rs{ring + (size_t)f * A, (size_t)F * A, (long)ring_mask, (long)sl->t0,
                           (long)sl->t_end, MASKED ? valid + f : nullptr, (size_t)F};
        const OutTarget ot{sl->out, (long)sl->out_windows, F, f};
        const double k = wave_number(channel_frequency(geo, f));
        tracker_windows<A, 2, SPLIT, true, kDirTrajectory, 4, true, MASKED>(
            rs, s_pos, da, ot, k, w_begin, w_end, (S + 15) / 16, 0);
    }
    const int wave = threadIdx.x / kWaveSize;
    for (int task = blockIdx.x * kWaves + wave; task < n_short; task += gridDim.x * kWaves) {
        const int sidx = owner_slot(off_s, task);
        const ServeSlot* sl = slots + sidx;
        const int local = task - off_s[sidx];
        const int blocks = (sl->B + 15) / 16;
        const int f = local % F;
        const int rest = local / F;
        const int block = rest % blocks;
        const int run = rest / blocks;
        const int w_first = work->w_lo[sidx];
        const int w_last = w_first + work->n_win[sidx];
        const int w_begin = w_first + run * kServeShortRun;
        const int w_end =
            w_begin + kServeShortRun < w_last ? w_begin + kServeShortRun : w_last;
        const double k = wave_number(channel_frequency(geo, f));
        serve_short_task<A, SPLIT, MASKED>(ring, valid, ring_mask, F, s_pos, sl, f, block * 16,
                                           w_begin, w_end,
                                   k);
    }
}

template <int A, int NT, bool SPLIT, bool INTEGRATE, int DIRSRC, bool ANCHORED>
void launch_mfma_nt(hipStream_t stream, const TrackerLaunch& L, const DirArgs& da) {
    const int tpw = (int)ceil_div((size_t)L.S, 16);
    constexpr int ST = A == 256 ? 2 : 4;  // two stage buffers of ST tiles in LDS
    // Chunk: ~1280 spectra of whole windows per workgroup.
    int wpc = (int)ceil_div(80, (size_t)tpw);
    if (wpc < 1) wpc = 1;
    const int n_windows = (int)L.n_windows;
    const int n_chunks = (int)ceil_div((size_t)n_windows, (size_t)wpc);
    const int n_items = n_chunks * (int)L.F;
    const int per_xcd = (int)ceil_div((size_t)n_items, kXcds);
    const int beam_blocks = (int)ceil_div(L.B, (size_t)(kWaves * NT * 8));
    hipLaunchKernelGGL((tracker_mfma_kernel<A, NT, SPLIT, INTEGRATE, DIRSRC, ST, ANCHORED>),
                       dim3((unsigned)(per_xcd * kXcds), (unsigned)beam_blocks), dim3(kThreads),
                       0, stream, L.x, L.geo, da, L.out, (int)L.T, (int)L.F, tpw,
                       n_windows, wpc, n_items, per_xcd);
}

// Beams per wave: as many 8-beam N-tiles as the beam count needs, up to 4
// (2 at A=128, 1 at A=256) with directly evaluated weights; short windows keep
// anchor phasors in registers and take at most 2.
template <int A, bool SPLIT, bool INTEGRATE, int DIRSRC>
void launch_mfma_a(hipStream_t stream, const TrackerLaunch& L, const DirArgs& da) {
    constexpr int kMaxNT = A <= 64 ? 4 : (A == 128 ? 2 : 1);
    constexpr int kMaxNT2 = kMaxNT >= 2 ? 2 : 1;
    const size_t need = ceil_div(L.B, (size_t)(kWaves * 8));
    if (L.S < (size_t)kDirectWeightMinSpectra) {
        if (need <= 1 || kMaxNT2 == 1)
            launch_mfma_nt<A, 1, SPLIT, INTEGRATE, DIRSRC, true>(stream, L, da);
        else
            launch_mfma_nt<A, kMaxNT2, SPLIT, INTEGRATE, DIRSRC, true>(stream, L, da);
        return;
    }
    if (need <= 1 || kMaxNT == 1)
        launch_mfma_nt<A, 1, SPLIT, INTEGRATE, DIRSRC, false>(stream, L, da);
    else if (need <= 2 || kMaxNT == 2)
        launch_mfma_nt<A, kMaxNT2, SPLIT, INTEGRATE, DIRSRC, false>(stream, L, da);
    else
        launch_mfma_nt<A, kMaxNT, SPLIT, INTEGRATE, DIRSRC, false>(stream, L, da);
}

template <bool SPLIT, bool INTEGRATE, int DIRSRC>
void launch_mfma(hipStream_t stream, const TrackerLaunch& L, const DirArgs& da) {
    switch (L.A) {
    case 32: launch_mfma_a<32, SPLIT, INTEGRATE, DIRSRC>(stream, L, da); break;
    case 64: launch_mfma_a<64, SPLIT, INTEGRATE, DIRSRC>(stream, L, da); break;
    case 128: launch_mfma_a<128, SPLIT, INTEGRATE, DIRSRC>(stream, L, da); break;
    case 256: launch_mfma_a<256, SPLIT, INTEGRATE, DIRSRC>(stream, L, da); break;
    default: invalid("n_ant must be 32, 64, 128, or 256");
    }
}

template <bool INTEGRATE, int DIRSRC>
void launch_fp32(hipStream_t stream, const TrackerLaunch& L, const DirArgs& da) {
    const dim3 grid((unsigned)L.n_windows, (unsigned)L.F, (unsigned)ceil_div(L.B, 16));
#define BF_FP32(A_)                                                                        \
    hipLaunchKernelGGL((tracker_fp32_kernel<A_, 16, INTEGRATE, DIRSRC>), grid, dim3(256), 0, \
                       stream, L.x, L.geo, da, L.out, (int)L.T, (int)L.F)
    switch (L.A) {
    case 32: BF_FP32(32); break;
    case 64: BF_FP32(64); break;
    case 128: BF_FP32(128); break;
    case 256: BF_FP32(256); break;
    default: invalid("n_ant must be 32, 64, 128, or 256");
    }
#undef BF_FP32
}

// Single-beam per-spectrum tracker as an HBM stream (FP32 VALU).  Every
// lane holds 32 antennas of one channel, LPC = A / 32 lanes per channel,
// so a wave covers 64 / LPC adjacent channels and reads, per spectrum, one
// contiguous 2 KB run of packed rows (two 16-byte loads per lane, with the
// next kB1Depth - 1 spectra in flight while one is summed).  A unit is (window,
// channel group, time chunk); its weights (reduced-phase FP64 -> FP32
// sincos, within about one float ulp of (float)cos / (float)sin of the FP64
// phase) stay in 64 registers for the whole chunk.  The complex sum runs
// over the lane's antennas in order in two interleaved accumulator pairs,
// then across the LPC lanes of the channel by xor shuffles.
constexpr int kB1WavesPerBlock = 4;
constexpr int kB1Depth = 4;  // spectra in flight per lane

template <int A, int DIRSRC>
__global__ __launch_bounds__(kB1WavesPerBlock * 64) void beam1_stream_kernel(
    const uint8_t* __restrict__ x, float* __restrict__ out, int T, int F, int S, int n_windows,
    int groups, int chunks, Geom geo, DirArgs da) {
#pragma clang fp contract(off)
    constexpr int LPC = A / 32;
    constexpr int CPW = 64 / LPC;  // channels per wave
    const int lane = threadIdx.x & 63;
    const long unit = (long)blockIdx.x * kB1WavesPerBlock + (threadIdx.x >> 6);
    if (unit >= (long)n_windows * groups * chunks) return;  // whole waves
    const int c = (int)(unit % chunks);
    const long r = unit / chunks;
    const int g = (int)(r % groups);
    const int w = (int)(r / groups);
    const int f = g * CPW + lane / LPC;
    const int part = lane % LPC;  // antennas 32 part .. 32 part + 31
    const bool live = f < F;
    const int fl = live ? f : F - 1;  // loads stay in bounds; nothing is stored
    const int tw0 = w * S;
    const int tw1 = min(tw0 + S, T);
    const int per = (tw1 - tw0 + chunks - 1) / chunks;
    const int t0 = min(tw0 + c * per, tw1);
    const int t1 = min(t0 + per, tw1);
    float d[3];
    beam_direction<DIRSRC>(da, w, 0, d);
    const double k = wave_number(channel_frequency(geo, fl));
    float wr[32], wi[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        float p[3];
        antenna_position(geo, 32 * part + j, p);
        const float2 v = steering_weight_reduced(p, d, k);
        wr[j] = v.x;
        wi[j] = v.y;
    }
    const uint8_t* base = x + (size_t)fl * A + 32 * part;
    const size_t row = (size_t)F * A;
    // a ring of kB1Depth spectra in flight: slot u holds spectrum tb + u and
    // is refilled with spectrum tb + u + kB1Depth once it has been summed
    uint4 ring[kB1Depth][2];
#pragma unroll
    for (int u = 0; u < kB1Depth; ++u)
        if (t0 + u < t1) {
            const uint4* q = reinterpret_cast<const uint4*>(base + (size_t)(t0 + u) * row);
            ring[u][0] = q[0];
            ring[u][1] = q[1];
        }
    for (int tb = t0; tb < t1; tb += kB1Depth) {
#pragma unroll
        for (int u = 0; u < kB1Depth; ++u) {
            const int t = tb + u;
            if (t >= t1) break;  // uniform
            float re[2] = {0.0f, 0.0f}, im[2] = {0.0f, 0.0f};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const unsigned wd[4] = {ring[u][h].x, ring[u][h].y, ring[u][h].z, ring[u][h].w};
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int j = 16 * h + q;
                    const int byte = (int)((wd[q >> 2] >> (8 * (q & 3))) & 0xFFu);
                    const float xr = (float)((byte << 28) >> 28);
                    const float xi = (float)((byte << 24) >> 28);
                    const int e = q & 1;
                    re[e] = fmaf(wr[j], xr, re[e]);
                    re[e] = fmaf(-wi[j], xi, re[e]);
                    im[e] = fmaf(wr[j], xi, im[e]);
                    im[e] = fmaf(wi[j], xr, im[e]);
                }
            }
            if (t + kB1Depth < t1) {
                const uint4* q = reinterpret_cast<const uint4*>(base + (size_t)(t + kB1Depth) * row);
                ring[u][0] = q[0];
                ring[u][1] = q[1];
            }
            float sr = re[0] + re[1], si = im[0] + im[1];
#pragma unroll
            for (int o = 1; o < LPC; o <<= 1) {
                sr += __shfl_xor(sr, o);
                si += __shfl_xor(si, o);
            }
            if (live && part == 0) out[(size_t)t * F + f] = sr * sr + si * si;
        }
    }
}

// Units (window, channel group, time chunk) for about four waves per SIMD,
// chunks of at least 8 spectra.
template <int DIRSRC>
void launch_beam1_stream(hipStream_t stream, const TrackerLaunch& L, const DirArgs& da) {
    const size_t cpw = 64 / (L.A / 32);
    const size_t groups = ceil_div(L.F, cpw);
    const size_t target = 4096;
    size_t chunks = std::max<size_t>(1, ceil_div(target, L.n_windows * groups));
    chunks = std::min(chunks, std::max<size_t>(1, L.S / 8));
    const size_t n_units = L.n_windows * groups * chunks;
    const size_t blocks = ceil_div(n_units, kB1WavesPerBlock);
    if (blocks > (size_t)UINT32_MAX) fail(BF_ERR_OVERFLOW, "tracker grid exceeds the supported size");
#define BF_B1(A_)                                                                              \
    hipLaunchKernelGGL((beam1_stream_kernel<A_, DIRSRC>), dim3((unsigned)blocks),               \
                       dim3(kB1WavesPerBlock * 64), 0, stream, L.x, L.out, (int)L.T, (int)L.F, \
                       (int)L.S, (int)L.n_windows, (int)groups, (int)chunks, L.geo, da)
    switch (L.A) {
    case 32: BF_B1(32); break;
    case 64: BF_B1(64); break;
    case 128: BF_B1(128); break;
    case 256: BF_B1(256); break;
    default: invalid("n_ant must be 32, 64, 128, or 256");
    }
#undef BF_B1
}

template <int DIRSRC>
void dispatch(hipStream_t stream, const TrackerLaunch& L, const DirArgs& da) {
    const bool integ = L.integrate;
    // One beam, per-spectrum output: the FP32 stream kernel (AUTO and FP32).
    if (L.B == 1 && !integ && (L.kernel == BF_KERNEL_AUTO || L.kernel == BF_KERNEL_FP32)) {
        launch_beam1_stream<DIRSRC>(stream, L, da);
        return;
    }
    switch (L.kernel) {
    case BF_KERNEL_SPLIT_F16:
    case BF_KERNEL_AUTO:
        integ ? launch_mfma<true, true, DIRSRC>(stream, L, da)
              : launch_mfma<true, false, DIRSRC>(stream, L, da);
        break;
    case BF_KERNEL_F16:
        integ ? launch_mfma<false, true, DIRSRC>(stream, L, da)
              : launch_mfma<false, false, DIRSRC>(stream, L, da);
        break;
    default:  // FP32
        integ ? launch_fp32<true, DIRSRC>(stream, L, da) : launch_fp32<false, DIRSRC>(stream, L, da);
        break;
    }
}

}  // namespace

void launch_tracker(hipStream_t stream, const TrackerLaunch& L) {
    if (L.T == 0 || L.F == 0 || L.B == 0 || L.S == 0 || L.n_windows == 0) return;
    if (L.T > (size_t)INT32_MAX / 2 || L.S > (size_t)INT32_MAX / 2 || L.F > 65535 ||
        L.n_windows > (size_t)INT32_MAX / 2)
        fail(BF_ERR_OVERFLOW, "tracker grid exceeds the supported size");
    DirArgs da{L.dir_table, L.trajectories, (long)L.window_offset, (int)L.S, (int)L.B};
    if (L.dir_source == kDirTable)
        dispatch<kDirTable>(stream, L, da);
    else
        dispatch<kDirTrajectory>(stream, L, da);
    hip_check(hipGetLastError(), "tracker kernel launch");
}

void launch_serve_step(hipStream_t stream, const ServeLaunch& L) {
    hipLaunchKernelGGL(serve_plan_kernel, dim3(1), dim3(kServeMaxSlots), 0, stream, L.slots,
                       L.ctrl, L.work, L.heads, (int)L.F);
#define BF_SERVE(A_, SPLIT_)                                                                  \
    do {                                                                                      \
        if (L.valid != nullptr)                                                               \
            hipLaunchKernelGGL((serve_items_kernel<A_, SPLIT_, true>), dim3((unsigned)L.grid), \
                               dim3(kThreads), 0, stream, L.ring, L.valid, L.ring_mask,       \
                               (int)L.F, L.geo, L.slots, L.work);                             \
        else                                                                                  \
            hipLaunchKernelGGL((serve_items_kernel<A_, SPLIT_, false>),                       \
                               dim3((unsigned)L.grid), dim3(kThreads), 0, stream, L.ring,     \
                               L.valid, L.ring_mask, (int)L.F, L.geo, L.slots, L.work);       \
    } while (0)
    if (L.A == 32 && L.split)
        BF_SERVE(32, true);
    else if (L.A == 32)
        BF_SERVE(32, false);
    else if (L.A == 64 && L.split)
        BF_SERVE(64, true);
    else if (L.A == 64)
        BF_SERVE(64, false);
    else
        invalid("the serving loop supports 32 or 64 antennas");
#undef BF_SERVE
    hip_check(hipGetLastError(), "serve step launch");
}

}  // namespace bfhip
