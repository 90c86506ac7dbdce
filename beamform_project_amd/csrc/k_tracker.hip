// k_tracker.hip -- dynamic beam tracker kernels for gfx950.
//
// The tracker re-points every beam once per integration window: window w
// uses the direction of its first spectrum (reference
// src/beam_tracker.cpp:166-190) and the steering weights
// w[b][f][a] = (float)cos(phi), (float)sin(phi) with the phase
// phi = (2 pi f / c) * (p_a . d) evaluated in FP64 (src/weights.cpp:37-55),
// then contracts like the fixed grid (src/cpu_beamformer.cpp:56-78).
// Weights are generated inside the kernels, per window, never stored:
// cos/sin are evaluated in FP64 as the reference does, except that the
// MFMA kernels reduce the FP64 phase to [-pi, pi] and evaluate it in FP32
// for windows shorter than 64 spectra (about one float ulp, below the
// FP16 split's rounding), where weight generation would dominate.
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
#include "bf_kernels.hpp"
#include "bf_serve.hpp"
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

// Channel frequency: a device table, or src/geometry.cpp:79-95.
__device__ __forceinline__ float channel_frequency(const Geom& geo, int f) {
    if (geo.freqs != nullptr) return geo.freqs[f];
    const float step = (float)f * geo.f_step;
    return geo.f_start + step;
}

// cos/sin of phi = k * (p . d) for short windows: the phase is formed and
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

// Windows of at least this many spectra take the weights from an FP64
// sincos, as the reference does; shorter windows, where weight generation
// would dominate, take the reduced FP32 evaluation above.
constexpr int kExactWeightMinSpectra = 64;

// B operand of one wave for one window.  Lanes c and c^1 hold the Re and
// Im columns of the same beam over the same antennas, so each computes
// half of the weights and takes the other half from its partner by DPP.
template <int A, int NT, bool SPLIT, int DIRSRC>
__device__ __forceinline__ void window_weights(const DirArgs& da, const Geom& geo,
                                               double k, int w, int beam0, int lane,
                                               bool exact, bf_half8 (&bhi)[NT][A / 16],
                                               bf_half8 (&blo)[NT][A / 16]) {
    constexpr int KS = A / 16;
    const int c = lane & 15;
    const int g = lane >> 4;
    const int comp_im = c & 1;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int b = beam0 + nt * 8 + (c >> 1);
        const bool live = b < da.B;
        float d[3] = {0.0f, 0.0f, 1.0f};
        if (live) beam_direction<DIRSRC>(da, w, b, d);
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            float2 own[2];
            float2 other[2];
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                const int a = g * (A / 4) + 4 * s + 2 * comp_im + jj;
                own[jj] = make_float2(0.0f, 0.0f);
                if (live) {
                    float p[3];
                    antenna_position(geo, a, p);
                    own[jj] = exact ? steering_weight(p, d, k)
                                    : steering_weight_reduced(p, d, k);
                }
            }
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                other[jj].x = dpp_swap_pairs(own[jj].x);
                other[jj].y = dpp_swap_pairs(own[jj].y);
            }
            bf_half8 h, l;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float2 v = (j >> 1) == comp_im ? own[j & 1] : other[j & 1];
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

// Time rows of one channel.  Spectrum t (absolute) lives at
// x_f + (t & ring_mask) * row_stride: a linear [T][F][A] array uses
// ring_mask = -1, a ring of 2^k spectra uses 2^k - 1.  Window w covers
// spectra [t_base + w*S, t_base + (w+1)*S) clipped to t_limit.
struct RowSource {
    const uint8_t* x_f;  // x + f * A
    size_t row_stride;   // F * A bytes
    long ring_mask;
    long t_base;
    long t_limit;
};

// Output of one channel: integrated [slot][F][B] with slot = w mod
// out_windows (0: slot = w), or per-spectrum [t - t_base][F][B].
struct OutTarget {
    float* out;
    long out_windows;
    int F;
    int f;
};

// Windows [w_begin, w_end) of one channel, for the NT*8 beams of this wave
// starting at beam0.  Entered by the whole workgroup (kThreads threads):
// stages of ST window-aligned tiles of 16 spectra go through LDS in the
// A-operand layout, the next stage is held in registers meanwhile, and the
// B operand (this window's weights) is regenerated at each window's first
// tile.
template <int A, int NT, bool SPLIT, bool INTEGRATE, int DIRSRC, int ST>
__device__ __forceinline__ void tracker_windows(const RowSource& rs, const Geom& geo,
                                                const DirArgs& da, const OutTarget& ot, double k,
                                                int w_begin, int w_end, int tpw, int beam0) {
    constexpr int KS = A / 16;
    constexpr int TILE_BYTES = KS * kWaveSize * 16;
    constexpr int DW_ROW = A / 4;
    constexpr int DW_TILE = 16 * DW_ROW;
    constexpr int DW_STAGE = ST * DW_TILE;
    constexpr int DW_THREAD = (DW_STAGE + kThreads - 1) / kThreads;
    __shared__ __attribute__((aligned(16))) unsigned char lds[ST * TILE_BYTES];

    const int S = da.S;
    const int B = da.B;
    const int F = ot.F;
    const int f = ot.f;
    const int tid = threadIdx.x;
    const int lane = tid & (kWaveSize - 1);
    const bool has_beams = beam0 < B;
    const int q_end = (w_end - w_begin) * tpw;  // tiles of this call

    unsigned pre[DW_THREAD];
    auto load_stage = [&](int q0) {
#pragma unroll
        for (int kk = 0; kk < DW_THREAD; ++kk) {
            const int q = tid + kk * kThreads;
            unsigned v = 0u;
            if (q < DW_STAGE) {
                const int tl = q / DW_TILE;
                const int rem = q - tl * DW_TILE;
                const int rho = rem / DW_ROW;
                const int j = rem - rho * DW_ROW;
                const int qt = q0 + tl;
                if (qt < q_end) {
                    const int w = w_begin + qt / tpw;
                    const int in_win = 16 * (qt % tpw) + rho;
                    const long t = rs.t_base + (long)w * S + in_win;
                    if (in_win < S && t < rs.t_limit)
                        v = *reinterpret_cast<const unsigned*>(
                            rs.x_f + (size_t)(t & rs.ring_mask) * rs.row_stride + 4 * j);
                }
            }
            pre[kk] = v;
        }
    };
    auto store_stage = [&]() {
#pragma unroll
        for (int kk = 0; kk < DW_THREAD; ++kk) {
            const int q = tid + kk * kThreads;
            if (q < DW_STAGE) {
                const int tl = q / DW_TILE;
                const int rem = q - tl * DW_TILE;
                const int rho = rem / DW_ROW;
                const int j = rem - rho * DW_ROW;
                const int g = j / (A / 16);
                const int s = j - g * (A / 16);
                unsigned h2[4];
                bf_int4x4_to_half2x4(pre[kk], h2);
                bf_uint4v u = {h2[0], h2[1], h2[2], h2[3]};
                *reinterpret_cast<bf_uint4v*>(lds + tl * TILE_BYTES +
                                              (s * kWaveSize + rho + 16 * g) * 16) = u;
            }
        }
    };

    bf_half8 bhi[NT][KS];
    bf_half8 blo[NT][KS];
    float win[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) win[nt] = 0.0f;
    const int col = lane & 15;
    const int G = lane >> 4;

    if (q_end > 0) load_stage(0);
    for (int q0 = 0; q0 < q_end; q0 += ST) {
        __syncthreads();
        store_stage();
        __syncthreads();
        if (q0 + ST < q_end) load_stage(q0 + ST);
        if (!has_beams) continue;
#pragma unroll 1
        for (int tl = 0; tl < ST && q0 + tl < q_end; ++tl) {
            const int qt = q0 + tl;
            const int wl = qt / tpw;        // window within this call
            const int jt = qt - wl * tpw;   // tile within the window
            const int w = w_begin + wl;
            if (jt == 0)
                window_weights<A, NT, SPLIT, DIRSRC>(da, geo, k, w, beam0, lane,
                                                     S >= kExactWeightMinSpectra, bhi, blo);
            bf_half8 av[KS];
#pragma unroll
            for (int s = 0; s < KS; ++s)
                av[s] = *reinterpret_cast<const bf_half8*>(lds + tl * TILE_BYTES +
                                                           (s * kWaveSize + lane) * 16);
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
            if (!INTEGRATE) {
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const int b = beam0 + nt * 8 + (col >> 1);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float sq = acc[nt][r] * acc[nt][r];
                        const float p = sq + dpp_swap_pairs(sq);
                        const int in_win = 16 * jt + 4 * G + r;
                        const long t = (long)w * S + in_win;
                        if ((col & 1) == 0 && b < B && in_win < S && rs.t_base + t < rs.t_limit)
                            ot.out[((size_t)t * F + f) * B + b] = p;
                    }
                }
                continue;
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                float a = win[nt];
#pragma unroll
                for (int r = 0; r < 4; ++r) a = fmaf(acc[nt][r], acc[nt][r], a);
                win[nt] = a;
            }
            if (jt == tpw - 1) {
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
    }
}

template <int A, int NT, bool SPLIT, bool INTEGRATE, int DIRSRC, int ST>
__global__ __launch_bounds__(kThreads) void tracker_mfma_kernel(
    const uint8_t* __restrict__ x, Geom geo, DirArgs da, float* __restrict__ out, int T, int F,
    int tpw, int n_windows, int windows_per_chunk, int n_items, int per_xcd) {
    const int wave = threadIdx.x / kWaveSize;
    const int item = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
    if (item >= n_items) return;
    const int n_chunks = n_items / F;
    const int f = item / n_chunks;
    const int chunk = item - f * n_chunks;
    const int beam0 = (blockIdx.y * kWaves + wave) * NT * 8;
    const double k = wave_number(channel_frequency(geo, f));
    const int w_begin = chunk * windows_per_chunk;
    int w_end = w_begin + windows_per_chunk;
    w_end = w_end < n_windows ? w_end : n_windows;
    const RowSource rs{x + (size_t)f * A, (size_t)F * A, -1L, 0L, (long)T};
    const OutTarget ot{out, 0L, F, f};
    tracker_windows<A, NT, SPLIT, INTEGRATE, DIRSRC, ST>(rs, geo, da, ot, k, w_begin, w_end, tpw,
                                                         beam0);
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
// Plan: one thread per request slot finds the windows whose last spectrum
// lies in (processed, head] and the items they make; a block-wide scan
// turns the counts into item offsets.  The previous head is read before
// the barrier and replaced after it, so each step sees its own range.
__global__ __launch_bounds__(kServeMaxSlots) void serve_plan_kernel(
    const ServeSlot* __restrict__ slots, ServeCtrl* __restrict__ ctrl,
    ServeWork* __restrict__ work, int F) {
    __shared__ int scan[kServeMaxSlots];
    const int s = threadIdx.x;
    const long long H = ctrl->head;
    const long long P = ctrl->processed;
    const ServeSlot& sl = slots[s];
    int n_win = 0;
    int w_lo = 0;
    int items = 0;
    if (sl.active) {
        const long long S = sl.S;
        const long long t0 = sl.t0;
        const long long te = sl.t_end;
        const long long W = (te - t0 + S - 1) / S;
        // the last window may be partial and end at te, before t0 + W*S
        const long long lo = P >= te ? W : (P <= t0 ? 0 : (P - t0) / S);
        const long long hi = H >= te ? W : (H < t0 ? 0 : (H - t0) / S);
        if (hi > lo) {
            n_win = (int)(hi - lo);
            w_lo = (int)lo;
            items = F * ((n_win + sl.wpi - 1) / sl.wpi);
        }
    }
    scan[s] = items;
    __syncthreads();
    for (int d = 1; d < kServeMaxSlots; d <<= 1) {
        const int v = s >= d ? scan[s - d] : 0;
        __syncthreads();
        scan[s] += v;
        __syncthreads();
    }
    work->off[s] = scan[s] - items;
    work->w_lo[s] = w_lo;
    work->n_win[s] = n_win;
    if (s == kServeMaxSlots - 1) {
        work->off[kServeMaxSlots] = scan[s];
        ctrl->n_items = scan[s];
        ctrl->processed = H;
    }
}

// Items: persistent workgroups stride over the step's items.  Item i
// belongs to the last slot s with off[s] <= i; within the slot it is
// (run of windows, channel) with channels fastest.  Four waves cover up to
// 64 beams (16 per wave).
template <int A, bool SPLIT>
__global__ __launch_bounds__(kThreads) void serve_items_kernel(
    const uint8_t* __restrict__ ring, long long ring_mask, int F, Geom geo,
    const ServeSlot* __restrict__ slots, const ServeWork* __restrict__ work) {
    __shared__ int off[kServeMaxSlots + 1];
    for (int i = threadIdx.x; i <= kServeMaxSlots; i += kThreads) off[i] = work->off[i];
    __syncthreads();
    const int n_items = off[kServeMaxSlots];
    const int wave = threadIdx.x / kWaveSize;
    for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
        int lo = 0;
        int hi = kServeMaxSlots;  // off[lo] <= item < off[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (off[mid] <= item)
                lo = mid;
            else
                hi = mid;
        }
        const ServeSlot* sl = slots + lo;
        const int local = item - off[lo];
        const int run = local / F;
        const int f = local - run * F;
        const int S = sl->S;
        const int w_first = work->w_lo[lo];
        const int w_last = w_first + work->n_win[lo];
        const int w_begin = w_first + run * sl->wpi;
        const int w_end = w_begin + sl->wpi < w_last ? w_begin + sl->wpi : w_last;
        const DirArgs da{nullptr, &sl->traj[0][0], 0L, S, sl->B};
        const RowSource rs{ring + (size_t)f * A, (size_t)F * A, (long)ring_mask, (long)sl->t0,
                           (long)sl->t_end};
        const OutTarget ot{sl->out, (long)sl->out_windows, F, f};
        const double k = wave_number(channel_frequency(geo, f));
        tracker_windows<A, 2, SPLIT, true, kDirTrajectory, 4>(rs, geo, da, ot, k, w_begin, w_end,
                                                              (S + 15) / 16, wave * 16);
    }
}

__global__ void serve_set_head_kernel(ServeCtrl* ctrl, long long head) {
    if (threadIdx.x == 0) ctrl->head = head;
}

template <int A, int NT, bool SPLIT, bool INTEGRATE, int DIRSRC>
void launch_mfma_nt(hipStream_t stream, const TrackerLaunch& L, const DirArgs& da) {
    const int tpw = (int)ceil_div((size_t)L.S, 16);
    constexpr int ST = 4;
    // Chunk: ~1280 spectra of whole windows per workgroup.
    int wpc = (int)ceil_div(80, (size_t)tpw);
    if (wpc < 1) wpc = 1;
    const int n_windows = (int)L.n_windows;
    const int n_chunks = (int)ceil_div((size_t)n_windows, (size_t)wpc);
    const int n_items = n_chunks * (int)L.F;
    const int per_xcd = (int)ceil_div((size_t)n_items, kXcds);
    const int beam_blocks = (int)ceil_div(L.B, (size_t)(kWaves * NT * 8));
    hipLaunchKernelGGL((tracker_mfma_kernel<A, NT, SPLIT, INTEGRATE, DIRSRC, ST>),
                       dim3((unsigned)(per_xcd * kXcds), (unsigned)beam_blocks), dim3(kThreads),
                       0, stream, L.x, L.geo, da, L.out, (int)L.T, (int)L.F, tpw,
                       n_windows, wpc, n_items, per_xcd);
}

template <int A, bool SPLIT, bool INTEGRATE, int DIRSRC>
void launch_mfma_a(hipStream_t stream, const TrackerLaunch& L, const DirArgs& da) {
    constexpr int kMaxNT = A <= 64 ? 4 : (A == 128 ? 2 : 1);
    const size_t need = ceil_div(L.B, (size_t)(kWaves * 8));
    if (need <= 1 || kMaxNT == 1)
        launch_mfma_nt<A, 1, SPLIT, INTEGRATE, DIRSRC>(stream, L, da);
    else if (need <= 2 || kMaxNT == 2)
        launch_mfma_nt<A, (kMaxNT >= 2 ? 2 : 1), SPLIT, INTEGRATE, DIRSRC>(stream, L, da);
    else
        launch_mfma_nt<A, kMaxNT, SPLIT, INTEGRATE, DIRSRC>(stream, L, da);
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

template <int DIRSRC>
void dispatch(hipStream_t stream, const TrackerLaunch& L, const DirArgs& da) {
    const bool integ = L.integrate;
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

void launch_serve_set_head(hipStream_t stream, ServeCtrl* ctrl, long long head) {
    hipLaunchKernelGGL(serve_set_head_kernel, dim3(1), dim3(64), 0, stream, ctrl, head);
    hip_check(hipGetLastError(), "serve head launch");
}

void launch_serve_step(hipStream_t stream, const ServeLaunch& L) {
    hipLaunchKernelGGL(serve_plan_kernel, dim3(1), dim3(kServeMaxSlots), 0, stream, L.slots,
                       L.ctrl, L.work, (int)L.F);
#define BF_SERVE(A_, SPLIT_)                                                                 \
    hipLaunchKernelGGL((serve_items_kernel<A_, SPLIT_>), dim3((unsigned)L.grid), dim3(kThreads), \
                       0, stream, L.ring, L.ring_mask, (int)L.F, L.geo, L.slots, L.work)
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
