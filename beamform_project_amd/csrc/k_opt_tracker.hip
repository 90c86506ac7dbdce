// k_opt_tracker.hip -- adaptive planar direction search for gfx950: the
// searching path of the reference's CpuOptBeamTracker::run_into
// (src/cpu_opt_beam_tracker.cpp:1619-2457), one integration window at a time.
//
// Per window (all stages stream-ordered, nothing returns to the host):
//   covariance      launch_doa_covariance (k_doa.hip): exact int32 sums of
//                   x x^H over the window's snapshots, forward smoothing over
//                   L = A - M + 1 sub-arrays (:559-636).
//   opt_fold        per channel, FP64: R_block = Rs / (K L), forward-backward
//                   average when smoothing (:637-656, without the in-place
//                   read of already-averaged entries), diagonal load for
//                   Capon (:244-253, before the fold, :1803-1804), forgetting
//                   fold R = lambda R + (1 - lambda) R_block (:1880-1892),
//                   trace.
//   opt_factor      Capon only, one wave per channel: Cholesky R = L L^H
//                   (:398-421) and its inverse W = L^-1, so that
//                   a^H R^-1 a = |W a|^2 (:423-437).
//   opt_scan        per (channel, cell): the steering vector
//                   a_j = exp(-j k p_j . d) (:96-106) and the Bartlett power
//                   tr(R) + 2 Re sum_{r<c} conj(a_r) R_rc a_c (:287-306) or
//                   the Capon power 1 / |W a|^2.
//   opt_decide      one workgroup: powers summed over channels, argmax, the
//                   next level's centre, quadratic peak interpolation
//                   (:830-853), and at the last level the window's direction.
// Level 0 scans the G x G grid centred on the previous estimate
// (build_centred_grid, :798-828); each refinement level scans the 3 x 3
// neighbourhood of the running centre at half the previous pitch
// (:2179-2388).
//
// Steering directions follow the reference's two cases.  For a coplanar
// array (every |z| <= 1e-6) the reference multiplies a centre phasor by a
// tabulated relative phasor, i.e. it steers at the nominal lattice point
// centre + offset in FP64 (:1971-2013, :1216-1296); a lattice point outside
// l^2 + m^2 <= 0.999 and every point of a non-coplanar array steers at the
// clamped float direction instead.  Here the FP64 phase of that direction is
// evaluated directly, so the steering vector equals the reference's up to
// the rounding of its phasor product.
#include <algorithm>

#include "bf_doa.hpp"
#include "gauss_jordan.hpp"

namespace bfhip {
namespace {

constexpr int kThreads = 256;

struct cd {
    double re, im;
};

// ------------------------------------------------------- search geometry
// Coarse cell (u, v) of the grid centred on `c` (build_centred_grid,
// src/cpu_opt_beam_tracker.cpp:798-828): its clamped float direction, and
// the direction the scan steers at.
__device__ void coarse_cell(const OptGridArgs& g, const float c[3], int u, int v, float dir[3],
                            double st[3]) {
    const int G = g.G;
    const float step_l = G == 1 ? 0.0f : (2.0f * g.fov_l) / (float)(G - 1);
    const float step_m = G == 1 ? 0.0f : (2.0f * g.fov_m) / (float)(G - 1);
    float l = (c[0] - g.fov_l) + (float)u * step_l;
    float m = (c[1] - g.fov_m) + (float)v * step_m;
    const float r2 = l * l + m * m;
    if (r2 > 0.999f) {
        const float s = sqrtf(0.999f / r2);
        l *= s;
        m *= s;
    }
    dir[0] = l;
    dir[1] = m;
    dir[2] = sqrtf(1.0f - (l * l + m * m));
    st[0] = (double)dir[0];
    st[1] = (double)dir[1];
    st[2] = (double)dir[2];
    if (g.coplanar) {
        // nominal lattice point (:1971-1993)
        const double sl = G == 1 ? 0.0 : (2.0 * (double)g.fov_l) / (double)(G - 1);
        const double sm = G == 1 ? 0.0 : (2.0 * (double)g.fov_m) / (double)(G - 1);
        const double mid = (double)(G - 1) * 0.5;
        const double nl = (double)c[0] + ((double)u - mid) * sl;
        const double nm = (double)c[1] + ((double)v - mid) * sm;
        if (!(nl * nl + nm * nm > 0.999)) {
            st[0] = nl;
            st[1] = nm;
            st[2] = 0.0;
        }
    }
}

// Candidate i = (dv + 1) * 3 + (du + 1) of refinement level `level` >= 1
// around centre `c` (:2179-2203, :2219-2233).
__device__ void refine_cell(const OptGridArgs& g, const float c[3], int level, int i,
                            float dir[3], double st[3]) {
    const int du = i % 3 - 1, dv = i / 3 - 1;
    float hl = g.fov_l / (float)g.G, hm = g.fov_m / (float)g.G;
    for (int k = 0; k < level; ++k) {
        hl *= 0.5f;
        hm *= 0.5f;
    }
    float l = c[0] + (float)du * hl;
    float m = c[1] + (float)dv * hm;
    const float r2 = l * l + m * m;
    const bool clamped = r2 > 0.999f;
    if (clamped) {
        const float s = sqrtf(0.999f / r2);
        l *= s;
        m *= s;
    }
    dir[0] = l;
    dir[1] = m;
    dir[2] = sqrtf(1.0f - (l * l + m * m));
    if (g.coplanar && !clamped) {
        st[0] = (double)c[0] + (double)du * (double)hl;
        st[1] = (double)c[1] + (double)dv * (double)hm;
        st[2] = 0.0;
    } else {
        st[0] = (double)dir[0];
        st[1] = (double)dir[1];
        st[2] = (double)dir[2];
    }
}

__device__ void cell_direction(const OptGridArgs& g, const OptSearch& s, int level, int cell,
                               float dir[3], double st[3]) {
    if (level == 0)
        coarse_cell(g, s.prev, cell % g.G, cell / g.G, dir, st);
    else
        refine_cell(g, s.centre, level, cell, dir, st);
}

// quadratic_interp (:830-853): offsets applied only where the 3-point
// curvature is positive, clamped to half a cell.
__device__ void quadratic_interp(float pc, float p_up, float p_down, float p_vp, float p_vm,
                                 float* du, float* dv) {
    *du = 0.0f;
    *dv = 0.0f;
    const float den_u = p_down - 2.0f * pc + p_up;
    if (den_u > 0.0f) {
        float d = (0.5f * (p_up - p_down)) / den_u;
        d = d > 0.5f ? 0.5f : d;
        d = d < -0.5f ? -0.5f : d;
        *du = d;
    }
    const float den_v = p_vm - 2.0f * pc + p_vp;
    if (den_v > 0.0f) {
        float d = (0.5f * (p_vp - p_vm)) / den_v;
        d = d > 0.5f ? 0.5f : d;
        d = d < -0.5f ? -0.5f : d;
        *dv = d;
    }
}

// direction_from_lm of (l, m) after the 0.999 horizon clamp (:2350-2357).
__device__ void clamp_direction(float l, float m, float out[3]) {
    const float r2 = l * l + m * m;
    if (r2 > 0.999f) {
        const float s = sqrtf(0.999f / r2);
        l *= s;
        m *= s;
    }
    out[0] = l;
    out[1] = m;
    out[2] = sqrtf(1.0f - (l * l + m * m));
}

// ------------------------------------------------------------------ fold
// One workgroup per channel, one covariance entry per thread and pass:
// R_block = Rs / (K L), forward-backward averaged when fb (the mirrored
// entry is read from Rs directly), diagonal load for Capon, and the fold
// into state.  Bartlett also writes its scan operand and the trace of the
// folded R; Capon's operand comes from opt_factor_kernel.
__global__ __launch_bounds__(kThreads) void opt_fold_kernel(
    const int* __restrict__ rs, double* __restrict__ state, int first, int M, int K, int L,
    int fb, int capon, float load, float lambda, float2* __restrict__ scanmat,
    float* __restrict__ trace_out, int* __restrict__ mode_out) {
    __shared__ double s_diag[64];
    __shared__ double s_eps;
    const int f = blockIdx.x;
    const int tid = threadIdx.x;
    const int MM = M * M;
    const double scale = 1.0 / ((double)K * (double)L);
    const int* rf = rs + (size_t)f * MM * 2;
    auto block = [&](int r, int c) -> cd {
        const int e = r * M + c;
        cd v{(double)rf[2 * e] * scale, (double)rf[2 * e + 1] * scale};
        if (fb) {
            // (R + J conj(R) J) / 2
            const int m = (M - 1 - r) * M + (M - 1 - c);
            v.re = 0.5 * (v.re + (double)rf[2 * m] * scale);
            v.im = 0.5 * (v.im - (double)rf[2 * m + 1] * scale);
        }
        return v;
    };
    if (tid < M) s_diag[tid] = block(tid, tid).re;
    __syncthreads();
    if (tid == 0) {
        double tr = 0.0;
        for (int i = 0; i < M; ++i) tr += s_diag[i];
        s_eps = (capon && load > 0.0f) ? (double)load * tr / (double)M : 0.0;
    }
    __syncthreads();
    const double eps = s_eps;
    const double lam = (double)lambda;
    const double comp = (double)(1.0f - lambda);
    const bool copy = first || lambda >= 1.0f;
    double* st = state + (size_t)f * MM * 2;
    float2* Q = scanmat + (size_t)f * MM;
    for (int e = tid; e < MM; e += kThreads) {
        const int r = e / M, c = e - (e / M) * M;
        cd v = block(r, c);
        if (r == c) v.re += eps;
        if (!copy) {
            v.re = lam * st[2 * e] + comp * v.re;
            v.im = lam * st[2 * e + 1] + comp * v.im;
        }
        st[2 * e] = v.re;
        st[2 * e + 1] = v.im;
        if (r == c) s_diag[r] = v.re;
        // Bartlett operand: the strict upper triangle, zero elsewhere.
        if (!capon)
            Q[e] = r < c ? make_float2((float)v.re, (float)v.im) : make_float2(0.0f, 0.0f);
    }
    if (capon) return;
    __syncthreads();
    if (tid == 0) {
        double tr = 0.0;
        for (int i = 0; i < M; ++i) tr += s_diag[i];
        trace_out[f] = (float)tr;
        mode_out[f] = 0;
    }
}

// ---------------------------------------------------------------- factor
constexpr int kGjBlock = 4;                                     // 4 x 4 blocks
constexpr int kFactorThreads = (64 / kGjBlock) * (64 / kGjBlock);  // 256
// Capon operand: Q = R^-1 of the folded state by in-place Gauss-Jordan
// elimination in FP64, one workgroup per channel.  Thread t holds the 4 x 4
// block (rows 4 (t / 16) .., columns 4 (t % 16) ..) of a 64 x 64 matrix in
// registers (identity beyond M); per pivot k the owners publish row k and
// column k to LDS (double-buffered: one barrier per pivot) and every thread
// updates its block.  Without pivoting on a Hermitian positive-definite R
// the k-th pivot is the Cholesky pivot d_k = L_kk^2, so the reference's
// factorisation test (d_k > 1e-12, :398-421) is applied to it; a failed
// channel falls back to the Bartlett form on the folded R (:354-375).
// Q is written in full (FP64) with its trace; the scan evaluates
// 1 / (tr Q + 2 Re sum_{r<c} conj(a_r) Q_rc a_c) = 1 / (a^H R^-1 a) in
// FP64, where the large entries of R^-1 cancel.
__global__ __launch_bounds__(kFactorThreads) void opt_factor_kernel(
    const double* __restrict__ state, int M, double2* __restrict__ scan64,
    double* __restrict__ trace64, float2* __restrict__ scanmat, float* __restrict__ trace_out,
    int* __restrict__ mode_out) {
    __shared__ GjShared<cd> sh;
    __shared__ double s_diag[64];
    const int f = blockIdx.x;
    const int tid = threadIdx.x;
    const int i0 = gj_row0<kGjBlock>(tid), j0 = gj_col0<kGjBlock>(tid);
    const int MM = M * M;
    const double* R = state + (size_t)f * MM * 2;
    cd a[kGjBlock][kGjBlock];
    gj_load<kGjBlock>(a, R, M, M, tid);
    const bool ok = gj_invert<kGjBlock>(a, M, 1e-12, sh, tid);
    if (!ok) {
        float2* Q = scanmat + (size_t)f * MM;
        for (int e = tid; e < MM; e += kFactorThreads) {
            const int r = e / M, c = e - (e / M) * M;
            Q[e] = r < c ? make_float2((float)R[2 * e], (float)R[2 * e + 1])
                         : make_float2(0.0f, 0.0f);
        }
        if (tid == 0) {
            double tr = 0.0;
            for (int i = 0; i < M; ++i) tr += R[2 * (i * M + i)];
            trace_out[f] = (float)tr;
            mode_out[f] = 0;
        }
        return;
    }
    double2* Q = scan64 + (size_t)f * MM;
#pragma unroll
    for (int u = 0; u < kGjBlock; ++u)
#pragma unroll
        for (int v = 0; v < kGjBlock; ++v) {
            const int i = i0 + u, j = j0 + v;
            if (i < M && j < M) {
                Q[i * M + j] = make_double2(a[u][v].re, a[u][v].im);
                if (i == j) s_diag[i] = a[u][v].re;
            }
        }
    __syncthreads();
    if (tid == 0) {
        double tr = 0.0;
        for (int i = 0; i < M; ++i) tr += s_diag[i];
        trace64[f] = tr;
        mode_out[f] = 1;
    }
}

// ------------------------------------------------------------------ scan
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One pass of a wave over NC cells: acc_lane = sum_r conj(a_r) R[r][lane]
// over the strict upper triangle of R, a_r broadcast from this wave's LDS
// steering vectors.  R is a zero-padded 64 x 64 tile and a_r = 0 for
// r >= M, so rows are taken four at a time (the four LDS reads issue
// together) up to M rounded up.
template <int NC>
__device__ __forceinline__ void scan_pass(const float2* __restrict__ Q,
                                          const float2* __restrict__ sa, int M, int lane,
                                          float2 (&acc)[NC]) {
    for (int r0 = 0; r0 < M; r0 += 4) {
        float2 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = Q[(r0 + u) * 64 + lane];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = r0 + u;
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                const float2 av = sa[j * 64 + r];  // broadcast
                const float ar = av.x, ai = av.y;
                acc[j].x = fmaf(ai, q[u].y, fmaf(ar, q[u].x, acc[j].x));
                acc[j].y = fmaf(-ai, q[u].x, fmaf(ar, q[u].y, acc[j].y));
            }
        }
    }
}

// Bartlett form: P = tr(R) + 2 Re sum_c y_c a_c (also Capon's fallback).
template <int NC>
__device__ void scan_cells(const float2* __restrict__ Q, float2* __restrict__ sa,
                           const OptSearch& s,
                           const OptGridArgs& g, int level, int c_begin, int c_end, int lane,
                           int wave, int f, int F, double kw, float tr, float px, float py,
                           float pz, float* __restrict__ part) {
    const bool live = lane < g.M;
    for (int c0 = c_begin + wave * NC; c0 < c_end; c0 += 4 * NC) {
        float2 a[NC], acc[NC];
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            a[j] = make_float2(0.0f, 0.0f);
            acc[j] = make_float2(0.0f, 0.0f);
            const int cell = min(c0 + j, c_end - 1);
            float dir[3];
            double st[3];
            cell_direction(g, s, level, cell, dir, st);
            if (live) {
                const double ph = kw * ((double)px * st[0] + (double)py * st[1] + (double)pz * st[2]);
                double sn, cs;
                sincos(ph, &sn, &cs);
                a[j] = make_float2((float)cs, -(float)sn);
            }
            sa[j * 64 + lane] = a[j];
        }
        // the wave's own LDS writes precede its reads in order
        __builtin_amdgcn_wave_barrier();
        scan_pass<NC>(Q, sa, g.M, lane, acc);
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            float t = acc[j].x * a[j].x - acc[j].y * a[j].y;  // Re(y_c a_c)
            t = wave_sum(t);
            float p = tr + 2.0f * t;
            if (!isfinite(p)) p = 0.0f;
            if (lane == 0 && c0 + j < c_end) part[(size_t)(c0 + j) * F + f] = p;
        }
    }
}

// Capon in FP64: P = 1 / (tr Q + 2 Re sum_{r<c} conj(a_r) Q_rc a_c) with
// Q = R^-1 (strict upper triangle in Q64, zero-padded to 64 x 64).
template <int NC>
__device__ void scan_cells_capon64(const double2* __restrict__ Q64, double2* __restrict__ sa,
                                   const OptSearch& s,
                                   const OptGridArgs& g, int level, int c_begin, int c_end,
                                   int lane, int wave, int f, int F, double kw, double tr,
                                   float px, float py, float pz, float* __restrict__ part) {
    const bool live = lane < g.M;
    for (int c0 = c_begin + wave * NC; c0 < c_end; c0 += 4 * NC) {
        double2 a[NC], acc[NC];
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            a[j] = make_double2(0.0, 0.0);
            acc[j] = make_double2(0.0, 0.0);
            const int cell = min(c0 + j, c_end - 1);
            float dir[3];
            double st[3];
            cell_direction(g, s, level, cell, dir, st);
            if (live) {
                const double ph = kw * ((double)px * st[0] + (double)py * st[1] + (double)pz * st[2]);
                double sn, cs;
                sincos(ph, &sn, &cs);
                a[j] = make_double2(cs, -sn);
            }
            sa[j * 64 + lane] = a[j];
        }
        __builtin_amdgcn_wave_barrier();
        for (int r0 = 0; r0 < g.M; r0 += 4) {
            double2 q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = Q64[(r0 + u) * 64 + lane];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int r = r0 + u;
#pragma unroll
                for (int j = 0; j < NC; ++j) {
                    const double2 av = sa[j * 64 + r];  // broadcast
                    const double ar = av.x, ai = av.y;
                    acc[j].x = fma(ai, q[u].y, fma(ar, q[u].x, acc[j].x));
                    acc[j].y = fma(-ai, q[u].x, fma(ar, q[u].y, acc[j].y));
                }
            }
        }
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            double t = acc[j].x * a[j].x - acc[j].y * a[j].y;
            t = wave_sum_f64(t);
            const double den = tr + 2.0 * t;
            float p = den > 1.0e-12 ? (float)(1.0 / den) : 0.0f;
            if (!isfinite(p)) p = 0.0f;
            if (lane == 0 && c0 + j < c_end) part[(size_t)(c0 + j) * F + f] = p;
        }
    }
}

// grid (F, chunks); part[cell][f] = the channel's power at the cell.
template <int NC, int NC64>
__global__ __launch_bounds__(kThreads) void opt_scan_kernel(
    const float2* __restrict__ scanmat, const float* __restrict__ trace,
    const double2* __restrict__ scan64, const double* __restrict__ trace64,
    const int* __restrict__ mode, const float* __restrict__ pos, const float* __restrict__ freqs,
    const OptSearch* __restrict__ search, OptGridArgs g, int level, int n_cells,
    int cells_per_chunk, int F, int capon_run, float* __restrict__ part) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    float2* Q = reinterpret_cast<float2*>(smem);
    double2* Q64 = reinterpret_cast<double2*>(smem);
    const int f = blockIdx.x;
    const bool capon64 = mode[f] == 1;
    // per-wave steering vectors [NC][64] after the operand tile
    unsigned char* steer = smem + 64 * 64 * (capon_run ? sizeof(double2) : sizeof(float2));
    const int M = g.M;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int c_begin = (int)blockIdx.y * cells_per_chunk;
    const int c_end = min(n_cells, c_begin + cells_per_chunk);
    if (capon64) {
        // strict upper triangle of R^-1, zero elsewhere
        const double2* src = scan64 + (size_t)f * M * M;
        for (int e = tid; e < 64 * 64; e += kThreads) {
            const int r = e >> 6, c = e & 63;
            Q64[e] = (r < c && c < M) ? src[r * M + c] : make_double2(0.0, 0.0);
        }
    } else {
        // [64][64], zero outside [M][M] (M may be odd: element by element)
        const float2* src = scanmat + (size_t)f * M * M;
        for (int e = tid; e < 64 * 64; e += kThreads) {
            const int r = e >> 6, c = e & 63;
            Q[e] = (r < M && c < M) ? src[r * M + c] : make_float2(0.0f, 0.0f);
        }
    }
    __syncthreads();
    const OptSearch s = *search;
    const float tr = trace[f];
    const double kw = kTwoPi * (double)freqs[f] / kSpeedOfLight;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    if (lane < M) {
        px = pos[3 * lane];
        py = pos[3 * lane + 1];
        pz = pos[3 * lane + 2];
    }
    if (capon64)
        scan_cells_capon64<NC64>(Q64, reinterpret_cast<double2*>(steer) + wave * NC64 * 64, s,
                                 g, level, c_begin, c_end, lane, wave, f, F, kw, trace64[f],
                           px, py, pz, part);
    else
        scan_cells<NC>(Q, reinterpret_cast<float2*>(steer) + wave * NC * 64, s, g,
                             level, c_begin, c_end, lane, wave, f, F, kw, tr, px, py, pz,
                         part);
}

// ---------------------------------------------------------------- decide
constexpr int kDecideThreads = 1024;

// One workgroup: each wave sums a cell's channel powers (FP64, a fixed
// lane-strided order and shuffle tree), then thread 0 takes the level's
// serial decision.
__global__ __launch_bounds__(kDecideThreads) void opt_decide_kernel(
    const float* __restrict__ part, int F, int n_cells, OptGridArgs g, int level,
    int last_level, int window, OptSearch* __restrict__ search, float* __restrict__ dirs) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    float* P = reinterpret_cast<float*>(smem);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    // four cells per wave at a time and kSteps channel steps of 64 per
    // round: every load of a round is issued (at a clamped, in-bounds
    // address) before the sums, which add in the order f = lane, lane + 64,
    // ... as one loop over f would
    constexpr int kWaves = kDecideThreads / 64;
    constexpr int kSteps = 6;
    for (int c0 = wave; c0 < n_cells; c0 += 4 * kWaves) {
        double sum[4] = {0.0, 0.0, 0.0, 0.0};
        for (int f0 = lane; f0 < F; f0 += 64 * kSteps) {
            float v[kSteps][4];
#pragma unroll
            for (int k = 0; k < kSteps; ++k)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int f = min(f0 + 64 * k, F - 1);
                    const int cell = min(c0 + j * kWaves, n_cells - 1);
                    v[k][j] = part[(size_t)cell * F + f];
                }
#pragma unroll
            for (int k = 0; k < kSteps; ++k)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (f0 + 64 * k < F && c0 + j * kWaves < n_cells) sum[j] += (double)v[k][j];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double v = wave_sum_f64(sum[j]);
            const int cell = c0 + j * kWaves;
            if (lane == 0 && cell < n_cells) P[cell] = (float)v;
        }
    }
    __syncthreads();
    if (wave != 0) return;
    // argmax with the reference's rule (strict >, so the first maximum):
    // each lane scans its cells in increasing order, then the lanes combine
    // preferring the larger value and, on a tie, the lower index
    float bv = -__builtin_huge_valf();
    int bi = n_cells;  // none yet
    for (int i = lane; i < n_cells; i += 64)
        if (bi == n_cells || P[i] > bv) {
            bv = P[i];
            bi = i;
        }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const float ov = __shfl_xor(bv, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (oi < n_cells && (bi == n_cells || ov > bv || (ov == bv && oi < bi))) {
            bv = ov;
            bi = oi;
        }
    }
    if (lane != 0) return;
    OptSearch s = *search;
    const int best = bi;
    float dir[3];
    double st[3];
    cell_direction(g, s, level, best, dir, st);
    float c[3] = {dir[0], dir[1], dir[2]};
    if (g.interp) {
        float du = 0.0f, dv = 0.0f;
        if (level == 0 && last_level == 0) {
            // coarse-grid interpolation when there is no refinement (:2362-2396)
            const int G = g.G, bu = best % G, bv = best / G;
            auto at = [&](int u, int v) -> float {
                const int uu = bu + u, vv = bv + v;
                return (uu < 0 || uu >= G || vv < 0 || vv >= G) ? 0.0f : P[vv * G + uu];
            };
            quadratic_interp(P[best], at(1, 0), at(-1, 0), at(0, 1), at(0, -1), &du, &dv);
            const float step_l = G == 1 ? 0.0f : (2.0f * g.fov_l) / (float)(G - 1);
            const float step_m = G == 1 ? 0.0f : (2.0f * g.fov_m) / (float)(G - 1);
            clamp_direction(c[0] + du * step_l, c[1] + dv * step_m, c);
        } else if (level > 0) {
            quadratic_interp(P[4], P[5], P[3], P[7], P[1], &du, &dv);
            float hl = g.fov_l / (float)g.G, hm = g.fov_m / (float)g.G;
            for (int k = 0; k < level; ++k) {
                hl *= 0.5f;
                hm *= 0.5f;
            }
            clamp_direction(c[0] + du * hl, c[1] + dv * hm, c);
        }
    }
    s.centre[0] = c[0];
    s.centre[1] = c[1];
    s.centre[2] = c[2];
    if (level == last_level) {
        s.prev[0] = c[0];
        s.prev[1] = c[1];
        s.prev[2] = c[2];
        dirs[3 * window] = c[0];
        dirs[3 * window + 1] = c[1];
        dirs[3 * window + 2] = c[2];
    }
    *search = s;
}

}  // namespace

void launch_opt_fold(hipStream_t stream, const int* rs, double* state, bool first, size_t F,
                     size_t M, size_t K, size_t L, bool fb, bool capon, float load,
                     float lambda, float2* scanmat, float* trace, double2* scan64,
                     double* trace64, int* mode) {
    hipLaunchKernelGGL(opt_fold_kernel, dim3((unsigned)F), dim3(kThreads), 0, stream, rs, state,
                       first ? 1 : 0, (int)M, (int)K, (int)L, fb ? 1 : 0, capon ? 1 : 0, load,
                       lambda, scanmat, trace, mode);
    hip_check(hipGetLastError(), "opt_fold_kernel launch");
    if (!capon) return;
    hipLaunchKernelGGL(opt_factor_kernel, dim3((unsigned)F), dim3(kFactorThreads), 0, stream, state,
                       (int)M, scan64, trace64, scanmat, trace, mode);
    hip_check(hipGetLastError(), "opt_factor_kernel launch");
}

void launch_opt_scan(hipStream_t stream, const float2* scanmat, const float* trace,
                     const double2* scan64, const double* trace64, const int* mode,
                     const float* pos, const float* freqs, const OptSearch* search,
                     const OptGridArgs& g, int level, size_t F, size_t n_cells, bool capon,
                     float* part) {
    // One pass per wave: cells are spread over as many workgroups as it
    // takes (each stages its channel's operand once), which keeps the
    // per-wave chain short; the 3 x 3 refinement uses three waves of 3.
    const bool coarse = level == 0;
    const size_t per_wave = coarse ? (capon ? 4 : 8) : 3;
    const size_t cpc = 4 * per_wave;
    const size_t chunks = ceil_div(n_cells, cpc);
    // operand tile, then per-wave steering vectors (sized for either path)
    const size_t nc = coarse ? 8 : 3, nc64 = coarse ? 4 : 3;  // the template arguments
    const size_t steer_per_wave =
        64 * std::max(nc * sizeof(float2), capon ? nc64 * sizeof(double2) : 0);
    const size_t smem = 64 * 64 * (capon ? sizeof(double2) : sizeof(float2)) + 4 * steer_per_wave;
    const void* fn = coarse ? reinterpret_cast<const void*>(opt_scan_kernel<8, 4>)
                            : reinterpret_cast<const void*>(opt_scan_kernel<3, 3>);
    if (smem > 64 * 1024)
        hip_check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem),
                  "hipFuncSetAttribute");
#define BF_OPT_SCAN(NC_, NC64_)                                                                \
    hipLaunchKernelGGL((opt_scan_kernel<NC_, NC64_>), dim3((unsigned)F, (unsigned)chunks),      \
                       dim3(kThreads), smem, stream, scanmat, trace, scan64, trace64, mode, pos,  \
                       freqs, search, g, level, (int)n_cells, (int)cpc, (int)F, capon ? 1 : 0,  \
                       part)
    if (coarse)
        BF_OPT_SCAN(8, 4);
    else
        BF_OPT_SCAN(3, 3);
#undef BF_OPT_SCAN
    hip_check(hipGetLastError(), "opt_scan_kernel launch");
}

void launch_opt_decide(hipStream_t stream, const float* part, size_t F, size_t n_cells,
                       const OptGridArgs& g, int level, int last_level, size_t window,
                       OptSearch* search, float* dirs) {
    hipLaunchKernelGGL(opt_decide_kernel, dim3(1), dim3(kDecideThreads), n_cells * sizeof(float),
                       stream,
                       part, (int)F, (int)n_cells, g, level, last_level, (int)window, search, dirs);
    hip_check(hipGetLastError(), "opt_decide_kernel launch");
}

}  // namespace bfhip
