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
//                   trace; Capon: Cholesky R = L L^H (:398-421) and its
//                   inverse W = L^-1, so a^H R^-1 a = |W a|^2 (:423-437).
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

namespace bfhip {
namespace {

constexpr int kThreads = 256;
constexpr int kNc = 8;  // cells per wave pass of the scan

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

// ---------------------------------------------------------- fold / factor
// One workgroup per channel.  LDS: two M x M complex double buffers.
__global__ __launch_bounds__(kThreads) void opt_fold_kernel(
    const int* __restrict__ rs, double* __restrict__ state, int first, int M, int K, int L,
    int fb, int capon, float load, float lambda, float2* __restrict__ scanmat,
    float* __restrict__ trace_out, int* __restrict__ mode_out) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    cd* b0 = reinterpret_cast<cd*>(smem);
    cd* b1 = b0 + M * M;
    __shared__ double s_eps;
    __shared__ int s_ok;
    const int f = blockIdx.x;
    const int tid = threadIdx.x;
    const int MM = M * M;
    const double scale = 1.0 / ((double)K * (double)L);
    for (int e = tid; e < MM; e += kThreads) {
        const int* p = rs + ((size_t)f * MM + e) * 2;
        b0[e] = cd{(double)p[0] * scale, (double)p[1] * scale};
    }
    __syncthreads();
    cd* R = b0;
    if (fb) {
        // R = (R + J conj(R) J) / 2
        for (int e = tid; e < MM; e += kThreads) {
            const int r = e / M, c = e - (e / M) * M;
            const cd a = b0[e], b = b0[(M - 1 - r) * M + (M - 1 - c)];
            b1[e] = cd{0.5 * (a.re + b.re), 0.5 * (a.im - b.im)};
        }
        __syncthreads();
        R = b1;
    }
    if (tid == 0) {
        double tr = 0.0;
        for (int i = 0; i < M; ++i) tr += R[i * M + i].re;
        s_eps = (capon && load > 0.0f) ? (double)load * tr / (double)M : 0.0;
        s_ok = 1;
    }
    __syncthreads();
    const double lam = (double)lambda;
    const double comp = (double)(1.0f - lambda);
    const bool copy = first || lambda >= 1.0f;
    double* st = state + (size_t)f * MM * 2;
    for (int e = tid; e < MM; e += kThreads) {
        cd v = R[e];
        if (e / M == e % M) v.re += s_eps;
        if (!copy) {
            v.re = lam * st[2 * e] + comp * v.re;
            v.im = lam * st[2 * e + 1] + comp * v.im;
        }
        st[2 * e] = v.re;
        st[2 * e + 1] = v.im;
        R[e] = v;
    }
    __syncthreads();
    float2* Q = scanmat + (size_t)f * MM;
    if (!capon) {
        // Bartlett operand: the strict upper triangle, zero elsewhere.
        for (int e = tid; e < MM; e += kThreads) {
            const int r = e / M, c = e - (e / M) * M;
            Q[e] = r < c ? make_float2((float)R[e].re, (float)R[e].im) : make_float2(0.0f, 0.0f);
        }
        if (tid == 0) {
            double tr = 0.0;
            for (int i = 0; i < M; ++i) tr += R[i * M + i].re;
            trace_out[f] = (float)tr;
            mode_out[f] = 0;
        }
        return;
    }
    // Keep the folded R for a failed factorisation's Bartlett fallback.
    cd* A = R == b0 ? b1 : b0;
    for (int e = tid; e < MM; e += kThreads) A[e] = R[e];
    __syncthreads();
    // Right-looking Cholesky of A (lower triangle); the pivot floor is the
    // reference's 1e-12.  Column j is copied to `col` as it is scaled, so
    // the rank-1 update reads it contiguously (a column of A is strided by
    // M * 16 bytes, which maps every row to one LDS bank).
    __shared__ cd col[64];
    for (int j = 0; j < M; ++j) {
        if (tid == 0) {
            const double d = A[j * M + j].re;
            if (!(d > 1e-12)) s_ok = 0;
            A[j * M + j] = cd{d > 1e-12 ? sqrt(d) : 1.0, 0.0};
        }
        __syncthreads();
        const double dinv = 1.0 / A[j * M + j].re;
        for (int i = j + 1 + tid; i < M; i += kThreads) {
            const cd v = A[i * M + j];
            const cd w{v.re * dinv, v.im * dinv};
            A[i * M + j] = w;
            col[i] = w;
        }
        __syncthreads();
        // A[i][k] -= L[i][j] conj(L[k][j]) for j < k <= i
        for (int i = j + 1 + (tid >> 4); i < M; i += kThreads / 16) {
            const cd a = col[i];
            for (int k = j + 1 + (tid & 15); k <= i; k += 16) {
                const cd b = col[k];
                A[i * M + k].re -= a.re * b.re + a.im * b.im;
                A[i * M + k].im -= a.im * b.re - a.re * b.im;
            }
        }
        __syncthreads();
    }
    const int ok = s_ok;
    if (!ok) {
        // Bartlett on the loaded, folded covariance (:354-375 fallback).
        for (int e = tid; e < MM; e += kThreads) {
            const int r = e / M, c = e - (e / M) * M;
            Q[e] = r < c ? make_float2((float)R[e].re, (float)R[e].im) : make_float2(0.0f, 0.0f);
        }
        if (tid == 0) {
            double tr = 0.0;
            for (int i = 0; i < M; ++i) tr += R[i * M + i].re;
            trace_out[f] = (float)tr;
            mode_out[f] = 0;
        }
        return;
    }
    // W = L^-1 (lower), one column per thread, written transposed:
    // Q[k][i] = W[i][k] so that lane i of the scan reads row k contiguously.
    cd* Wm = R;  // R is no longer needed
    for (int j = tid; j < M; j += kThreads) {
        for (int i = 0; i < M; ++i) {
            if (i < j) {
                Wm[i * M + j] = cd{0.0, 0.0};
                continue;
            }
            cd s{i == j ? 1.0 : 0.0, 0.0};
            for (int k = j; k < i; ++k) {
                const cd a = A[i * M + k], w = Wm[k * M + j];
                s.re -= a.re * w.re - a.im * w.im;
                s.im -= a.re * w.im + a.im * w.re;
            }
            const double inv = 1.0 / A[i * M + i].re;
            Wm[i * M + j] = cd{s.re * inv, s.im * inv};
        }
    }
    __syncthreads();
    for (int e = tid; e < MM; e += kThreads) {
        const int k = e / M, i = e - (e / M) * M;
        const cd w = Wm[i * M + k];
        Q[e] = make_float2((float)w.re, (float)w.im);
    }
    if (tid == 0) {
        trace_out[f] = 0.0f;
        mode_out[f] = 1;
    }
}

// ------------------------------------------------------------------ scan
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One pass of a wave over kNc cells: acc_lane = sum_r op(a_r, Q[r][lane]),
// the a_r broadcast from lane r.  BART: conj(a_r) R[r][lane] over the strict
// upper triangle; else W[lane][r] a_r (Q holds W transposed).  Q is a
// zero-padded 64 x 64 tile and a_r = 0 for r >= M, so rows are taken four
// at a time (the four LDS reads issue together) up to M rounded up.
template <bool BART>
__device__ __forceinline__ void scan_pass(const float2* __restrict__ Q, int M, int lane,
                                          const float2 (&a)[kNc], float2 (&acc)[kNc]) {
    for (int r0 = 0; r0 < M; r0 += 4) {
        float2 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = Q[(r0 + u) * 64 + lane];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = r0 + u;
#pragma unroll
            for (int j = 0; j < kNc; ++j) {
                const float ar = __builtin_bit_cast(
                    float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, a[j].x), r));
                const float ai = __builtin_bit_cast(
                    float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, a[j].y), r));
                if (BART) {
                    acc[j].x = fmaf(ai, q[u].y, fmaf(ar, q[u].x, acc[j].x));
                    acc[j].y = fmaf(-ai, q[u].x, fmaf(ar, q[u].y, acc[j].y));
                } else {
                    acc[j].x = fmaf(-q[u].y, ai, fmaf(q[u].x, ar, acc[j].x));
                    acc[j].y = fmaf(q[u].y, ar, fmaf(q[u].x, ai, acc[j].y));
                }
            }
        }
    }
}

template <bool BART>
__device__ void scan_cells(const float2* __restrict__ Q, const OptSearch& s,
                           const OptGridArgs& g, int level, int c_begin, int c_end, int lane,
                           int wave, int f, int F, double kw, float tr, float px, float py,
                           float pz, float* __restrict__ part) {
    const bool live = lane < g.M;
    for (int c0 = c_begin + wave * kNc; c0 < c_end; c0 += 4 * kNc) {
        float2 a[kNc], acc[kNc];
#pragma unroll
        for (int j = 0; j < kNc; ++j) {
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
        }
        scan_pass<BART>(Q, g.M, lane, a, acc);
#pragma unroll
        for (int j = 0; j < kNc; ++j) {
            float t;
            if (BART)
                t = acc[j].x * a[j].x - acc[j].y * a[j].y;  // Re(y_c a_c)
            else
                t = acc[j].x * acc[j].x + acc[j].y * acc[j].y;
            t = wave_sum(t);
            float p;
            if (BART)
                p = tr + 2.0f * t;
            else
                p = t > 1.0e-12f ? 1.0f / t : 0.0f;
            if (!isfinite(p)) p = 0.0f;
            if (lane == 0 && c0 + j < c_end) part[(size_t)(c0 + j) * F + f] = p;
        }
    }
}

// grid (F, chunks); part[cell][f] = the channel's power at the cell.
__global__ __launch_bounds__(kThreads) void opt_scan_kernel(
    const float2* __restrict__ scanmat, const float* __restrict__ trace,
    const int* __restrict__ mode, const float* __restrict__ pos, const float* __restrict__ freqs,
    const OptSearch* __restrict__ search, OptGridArgs g, int level, int n_cells,
    int cells_per_chunk, int F, float* __restrict__ part) {
    __shared__ __attribute__((aligned(16))) float2 Q[64 * 64];
    const int f = blockIdx.x;
    const int M = g.M;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int c_begin = (int)blockIdx.y * cells_per_chunk;
    const int c_end = min(n_cells, c_begin + cells_per_chunk);
    {
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
    if (mode[f] == 0)
        scan_cells<true>(Q, s, g, level, c_begin, c_end, lane, wave, f, F, kw, tr, px, py, pz,
                         part);
    else
        scan_cells<false>(Q, s, g, level, c_begin, c_end, lane, wave, f, F, kw, tr, px, py, pz,
                          part);
}

// ---------------------------------------------------------------- decide
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

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
    for (int cell = wave; cell < n_cells; cell += kDecideThreads / 64) {
        const float* pc = part + (size_t)cell * F;
        double sum = 0.0;
        for (int f = lane; f < F; f += 64) sum += (double)pc[f];
        sum = wave_sum_f64(sum);
        if (lane == 0) P[cell] = (float)sum;
    }
    __syncthreads();
    if (tid != 0) return;
    OptSearch s = *search;
    int best = 0;
    for (int i = 1; i < n_cells; ++i)
        if (P[i] > P[best]) best = i;
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
                     float lambda, float2* scanmat, float* trace, int* mode) {
    const size_t smem = 2 * M * M * sizeof(cd);
    hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(opt_fold_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem),
              "hipFuncSetAttribute");
    hipLaunchKernelGGL(opt_fold_kernel, dim3((unsigned)F), dim3(kThreads), smem, stream, rs,
                       state, first ? 1 : 0, (int)M, (int)K, (int)L, fb ? 1 : 0, capon ? 1 : 0,
                       load, lambda, scanmat, trace, mode);
    hip_check(hipGetLastError(), "opt_fold_kernel launch");
}

void launch_opt_scan(hipStream_t stream, const float2* scanmat, const float* trace,
                     const int* mode, const float* pos, const float* freqs,
                     const OptSearch* search, const OptGridArgs& g, int level, size_t F,
                     size_t n_cells, float* part) {
    // enough workgroups to cover the CUs twice; each stages its channel's
    // operand once
    const size_t per_pass = 4 * kNc;
    size_t chunks = std::max<size_t>(1, ceil_div(512, F));
    chunks = std::min(chunks, ceil_div(n_cells, per_pass));
    const size_t cpc = ceil_div(ceil_div(n_cells, chunks), per_pass) * per_pass;
    chunks = ceil_div(n_cells, cpc);
    hipLaunchKernelGGL(opt_scan_kernel, dim3((unsigned)F, (unsigned)chunks), dim3(kThreads), 0,
                       stream, scanmat, trace, mode, pos, freqs, search, g, level, (int)n_cells,
                       (int)cpc, (int)F, part);
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
