// k_doa.hip -- adaptive direction-of-arrival estimation for gfx950.
//
// Per channel f and update (K snapshots t0 .. t0+K-1):
//   1. covariance_kernel (int8 MFMA, exact): R[r][c] = sum_k x_r conj(x_c)
//      over the packed int4 snapshots (reference
//      src/cpu_opt_beam_tracker.cpp:559-605 without the 1/K scale), then
//      forward spatial smoothing Rs = sum_{l<L} R[l:l+M, l:l+M], L = A-M+1
//      (:607-636, forward part).  Integers throughout, so bit-exact.
//   2. fold_scan_kernel (FP64): R_block = Rs/(K L), diagonal load
//      eps = load * tr(R_block) / M added before the fold (:244-253,
//      :1874-1890), R = alpha R + (1-alpha) R_block (first update copies),
//      Q = R^-1 by register-blocked Gauss-Jordan (gauss_jordan.hpp; its
//      pivots are the Cholesky pivots of :398-421), and the Bartlett
//      a^H R a and Capon 1/(a^H Q a) scans over n_theta angles of a uniform
//      linear array, steering a_m = exp(-j k x_m sin(theta)) (:96-106).
//      For a ULA a^H Q a = sum_d q_d exp(j d k dx sin(theta)) with q_d the
//      sum of Q's d-th diagonal, so each angle costs O(M), not O(M^2).
//   3. broadband_kernel + peaks_kernel: broadband spectra summed over
//      channels, local maxima, 3-point parabolic refinement.
//
// MFMA mapping: with y = [16 re, 16 im] int8 components per antenna (the
// nibble placed in the top half of a byte is 16x the signed value), the
// real Gram matrix G = Y^T Y (128 x 128 for A = 64) is accumulated with
// v_mfma_i32_16x16x64_i8 over 64-snapshot steps; both operands are
// "component x 16 snapshots" rows of a transposed LDS tile.  Then
// Re R = (G_rr + G_ii)/256 and Im R = (G_ir - G_ri)/256 exactly.
#include <algorithm>

#include "bf_doa.hpp"
#include "gauss_jordan.hpp"

namespace bfhip {
namespace {

constexpr int kThreads = 256;
constexpr int kTStride = 80;  // bytes per component row of the transposed tile

// Transposes the 4 x 4 byte block of d[0..3] (rows) into e[0..3] (columns).
__device__ __forceinline__ void transpose4x4(const unsigned d[4], unsigned e[4]) {
    const unsigned lo01 = __builtin_amdgcn_perm(d[1], d[0], 0x05010400u);
    const unsigned hi01 = __builtin_amdgcn_perm(d[1], d[0], 0x07030602u);
    const unsigned lo23 = __builtin_amdgcn_perm(d[3], d[2], 0x05010400u);
    const unsigned hi23 = __builtin_amdgcn_perm(d[3], d[2], 0x07030602u);
    e[0] = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);
    e[1] = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);
    e[2] = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);
    e[3] = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);
}

// Tile t of the upper triangle (I <= J) of an n x n tile grid, row by row.
__device__ __forceinline__ void upper_tile(int t, int n, int* I, int* J) {
    int i = 0;
    while (t >= n - i) {
        t -= n - i;
        ++i;
    }
    *I = i;
    *J = i + t;
}

// One workgroup per (channel, part of the snapshot range); A in {32, 64}
// -> 2A components.  Part p (blockIdx.y) covers snapshots kpart p ..
// kpart (p + 1) - 1 of the update (never empty: the launcher sizes the
// parts) and writes its own smoothed partial sum to rs_out + p F M M 2;
// the consumer adds the parts (integers, so in any order).  G is
// symmetric, so only the tiles on and above the diagonal are accumulated;
// every entry the covariance reads for r <= c lies there, and r > c is the
// conjugate of (c, r).
template <int A>
__global__ __launch_bounds__(kThreads) void covariance_kernel(
    const uint8_t* __restrict__ x, int T, int F, int t0, int K, int M, int kpart,
    int* __restrict__ rs_out /* [parts][F][M][M][2] */,
    int* __restrict__ r_full /* [F][A][A][2] or null (one part only) */) {
    constexpr int C = 2 * A;         // components
    constexpr int NTILE = C / 16;    // tile rows/cols of G
    constexpr int TILES = NTILE * (NTILE + 1) / 2;  // upper triangle
    constexpr int PER_WAVE = (TILES + 3) / 4;
    __shared__ __attribute__((aligned(16))) unsigned char tr[C * kTStride];
    __shared__ int g[C * (C + 1)];  // G, row stride C+1
    // Workgroups are dealt round-robin to the 8 XCDs; give each XCD a
    // contiguous range of channels so that channels f and f+1, which share
    // the 128-byte lines of every packed row, meet in one L2.
    const int per_xcd = (F + 7) / 8;
    const int f = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (f >= F) return;  // whole workgroup
    {
        const int kb = (int)blockIdx.y * kpart;
        t0 += kb;
        K = min(K - kb, kpart);
        rs_out += (size_t)blockIdx.y * F * M * M * 2;
    }
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    bf_int4v acc[PER_WAVE];
    int tI[PER_WAVE], tJ[PER_WAVE];  // this wave's tiles, wave-uniform
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i) {
        acc[i] = bf_int4v{0, 0, 0, 0};
        upper_tile(min(wave * PER_WAVE + i, TILES - 1), NTILE, &tI[i], &tJ[i]);
    }
    const size_t row_stride = (size_t)F * A;
    const uint8_t* xf = x + (size_t)f * A;
    // Staging work: 64 snapshots x A antennas = 16 snapshot quads x A/4
    // antenna quads, QPT per thread.  A ring of kDepth chunks is in flight
    // in registers: chunk c is staged while chunks c+1 .. c+kDepth-1 load.
    // Rows past the window read row t0 (in bounds: the launcher checks
    // t0 + K <= T), so the loads need no branches; their values are replaced
    // by zeros when the chunk is consumed (a select at load time would wait
    // for the load).
    constexpr int QUADS = 16 * (A / 4);
    constexpr int QPT = (QUADS + kThreads - 1) / kThreads;
    constexpr int kDepth = 4;
    unsigned ring[kDepth][QPT][4];
    auto in_window = [&](int j, int i, int k0) {
        const int q = tid + j * kThreads;
        return q < QUADS && k0 + 4 * (q / (A / 4)) + i < K;
    };
    auto load_chunk = [&](unsigned (&dst)[QPT][4], int k0) {
#pragma unroll
        for (int j = 0; j < QPT; ++j) {
            const int q = tid + j * kThreads;
            const int k4 = q / (A / 4);
            const int a4 = q < QUADS ? q - k4 * (A / 4) : 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int kk = in_window(j, i, k0) ? k0 + 4 * k4 + i : 0;
                dst[j][i] = *reinterpret_cast<const unsigned*>(
                    xf + (size_t)(t0 + kk) * row_stride + 4 * a4);
            }
        }
    };
    // one chunk: transpose into LDS, then the MFMA step of every tile
    auto process = [&](const unsigned (&cur)[QPT][4]) {
#pragma unroll
        for (int jq = 0; jq < QPT; ++jq) {
            const int q = tid + jq * kThreads;
            if (q >= QUADS) continue;
            const int k4 = q / (A / 4);
            const int a4 = q - k4 * (A / 4);
            const unsigned* d = cur[jq];
            unsigned e[4];
            transpose4x4(d, e);
            // Lanes of one instruction share k4 in groups of 16 consecutive
            // a4, whose rows lie 8 apart (a multiple of 32 banks at this
            // stride); writing antenna (jj + a4) mod 4 in instruction jj
            // spreads a group over 4 bank groups.
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int j = (jj + a4) & 3;
                const unsigned ej = j == 0 ? e[0] : j == 1 ? e[1] : j == 2 ? e[2] : e[3];
                const int a = 4 * a4 + j;
                const unsigned re = (ej << 4) & 0xF0F0F0F0u;  // 16 * real
                const unsigned im = ej & 0xF0F0F0F0u;         // 16 * imag
                *reinterpret_cast<unsigned*>(tr + (2 * a) * kTStride + 4 * k4) = re;
                *reinterpret_cast<unsigned*>(tr + (2 * a + 1) * kTStride + 4 * k4) = im;
            }
        }
        __syncthreads();
        const int rc = lane & 15;
        const int kg = lane >> 4;
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i) {
            const int tile = wave * PER_WAVE + i;
            if (tile < TILES) {
                const int I = tI[i], J = tJ[i];
                const bf_int4v av =
                    *reinterpret_cast<const bf_int4v*>(tr + (16 * I + rc) * kTStride + 16 * kg);
                const bf_int4v bv =
                    *reinterpret_cast<const bf_int4v*>(tr + (16 * J + rc) * kTStride + 16 * kg);
                acc[i] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bv, acc[i], 0, 0, 0);
            }
        }
        __syncthreads();
    };
    auto take = [&](int s_k0, const unsigned (&slot)[QPT][4], unsigned (&cur)[QPT][4]) {
#pragma unroll
        for (int j = 0; j < QPT; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) cur[j][i] = in_window(j, i, s_k0) ? slot[j][i] : 0u;
    };
#pragma unroll
    for (int s = 0; s < kDepth; ++s) load_chunk(ring[s], 64 * s);
    // full groups of kDepth chunks: each slot is consumed, then refilled
    // unconditionally (past the window it reads row t0), so every slot has
    // one producer and the ring needs no register copies
    int kr = 0;
    for (; kr + 64 * kDepth <= K; kr += 64 * kDepth) {
#pragma unroll
        for (int s = 0; s < kDepth; ++s) {
            unsigned cur[QPT][4];
            take(kr + 64 * s, ring[s], cur);
            load_chunk(ring[s], kr + 64 * (s + kDepth));
            process(cur);
        }
    }
    // the last K mod (64 kDepth) snapshots: up to kDepth chunks, already loaded
#pragma unroll
    for (int s = 0; s < kDepth; ++s) {
        if (kr + 64 * s < K) {  // uniform
            unsigned cur[QPT][4];
            take(kr + 64 * s, ring[s], cur);
            process(cur);
        }
    }
    // G tiles -> LDS: lane holds column l%16, rows 4*(l/16) + r.
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i) {
        const int tile = wave * PER_WAVE + i;
        if (tile < TILES) {
            const int I = tI[i], J = tJ[i];
#pragma unroll
            for (int r = 0; r < 4; ++r)
                g[(16 * I + 4 * (lane >> 4) + r) * (C + 1) + 16 * J + (lane & 15)] = acc[i][r];
        }
    }
    __syncthreads();
    if (r_full != nullptr) {
        for (int e = tid; e < A * A; e += kThreads) {
            const int r = e / A, c = e - (e / A) * A;
            const int a = r <= c ? r : c, b = r <= c ? c : r;  // upper entry
            const int gre = g[(2 * a) * (C + 1) + 2 * b] + g[(2 * a + 1) * (C + 1) + 2 * b + 1];
            const int gim = g[(2 * a + 1) * (C + 1) + 2 * b] - g[(2 * a) * (C + 1) + 2 * b + 1];
            r_full[(((size_t)f * A + r) * A + c) * 2] = gre / 256;
            r_full[(((size_t)f * A + r) * A + c) * 2 + 1] = (r <= c ? gim : -gim) / 256;
        }
    }
    // Rs[a][b] = sum_{l<L} G'(a+l, b+l) on the diagonals b = a + d >= 0,
    // with G'(r, c) = (G_rr + G_ii, G_ir - G_ri) / 256 (every G entry is a
    // multiple of 256).  Each thread walks a segment of kSeg entries of one
    // diagonal as a running sum, Rs(a+1, b+1) = Rs(a, b) - G'(a, b) +
    // G'(a+L, b+L), and writes the entry and its conjugate (b, a): integers,
    // so the order of the sums does not matter.
    constexpr int kSeg = 8;
    const int L = A - M + 1;
    auto gp = [&](int r, int c, int& re, int& im) {  // r <= c
        re = (g[(2 * r) * (C + 1) + 2 * c] + g[(2 * r + 1) * (C + 1) + 2 * c + 1]) / 256;
        im = (g[(2 * r + 1) * (C + 1) + 2 * c] - g[(2 * r) * (C + 1) + 2 * c + 1]) / 256;
    };
    int2* rs2 = reinterpret_cast<int2*>(rs_out) + (size_t)f * M * M;
    for (int w = tid;; w += kThreads) {
        int d = 0, seg = w;
        while (d < M) {
            const int ns = (M - d + kSeg - 1) / kSeg;
            if (seg < ns) break;
            seg -= ns;
            ++d;
        }
        if (d >= M) break;
        const int a0 = seg * kSeg;
        const int a1 = min(a0 + kSeg, M - d);
        int sre = 0, sim = 0;
        for (int l = 0; l < L; ++l) {
            int re, im;
            gp(a0 + l, a0 + d + l, re, im);
            sre += re;
            sim += im;
        }
        for (int a = a0; a < a1; ++a) {
            if (a > a0) {
                int re, im;
                gp(a - 1, a - 1 + d, re, im);
                sre -= re;
                sim -= im;
                gp(a + L - 1, a + L - 1 + d, re, im);
                sre += re;
                sim += im;
            }
            rs2[a * M + a + d] = make_int2(sre, sim);
            if (d > 0) rs2[(a + d) * M + a] = make_int2(sre, -sim);
        }
    }
}

// ------------------------------------------------------------- FP64 stage
struct cd {
    double re, im;
};
__device__ __forceinline__ cd cmul(cd a, cd b) {
    return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// out[d] = sum_c X[c+d][c] for d < M <= 64 (row stride MP): four lanes per
// diagonal (4 M <= 256 threads), each summing every fourth entry, combined
// in a fixed order.
__device__ __forceinline__ void diagonal_sums(const cd* X, int M, int MP, cd* out, int tid) {
    if (tid >= 4 * M) return;  // whole groups of four lanes
    const int d = tid >> 2, part = tid & 3;
    cd s{0.0, 0.0};
    for (int c = part; c + d < M; c += 4) {
        const cd v = X[(c + d) * MP + c];
        s.re += v.re;
        s.im += v.im;
    }
    s.re += __shfl_xor(s.re, 1);
    s.im += __shfl_xor(s.im, 1);
    s.re += __shfl_xor(s.re, 2);
    s.im += __shfl_xor(s.im, 2);
    if (part == 0) out[d] = s;
}

// One workgroup of 256 threads (16 x 16 Gauss-Jordan blocks) per channel;
// 2 x 2 blocks on 1024 threads measured slower (16-wave barriers, twice
// the LDS reads).
// buf0: M x M complex double in LDS, row stride M + 1 (a stride of M
// complex doubles is a multiple of the 32 banks, which would put every
// lane of the diagonal walks below on one bank).
constexpr int kFoldSide = 16;                          // Gauss-Jordan blocks per side
constexpr int kFoldThreads = kFoldSide * kFoldSide;   // 256
// BS x BS Gauss-Jordan blocks: 3 for M <= 48, 4 for M <= 64.
template <int BS>
__global__ __launch_bounds__(kFoldThreads) void fold_scan_kernel(
    const int* __restrict__ rs /* [parts][F][M][M][2] */, int parts,
    double* __restrict__ state /* [F][M][M][2] */, int first,
    int M, int K, int L, double alpha, double load, const float* __restrict__ freqs,
    double x_step, int n_theta, const double* __restrict__ sin_theta,
    double* __restrict__ capon, double* __restrict__ bartlett, int* __restrict__ chol_ok) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int MP = M + 1;
    cd* buf0 = reinterpret_cast<cd*>(smem);
    cd* qb = buf0 + M * MP;     // [M] Bartlett diagonal sums, d = 0..M-1
    cd* qc = qb + M;            // [M] Capon diagonal sums
    __shared__ double s_eps;
    __shared__ int s_ok;
    __shared__ GjShared<cd> gj;
    const int f = blockIdx.x;
    const int tid = threadIdx.x;
    const int MM = M * M;
    const double scale = 1.0 / ((double)K * (double)L);
    // R_block = Rs / (K L), folded into the carried (unloaded) state
    double* st = state + (size_t)f * MM * 2;
    const int F = gridDim.x;
    for (int e = tid; e < MM; e += kFoldThreads) {
        int sre = 0, sim = 0;  // the parts' integer sums (exact)
        for (int q = 0; q < parts; ++q) {
            const int2 p = reinterpret_cast<const int2*>(rs)[((size_t)q * F + f) * MM + e];
            sre += p.x;
            sim += p.y;
        }
        cd v{(double)sre * scale, (double)sim * scale};
        if (!first) {
            v.re = alpha * st[2 * e] + (1.0 - alpha) * v.re;
            v.im = alpha * st[2 * e + 1] + (1.0 - alpha) * v.im;
        }
        st[2 * e] = v.re;
        st[2 * e + 1] = v.im;
        buf0[(e / M) * MP + e % M] = v;
    }
    __syncthreads();
    // The reference loads each R_block by load * tr(R_block) / M before its
    // fold, for Capon only.  The trace is linear, so that loaded recursion
    // equals the unloaded one plus load * tr(R) / M on the diagonal: Bartlett
    // scans R, Capon inverts R + eps I.
    if (tid < 64) {  // the trace of the folded R, reduced across wave 0 (M <= 64)
        double tr = tid < M ? buf0[tid * MP + tid].re : 0.0;
        for (int o = 32; o > 0; o >>= 1) tr += __shfl_xor(tr, o);
        if (tid == 0) {
            s_eps = load > 0.0 ? load * tr / (double)M : 0.0;
            s_ok = 1;
        }
    }
    // Bartlett diagonal sums of R: qb[d] = sum_c R[c+d][c]
    diagonal_sums(buf0, M, MP, qb, tid);
    __syncthreads();
    if (tid < M) buf0[tid * MP + tid].re += s_eps;
    __syncthreads();
    // Q = R^-1 in place by register-blocked Gauss-Jordan; its pivots are
    // the Cholesky pivots, so R fails the factorisation exactly when a
    // pivot is not positive.
    {
        cd blk[BS][BS];
        gj_load<BS, kFoldSide>(blk, reinterpret_cast<const double*>(buf0), M, MP, tid);
        __syncthreads();  // every block read before Q overwrites buf0
        const bool ok_gj = gj_invert<BS, kFoldSide>(blk, M, 0.0, gj, tid);
        if (tid == 0) s_ok = ok_gj ? 1 : 0;
        if (ok_gj) {
            const int i0 = gj_row0<BS, kFoldSide>(tid), j0 = gj_col0<BS, kFoldSide>(tid);
#pragma unroll
            for (int u = 0; u < BS; ++u)
#pragma unroll
                for (int v = 0; v < BS; ++v)
                    if (i0 + u < M && j0 + v < M) buf0[(i0 + u) * MP + j0 + v] = blk[u][v];
        }
        __syncthreads();
    }
    const int ok = s_ok;
    if (tid == 0) chol_ok[f] = ok;
    diagonal_sums(buf0, M, MP, qc, tid);
    __syncthreads();
    // a^H X a = q_0 + 2 Re sum_{d>0} q_d e^{j d phi}, phi = k dx sin(theta).
    // kAngles angles per thread side by side, so that their recursions
    // e <- e * step (a dependent chain per angle) overlap.
    constexpr int kAngles = 3;
    const double kw = kTwoPi * (double)freqs[f] / kSpeedOfLight;
    for (int q0 = tid; q0 < n_theta; q0 += kAngles * kFoldThreads) {
        cd step[kAngles], e[kAngles];
        double pb[kAngles], pc[kAngles];
#pragma unroll
        for (int i = 0; i < kAngles; ++i) {
            const int q = min(q0 + i * kFoldThreads, n_theta - 1);
            const double phi = kw * x_step * sin_theta[q];
            double sn, cs;
            sincos(phi, &sn, &cs);
            step[i] = cd{cs, sn};
            e[i] = cd{1.0, 0.0};
            pb[i] = qb[0].re;
            pc[i] = qc[0].re;
        }
        for (int d = 1; d < M; ++d) {
            const cd b = qb[d], c = qc[d];
#pragma unroll
            for (int i = 0; i < kAngles; ++i) {
                e[i] = cmul(e[i], step[i]);
                pb[i] += 2.0 * (b.re * e[i].re - b.im * e[i].im);
                pc[i] += 2.0 * (c.re * e[i].re - c.im * e[i].im);
            }
        }
#pragma unroll
        for (int i = 0; i < kAngles; ++i) {
            const int q = q0 + i * kFoldThreads;
            if (q < n_theta) {
                bartlett[(size_t)f * n_theta + q] = pb[i];
                capon[(size_t)f * n_theta + q] = (ok && pc[i] > 0.0) ? 1.0 / pc[i] : 0.0;
            }
        }
    }
}

// Broadband spectra: sum over channels.  A workgroup covers 64 angles; its
// kBbParts waves each sum every kBbParts-th channel (f = part, part +
// kBbParts, ...) into four running sums (channels taken four at a time, so
// that eight loads are in flight), combined in a fixed order, and the
// partial sums are added in part order.
constexpr int kBbParts = 16;

__global__ __launch_bounds__(64 * kBbParts) void broadband_kernel(
    const double* __restrict__ capon, const double* __restrict__ bartlett, int F, int n_theta,
    double* __restrict__ broadband /* [2][n] */) {
    __shared__ double part_c[kBbParts][64];
    __shared__ double part_b[kBbParts][64];
    const int ql = threadIdx.x & 63;
    const int part = threadIdx.x >> 6;
    const int q = blockIdx.x * 64 + ql;
    double sc = 0.0, sb = 0.0;
    if (q < n_theta) {
        double c4[4] = {0.0, 0.0, 0.0, 0.0}, b4[4] = {0.0, 0.0, 0.0, 0.0};
        int f = part;
        for (; f + 3 * kBbParts < F; f += 4 * kBbParts) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                c4[u] += capon[(size_t)(f + u * kBbParts) * n_theta + q];
                b4[u] += bartlett[(size_t)(f + u * kBbParts) * n_theta + q];
            }
        }
        for (; f < F; f += kBbParts) {
            c4[0] += capon[(size_t)f * n_theta + q];
            b4[0] += bartlett[(size_t)f * n_theta + q];
        }
        sc = (c4[0] + c4[1]) + (c4[2] + c4[3]);
        sb = (b4[0] + b4[1]) + (b4[2] + b4[3]);
    }
    part_c[part][ql] = sc;
    part_b[part][ql] = sb;
    __syncthreads();
    if (part == 0 && q < n_theta) {
        double c = part_c[0][ql], b = part_b[0][ql];
        for (int i = 1; i < kBbParts; ++i) {
            c += part_c[i][ql];
            b += part_b[i][ql];
        }
        broadband[q] = c;
        broadband[n_theta + q] = b;
    }
}

constexpr int kMaxCandidates = 4096;
constexpr int kPeaksLds = 4096;  // spectra up to this many angles staged in LDS
constexpr int kRankMax = 1024;   // candidate lists up to this long are ranked in parallel

// Angle of the parabola vertex through (q-1, q, q+1), the offset clamped to
// +-0.5 cell (0 where the three points are not concave).
__device__ __forceinline__ double refine_peak(const double* P, const double* theta, int q) {
    const double pm = P[q - 1], pc = P[q], pp = P[q + 1];
    const double den = pm - 2.0 * pc + pp;
    double dlt = 0.0;
    if (den < 0.0) {
        dlt = 0.5 * (pm - pp) / den;
        dlt = dlt > 0.5 ? 0.5 : (dlt < -0.5 ? -0.5 : dlt);
    }
    return theta[q] + dlt * (theta[q + 1] - theta[q]);
}

// Peaks of one broadband spectrum (workgroup 0: Capon, 1: Bartlett): local
// maxima (P[q] > P[q-1] and P[q] >= P[q+1]) are found in parallel and
// compacted in increasing q, each is refined with the parabola through
// (q-1, q, q+1), and the max_peaks strongest are kept, earlier angles first
// on ties.  Up to kRankMax candidates, every candidate's slot is its rank in
// that order, counted in parallel; longer lists go through one thread's
// insertion into the output, which gives the same list.
__global__ __launch_bounds__(kThreads) void peaks_kernel(
    const double* __restrict__ broadband, int n_theta, const double* __restrict__ theta,
    int max_peaks, double* __restrict__ peaks /* [2][max_peaks][2] theta, power */,
    int* __restrict__ n_peaks /* [2] */) {
    __shared__ int cand[kMaxCandidates];
    __shared__ double cpow[kRankMax];
    __shared__ int wave_count[kThreads / 64];
    __shared__ int total;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    // the spectrum, staged in LDS when it fits (detection and refinement
    // read it three times per angle)
    const double* Pg = broadband + (size_t)blockIdx.x * n_theta;
    const double* P = Pg;
    if (n_theta <= kPeaksLds) {
        double* Ps = reinterpret_cast<double*>(smem);
        for (int q = threadIdx.x; q < n_theta; q += kThreads) Ps[q] = Pg[q];
        __syncthreads();
        P = Ps;
    }
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    if (tid == 0) total = 0;
    __syncthreads();
    for (int q0 = 0; q0 < n_theta; q0 += kThreads) {
        const int q = q0 + tid;
        const bool peak = q >= 1 && q + 1 < n_theta && P[q] > P[q - 1] && P[q] >= P[q + 1];
        const unsigned long long m = __ballot(peak);
        if (lane == 0) wave_count[wave] = __popcll(m);
        __syncthreads();
        int base = total;
        for (int w = 0; w < wave; ++w) base += wave_count[w];
        if (peak) {
            const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
            if (pos < kMaxCandidates) cand[pos] = q;
        }
        __syncthreads();
        if (tid == 0)
            for (int w = 0; w < kThreads / 64; ++w) total += wave_count[w];
        __syncthreads();
    }
    const int n_cand = total;  // uniform: read after the last barrier
    double* out = peaks + (size_t)blockIdx.x * max_peaks * 2;
    if (n_cand <= kRankMax) {
        for (int i = tid; i < n_cand; i += kThreads) cpow[i] = P[cand[i]];
        const int n_out = n_cand < max_peaks ? n_cand : max_peaks;
        for (int i = n_out + tid; i < max_peaks; i += kThreads) out[2 * i] = out[2 * i + 1] = 0.0;
        __syncthreads();
        // rank = candidates ahead in (power descending, q ascending); the
        // ranks are a permutation of 0 .. n_cand-1, so every slot below
        // n_out is written exactly once
        for (int i = tid; i < n_cand; i += kThreads) {
            const double pc = cpow[i];
            int rank = 0;
            for (int j = 0; j < n_cand; ++j) {
                const double pj = cpow[j];
                rank += (pj > pc || (pj == pc && j < i)) ? 1 : 0;
            }
            if (rank < max_peaks) {
                out[2 * rank] = refine_peak(P, theta, cand[i]);
                out[2 * rank + 1] = pc;
            }
        }
        if (tid == 0) n_peaks[blockIdx.x] = n_out;
        return;
    }
    if (tid != 0) return;
    for (int i = 0; i < max_peaks; ++i) out[2 * i] = out[2 * i + 1] = 0.0;
    const bool listed = n_cand <= kMaxCandidates;
    const int n_scan = listed ? n_cand : n_theta;
    int n = 0;
    for (int i = 0; i < n_scan; ++i) {
        const int q = listed ? cand[i] : i;
        if (!listed && !(q >= 1 && q + 1 < n_theta && P[q] > P[q - 1] && P[q] >= P[q + 1]))
            continue;
        const double pc = P[q];
        const double th = refine_peak(P, theta, q);
        // insertion into the top-max_peaks list by power
        int pos = n < max_peaks ? n : max_peaks;
        while (pos > 0 && out[2 * (pos - 1) + 1] < pc) {
            if (pos < max_peaks) {
                out[2 * pos] = out[2 * (pos - 1)];
                out[2 * pos + 1] = out[2 * (pos - 1) + 1];
            }
            --pos;
        }
        if (pos < max_peaks) {
            out[2 * pos] = th;
            out[2 * pos + 1] = pc;
            if (n < max_peaks) ++n;
        }
    }
    n_peaks[blockIdx.x] = n;
}

}  // namespace

size_t doa_covariance_parts(size_t K, size_t max_parts) {
    const size_t chunks = ceil_div(K, 64);
    const size_t per = ceil_div(chunks, std::max<size_t>(1, std::min(max_parts, chunks)));
    return ceil_div(chunks, per);
}

size_t launch_doa_covariance(hipStream_t stream, const uint8_t* x, size_t T, size_t F,
                             size_t A, size_t t0, size_t K, size_t M, int* rs, int* r_full,
                             size_t max_parts) {
    if (K == 0 || t0 + K > T) invalid("the covariance window exceeds the time axis");
    // Parts of whole 64-snapshot chunks, none empty: part p starts at
    // kpart p <= 64 (chunks - 1) < K.
    const size_t parts = r_full != nullptr ? 1 : doa_covariance_parts(K, max_parts);
    const size_t kpart = 64 * ceil_div(ceil_div(K, 64), parts);
    const dim3 grid((unsigned)(8 * ceil_div(F, 8)), (unsigned)parts);
#define BF_COV(A_)                                                                            \
    hipLaunchKernelGGL(covariance_kernel<A_>, grid, dim3(kThreads), 0, stream, x, (int)T,    \
                       (int)F, (int)t0, (int)K, (int)M, (int)kpart, rs, r_full)
    switch (A) {
    case 32: BF_COV(32); break;
    case 64: BF_COV(64); break;
    default: invalid("the DOA covariance supports 32 or 64 antennas");
    }
#undef BF_COV
    hip_check(hipGetLastError(), "covariance_kernel launch");
    return parts;
}

void launch_doa_fold_scan(hipStream_t stream, const int* rs, size_t parts, double* state,
                          bool first,
                          size_t F, size_t M, size_t K, size_t L, double alpha, double load,
                          const float* freqs, double x_step, size_t n_theta,
                          const double* sin_theta, double* capon, double* bartlett,
                          int* chol_ok) {
    if (M == 0 || M > 4 * kFoldSide) invalid("the DOA fold supports 1 <= subarray <= 64");
    const size_t smem = (M * (M + 1) + 2 * M) * sizeof(double) * 2;
    auto go = [&](auto kernel) {
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem),
                  "hipFuncSetAttribute");
        hipLaunchKernelGGL(kernel, dim3((unsigned)F), dim3(kFoldThreads), smem, stream, rs,
                           (int)parts, state, first ? 1 : 0, (int)M, (int)K, (int)L, alpha, load, freqs, x_step,
                           (int)n_theta, sin_theta, capon, bartlett, chol_ok);
    };
    if (M <= 3 * kFoldSide)
        go(fold_scan_kernel<3>);
    else
        go(fold_scan_kernel<4>);
    hip_check(hipGetLastError(), "fold_scan_kernel launch");
}

void launch_doa_estimate(hipStream_t stream, const double* capon, const double* bartlett,
                         size_t F, size_t n_theta, const double* theta, size_t max_peaks,
                         double* broadband, double* peaks, int* n_peaks) {
    hipLaunchKernelGGL(broadband_kernel, dim3((unsigned)ceil_div(n_theta, 64)),
                       dim3(64 * kBbParts), 0, stream, capon, bartlett, (int)F, (int)n_theta,
                       broadband);
    const size_t smem = n_theta <= (size_t)kPeaksLds ? n_theta * sizeof(double) : 0;
    hipLaunchKernelGGL(peaks_kernel, dim3(2), dim3(kThreads), smem, stream, broadband,
                       (int)n_theta, theta, (int)max_peaks, peaks, n_peaks);
    hip_check(hipGetLastError(), "estimate kernels launch");
}

}  // namespace bfhip
