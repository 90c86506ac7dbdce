// gauss_jordan.hpp -- in-place inverse of a Hermitian positive-definite
// complex FP64 matrix (M <= NB BS <= 64) by Gauss-Jordan elimination
// without pivoting, for one workgroup of NB^2 threads.  Thread t holds the
// BS x BS block of rows BS (t / NB) .. and columns BS (t % NB) .. in
// registers (identity outside M x M); NB BS just above M keeps the padding,
// which costs the same arithmetic as the matrix, small (NB = 16: BS = 3 for
// M <= 48, BS = 4 for M <= 64).  Per pivot k the owners publish row k and
// column k to LDS (double-buffered, so one barrier per pivot), every thread
// applies the elimination update to its whole block, and the threads
// holding row or column k then overwrite those entries.  Without pivoting
// on a Hermitian positive-definite matrix the k-th pivot equals the
// Cholesky pivot L_kk^2, so a factorisation test on it carries over
// unchanged.
#pragma once

#include <hip/hip_runtime.h>

namespace bfhip {

template <typename Cd>
struct GjShared {
    Cd row[2][64];
    Cd col[2][64];
};

// Loads the block of a row-major M x M matrix with row stride ld (re, im at
// 2 (i ld + j)) held in `src` (global or LDS) into registers.
template <int BS, int NB = 64 / BS>
__device__ __forceinline__ int gj_row0(int tid) { return BS * (tid / NB); }
template <int BS, int NB = 64 / BS>
__device__ __forceinline__ int gj_col0(int tid) { return BS * (tid % NB); }

template <int BS, int NB = 64 / BS, typename Cd>
__device__ __forceinline__ void gj_load(Cd (&a)[BS][BS], const double* src, int M, int ld,
                                        int tid) {
    static_assert(NB * BS <= 64, "the published row and column hold 64 entries");
    const int i0 = gj_row0<BS, NB>(tid), j0 = gj_col0<BS, NB>(tid);
#pragma unroll
    for (int u = 0; u < BS; ++u)
#pragma unroll
        for (int v = 0; v < BS; ++v) {
            const int i = i0 + u, j = j0 + v;
            if (i < M && j < M)
                a[u][v] = Cd{src[2 * (i * ld + j)], src[2 * (i * ld + j) + 1]};
            else
                a[u][v] = Cd{i == j ? 1.0 : 0.0, 0.0};
        }
}

// Inverts in place (M <= NB BS, which the caller guarantees); returns false
// (uniformly) at the first pivot k < M that is not above `floor`, leaving
// the block unspecified.
template <int BS, int NB = 64 / BS, typename Cd>
__device__ bool gj_invert(Cd (&a)[BS][BS], int M, double floor, GjShared<Cd>& sh, int tid) {
    static_assert(NB * BS <= 64, "the published row and column hold 64 entries");
    const int i0 = gj_row0<BS, NB>(tid), j0 = gj_col0<BS, NB>(tid);
    // pivots beyond M act on the identity padding only: skipped
    for (int k = 0; k < M; ++k) {
        const int buf = k & 1;
        // static register indices only (a runtime index would spill a[][])
#pragma unroll
        for (int u = 0; u < BS; ++u)
            if (i0 + u == k) {
#pragma unroll
                for (int v = 0; v < BS; ++v) sh.row[buf][j0 + v] = a[u][v];
            }
#pragma unroll
        for (int v = 0; v < BS; ++v)
            if (j0 + v == k) {
#pragma unroll
                for (int u = 0; u < BS; ++u) sh.col[buf][i0 + u] = a[u][v];
            }
        __syncthreads();
        const double piv = sh.row[buf][k].re;
        if (!(piv > floor)) return false;  // uniform: every thread reads the same pivot
        const double p = 1.0 / piv;
        Cd rj[BS], cp[BS];
#pragma unroll
        for (int v = 0; v < BS; ++v) rj[v] = sh.row[buf][j0 + v];
#pragma unroll
        for (int u = 0; u < BS; ++u) {
            const Cd ci = sh.col[buf][i0 + u];
            cp[u] = Cd{ci.re * p, ci.im * p};
        }
        // every entry: a_ij - a_ik a_kj / a_kk
#pragma unroll
        for (int u = 0; u < BS; ++u)
#pragma unroll
            for (int v = 0; v < BS; ++v) {
                const double gre =
                    fma(-cp[u].re, rj[v].re, fma(cp[u].im, rj[v].im, a[u][v].re));
                const double gim =
                    fma(-cp[u].re, rj[v].im, fma(-cp[u].im, rj[v].re, a[u][v].im));
                a[u][v] = Cd{gre, gim};
            }
        // then the pivot row (a_kj / a_kk; the pivot 1 / a_kk) and the pivot
        // column (-a_ik / a_kk) over what the line above wrote there: only
        // the threads holding row or column k take these branches
#pragma unroll
        for (int u = 0; u < BS; ++u)
            if (i0 + u == k) {
#pragma unroll
                for (int v = 0; v < BS; ++v)
                    a[u][v] = j0 + v == k ? Cd{p, 0.0} : Cd{rj[v].re * p, rj[v].im * p};
            }
#pragma unroll
        for (int v = 0; v < BS; ++v)
            if (j0 + v == k) {
#pragma unroll
                for (int u = 0; u < BS; ++u)
                    if (i0 + u != k) a[u][v] = Cd{-cp[u].re, -cp[u].im};
            }
    }
    return true;
}

}  // namespace bfhip
