// bf_doa.hpp -- DOA kernel launchers (internal).
#pragma once

#include "bf_common.hpp"

namespace bfhip {

// Exact smoothed integer covariance of snapshots t0 .. t0+K-1 as `parts`
// partial sums rs [parts][F][M][M][2] over disjoint snapshot ranges (their
// sum is the covariance), parts = doa_covariance_parts(K, max_parts), one
// with r_full, which then also receives the full [F][A][A][2] covariance.
// Returns parts.
size_t doa_covariance_parts(size_t K, size_t max_parts);
size_t launch_doa_covariance(hipStream_t stream, const uint8_t* x, size_t T, size_t F,
                             size_t A, size_t t0, size_t K, size_t M, int* rs, int* r_full,
                             size_t max_parts = 1);

// FP64 fold, Cholesky, inverse and Bartlett/Capon scans per channel; rs
// holds `parts` partial covariances, summed here.
void launch_doa_fold_scan(hipStream_t stream, const int* rs, size_t parts, double* state,
                          bool first,
                          size_t F, size_t M, size_t K, size_t L, double alpha, double load,
                          const float* freqs, double x_step, size_t n_theta,
                          const double* sin_theta, double* capon, double* bartlett,
                          int* chol_ok);

// Broadband spectra and the strongest local maxima with parabolic refinement.
void launch_doa_estimate(hipStream_t stream, const double* capon, const double* bartlett,
                         size_t F, size_t n_theta, const double* theta, size_t max_peaks,
                         double* broadband, double* peaks, int* n_peaks);

// ----------------------------------------------- planar direction search
// Device state of one search: the centre of the window's coarse grid (the
// previous window's estimate) and the running centre of the refinement.
struct OptSearch {
    float prev[3];
    float centre[3];
};

struct OptGridArgs {
    int G;         // coarse cells per axis
    int M;         // sub-array size (steered antennas 0 .. M-1)
    float fov_l;   // half-extent of the coarse grid
    float fov_m;
    int coplanar;  // steer at nominal lattice points (every |z| <= 1e-6)
    int interp;    // quadratic peak interpolation
};

// Per channel: R_block = Rs / (K L) (forward-backward averaged when fb),
// diagonal load (Capon), forgetting fold into state [F][M][M][2] FP64, and
// the scan operand: Bartlett, the strict upper triangle of R in scanmat
// [F][M][M] with its trace in trace[F] (mode 0); Capon, R^-1 in scan64
// [F][M][M] with its trace in trace64[F] (mode 1), or the Bartlett form
// where the factorisation fails.
void launch_opt_fold(hipStream_t stream, const int* rs, double* state, bool first, size_t F,
                     size_t M, size_t K, size_t L, bool fb, bool capon, float load,
                     float lambda, float2* scanmat, float* trace, double2* scan64,
                     double* trace64, int* mode);

// part[n_cells][F]: each channel's power at the cells of `level` (0: the
// G x G grid around search->prev, else the 3 x 3 refinement candidates).
void launch_opt_scan(hipStream_t stream, const float2* scanmat, const float* trace,
                     const double2* scan64, const double* trace64, const int* mode,
                     const float* pos, const float* freqs, const OptSearch* search,
                     const OptGridArgs& g, int level, size_t F, size_t n_cells, bool capon,
                     float* part);

// Sums part over channels and takes the level's decision; at last_level
// writes the window's direction to dirs[window][3].
void launch_opt_decide(hipStream_t stream, const float* part, size_t F, size_t n_cells,
                       const OptGridArgs& g, int level, int last_level, size_t window,
                       OptSearch* search, float* dirs);

}  // namespace bfhip
