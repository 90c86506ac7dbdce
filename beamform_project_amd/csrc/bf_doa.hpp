// bf_doa.hpp -- DOA kernel launchers (internal).
#pragma once

#include "bf_common.hpp"

namespace bfhip {

// Exact smoothed integer covariance rs [F][M][M][2] (and optionally the full
// [F][A][A][2] covariance) of snapshots t0 .. t0+K-1.
void launch_doa_covariance(hipStream_t stream, const uint8_t* x, size_t T, size_t F, size_t A,
                           size_t t0, size_t K, size_t M, int* rs, int* r_full);

// FP64 fold, Cholesky, inverse and Bartlett/Capon scans per channel.
void launch_doa_fold_scan(hipStream_t stream, const int* rs, double* state, bool first,
                          size_t F, size_t M, size_t K, size_t L, double alpha, double load,
                          const float* freqs, double x_step, size_t n_theta,
                          const double* sin_theta, double* capon, double* bartlett,
                          int* chol_ok);

// Broadband spectra and the strongest local maxima with parabolic refinement.
void launch_doa_estimate(hipStream_t stream, const double* capon, const double* bartlett,
                         size_t F, size_t n_theta, const double* theta, size_t max_peaks,
                         double* broadband, double* peaks, int* n_peaks);

}  // namespace bfhip
