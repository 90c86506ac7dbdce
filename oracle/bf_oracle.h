/*
 * bf_oracle.h -- CPU oracle for the CHARTS beamformer / tracker / DOA path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded exclusively by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py.  The product library (libbeamform_hip.so)
 * never calls it and has no CPU fallback.
 *
 * Every function restates, in plain C, the algorithm of a reference C++
 * function (cited as file:line relative to the reference repository root).
 * Floating-point semantics follow the reference's compiled semantics:
 * float32 arithmetic in the same operation order, compiled with
 * -ffp-contract=off (the reference builds with -std=c++17, i.e. ISO mode, in
 * which GCC does not contract a*b+c into FMA), libm double cos/sin for the
 * phases, and IEEE division.
 *
 * Layout conventions (include/beamformer/indexing.hpp:13-66):
 *   packed voltage  [t][f][a]     uint8, low nibble = real, high nibble = imag
 *   weights         [b][f][a]     complex float32 interleaved (re, im)
 *   tiled weights   [f][b/32][a][32]
 *   intensity       [t][f][b]     float32
 *   integrated      [w][f][b]     float32
 *   int8 params     [wt][ft][bt]  {offset, scale} float32 per 1x16x16 chunk
 *
 * Return codes: 0 = ok, negative = the reference would throw
 * (-1 invalid_argument, -2 out_of_range).
 */
#ifndef BF_ORACLE_H
#define BF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_OK 0
#define OR_INVALID (-1)
#define OR_RANGE (-2)

/* ---- int4 codec: include/beamformer/int4.hpp:19-39 --------------------- */
int or_decode_nibble(unsigned nibble);
void or_unpack_bytes(const uint8_t* packed, size_t n, int8_t* re, int8_t* im);
int or_pack(int re, int im, uint8_t* out);

/* ---- contract: include/beamformer/config.hpp:81-94 ---------------------- */
int or_validate_dimensions(size_t n_time, size_t n_freq, size_t n_ant, size_t n_beams);

/* ---- geometry: src/geometry.cpp ----------------------------------------- */
int or_regular_array(size_t rows, size_t cols, float spacing, float* pos3);
int or_default_positions(size_t n_ant, float spacing, float* pos3);
int or_channelized_frequencies(size_t n, float start_hz, float width_hz, float* out);
int or_direction_from_lm(float l, float m, float* out3);
int or_default_beam_grid(size_t n_beams, float l_step, float m, float* out3);
int or_rectangular_beam_grid(size_t n_ant, float spacing, float design_hz, float* out3);

/* ---- weights: src/weights.cpp:11-122 ------------------------------------ */
int or_generate_weights(size_t n_beams, size_t n_freq, size_t n_ant,
                        const float* pos3, const float* freqs, const float* dirs3,
                        float* w);
int or_generate_tiled_weights(size_t n_beams, size_t n_freq, size_t n_ant,
                              const float* pos3, const float* freqs,
                              const float* dirs3, float* w);
/* FP64 weights (no float narrowing) for the FP64 oracle. */
int or_generate_weights_f64(size_t n_beams, size_t n_freq, size_t n_ant,
                            const float* pos3, const float* freqs,
                            const float* dirs3, double* w);

/* ---- fixed-grid beamformer: src/cpu_beamformer.cpp:40-79 ---------------- */
void or_beamform_packed(const uint8_t* x, const float* w, size_t T, size_t F,
                        size_t A, size_t B, float* out);
/* Same per-cell arithmetic, cells distributed over OpenMP threads
 * (labelled non-reference: the reference fixed grid is single-threaded). */
void or_beamform_packed_omp(const uint8_t* x, const float* w, size_t T, size_t F,
                            size_t A, size_t B, float* out);
/* FP64 oracle: sum_a W*X in double from the given float weights. */
void or_beamform_packed_f64(const uint8_t* x, const float* w, size_t T, size_t F,
                            size_t A, size_t B, double* out);
/* FP64 oracle from double weights. */
void or_beamform_packed_f64w(const uint8_t* x, const double* w, size_t T, size_t F,
                             size_t A, size_t B, double* out);

/* ---- integration: src/temporal_integration.cpp:16-110 ------------------- */
size_t or_integrated_time_count(size_t n_time, size_t spectra);
void or_integrate(const float* in, size_t T, size_t F, size_t B, size_t S, float* out);
void or_integrate_f64(const double* in, size_t T, size_t F, size_t B, size_t S,
                      double* out);

/* ---- int8 quantiser: src/quantization.cpp:11-193 ------------------------ */
size_t or_quant_param_count(size_t Tw, size_t F, size_t B);
void or_quantize(const float* in, size_t Tw, size_t F, size_t B, int8_t* codes,
                 float* params);
void or_dequantize(const int8_t* codes, const float* params, size_t Tw, size_t F,
                   size_t B, float* out);

/* ---- synthetic data: src/synthetic_data.cpp ----------------------------- */
void or_make_noise(size_t count, uint32_t seed, uint8_t* out);
int or_make_point_source(size_t T, size_t F, size_t A, const float* pos3,
                         const float* freqs, const float* dir3, float amplitude,
                         uint8_t* out);
/* Counter-hash generators shared with the product's device generators
 * (same integer algorithm, so the bytes are identical). */
uint64_t or_hash64(uint64_t key, uint64_t seed);
void or_make_uniform_int4(size_t count, uint64_t seed, uint8_t* out);

/* ---- dynamic tracker: src/beam_tracker.cpp:158-260 ---------------------- */
size_t or_tracker_window_count(size_t n_time, size_t spectra);
int or_tracker_direction(float l0, float m0, float rate_l, float rate_m, size_t t,
                         float* out3);
int or_beam_tracker(const uint8_t* x, size_t T, size_t F, size_t A, float l0,
                    float m0, float rate_l, float rate_m, size_t S, float* out);
/* src/beam_tracker_opt_v2.cpp:143-311: same cells, OpenMP two-pass. */
int or_beam_tracker_omp(const uint8_t* x, size_t T, size_t F, size_t A, float l0,
                        float m0, float rate_l, float rate_m, size_t S, float* out);
/* src/beam_tracker.cpp:262-306 */
int or_make_moving_point_source(size_t T, size_t F, size_t A, const float* pos3,
                                const float* freqs, float l0, float m0,
                                float rate_l, float rate_m, float amplitude,
                                uint8_t* out);

/* Multi-source tracker (the build's multi-beam extension of the tracker):
 * per window w and source b the direction dirs3[(w*B+b)*3], weights as
 * src/weights.cpp:42-55, contraction as src/cpu_beamformer.cpp:56-78, then
 * window integration as src/temporal_integration.cpp:95-109 (S spectra per
 * window, partial last window).  Output [W][F][B] float.  When out64 is not
 * NULL the FP64 oracle (double weights, double sums) is written there too. */
int or_multibeam_tracker(const uint8_t* x, size_t T, size_t F, size_t A, size_t B,
                         size_t S, const float* pos3, const float* freqs,
                         const float* dirs3, float* out, double* out64);
/* Per-spectrum variant (S only selects the direction window), output [T][F][B]. */
int or_multibeam_tracker_spectra(const uint8_t* x, size_t T, size_t F, size_t A,
                                 size_t B, size_t S, const float* pos3,
                                 const float* freqs, const float* dirs3,
                                 float* out, double* out64);

/* ---- adaptive DOA primitives: src/cpu_opt_beam_tracker.cpp -------------- */
/* Exact integer sample covariance R[r][c] = sum_k x_r conj(x_c) over the K
 * snapshots t0..t0+K-1 of channel f (src/cpu_opt_beam_tracker.cpp:559-605
 * without the 1/K scale).  out: [A][A][2] int64 (re, im). */
void or_covariance_i64(const uint8_t* x, size_t T, size_t F, size_t A, size_t t0,
                       size_t K, size_t f, int64_t* out);
/* Forward spatial smoothing as a sum of the L = A-M+1 diagonal M x M
 * sub-blocks (src/cpu_opt_beam_tracker.cpp:607-636, forward part only). */
void or_smooth_forward_i64(const int64_t* R, size_t A, size_t M, int64_t* out);
/* Diagonal load added to R before the fold, eps = load*tr(R)/M
 * (src/cpu_opt_beam_tracker.cpp:244-253, 1874-1890). */
void or_diagonal_load_f64(double* R, size_t M, double load);
/* Hermitian Cholesky R = L L^H, lower triangle, row-major complex
 * (src/cpu_opt_beam_tracker.cpp:398-421, in double).  Returns 1 on success. */
int or_cholesky_f64(double* R, size_t M);
/* Capon power 1/||L^{-1} a||^2 (src/cpu_opt_beam_tracker.cpp:423-437). */
double or_capon_from_factor_f64(const double* L, const double* a, size_t M);
/* Bartlett power Re(a^H R a) (src/cpu_opt_beam_tracker.cpp:255-306). */
double or_bartlett_f64(const double* R, const double* a, size_t M);
/* 3-point parabolic peak offset in cells, clamped to +-0.5.  The reference
 * (src/cpu_opt_beam_tracker.cpp:830-853) writes the denominator with the
 * opposite sign, so its guard only passes at minima; this restates the
 * textbook vertex of the parabola through the three samples. */
double or_quadratic_peak_offset(double p_minus, double p_centre, double p_plus);

/* Peak list of one broadband spectrum P[0..n): the local maxima
 * (P[q] > P[q-1] and P[q] >= P[q+1], 0 < q < n-1), each moved to the vertex
 * of its parabola (or_quadratic_peak_offset times theta[q+1] - theta[q]),
 * the max_peaks strongest, strongest first and the earlier angle first on
 * equal power.  out[2i] = angle, out[2i+1] = power; returns the count. */
size_t or_doa_peaks(const double* P, const double* theta, size_t n, size_t max_peaks,
                    double* out);

/* Full C5 update over all channels: covariance -> smoothing -> /(K*L) ->
 * fold with alpha (first update copies) -> Bartlett scan of the folded R;
 * Capon from the Cholesky factor of R + e I, where e is the fold of the
 * per-update diagonal loads eps = load * tr(R_block) / M (the reference
 * loads R_block before its fold for Capon only,
 * src/cpu_opt_beam_tracker.cpp:1879-1892).  Scans cover n_theta angles of a
 * ULA with element x positions xpos (metres), steering exp(-j k x sin(theta))
 * (src/cpu_opt_beam_tracker.cpp:96-106).  R_state: [F][M][M][2] double
 * (unloaded) and load_state: [F] double (e), carried between calls;
 * first != 0 marks the first update.
 * Outputs: capon[F][n_theta], bartlett[F][n_theta] (double). */
int or_doa_update(const uint8_t* x, size_t T, size_t F, size_t A, size_t t0,
                  size_t K, size_t M, double alpha, double load, int first,
                  const float* freqs, const float* xpos, size_t n_theta,
                  const double* theta, double* R_state, double* load_state,
                  double* capon, double* bartlett, int* chol_ok);

/* ---- accuracy metrics (BASELINE.md measurement definitions) ------------- */
/* max_i |g_i - r_i| / max(|r_i|, floor) */
double or_max_rel_err_floored(const float* got, const double* ref, size_t n, double floor);

/* ---- mt19937 (std::mt19937, used by make_noise) ------------------------- */
typedef struct {
    uint32_t mt[624];
    int index;
} or_mt19937;
void or_mt_seed(or_mt19937* g, uint32_t seed);
uint32_t or_mt_next(or_mt19937* g);

#ifdef __cplusplus
}
#endif

#endif /* BF_ORACLE_H */
