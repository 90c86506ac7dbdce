/*
 * bf_oracle.c -- CPU oracle (test infrastructure only; see bf_oracle.h).
 *
 * Build: oracle/Makefile (gcc -O2 -std=c11 -ffp-contract=off -fopenmp).
 * -ffp-contract=off is load-bearing: the reference compiles its CPU core in
 * ISO C++17 mode where GCC keeps a*b+c as two rounded operations, and the
 * bit-exact fixtures depend on that.
 */
#include "bf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static const double kTwoPi = 6.283185307179586476925286766559;   /* include/beamformer/physics.hpp:6 */
static const double kSpeedOfLight = 299792458.0;                /* include/beamformer/physics.hpp:5 */

/* ------------------------------------------------------------------ int4 */
/* include/beamformer/int4.hpp:19-24 */
int or_decode_nibble(unsigned nibble) {
    const unsigned bits = nibble & 0x0Fu;
    return bits < 8u ? (int)bits : (int)bits - 16;
}

/* include/beamformer/int4.hpp:26-30 */
void or_unpack_bytes(const uint8_t* packed, size_t n, int8_t* re, int8_t* im) {
    for (size_t i = 0; i < n; ++i) {
        re[i] = (int8_t)or_decode_nibble(packed[i]);
        im[i] = (int8_t)or_decode_nibble((unsigned)(packed[i] >> 4));
    }
}

/* include/beamformer/int4.hpp:32-39 */
int or_pack(int re, int im, uint8_t* out) {
    if (re < -8 || re > 7 || im < -8 || im > 7) return OR_RANGE;
    *out = (uint8_t)(((unsigned)re & 0x0Fu) | (((unsigned)im & 0x0Fu) << 4));
    return OR_OK;
}

/* include/beamformer/config.hpp:81-94 */
int or_validate_dimensions(size_t n_time, size_t n_freq, size_t n_ant, size_t n_beams) {
    if (n_time == 0) return OR_INVALID;
    if (n_freq != 336) return OR_INVALID;
    if (n_ant != 32 && n_ant != 64 && n_ant != 128 && n_ant != 256) return OR_INVALID;
    if (n_beams == 0 || n_beams > 128) return OR_INVALID;
    return OR_OK;
}

/* -------------------------------------------------------------- geometry */
/* src/geometry.cpp:33-51: column-major x, row y, z = 0. */
int or_regular_array(size_t rows, size_t cols, float spacing, float* pos3) {
    if (rows == 0 || cols == 0 || !(spacing > 0.0f)) return OR_INVALID;
    size_t i = 0;
    for (size_t r = 0; r < rows; ++r) {
        for (size_t c = 0; c < cols; ++c, ++i) {
            pos3[3 * i + 0] = (float)c * spacing;
            pos3[3 * i + 1] = (float)r * spacing;
            pos3[3 * i + 2] = 0.0f;
        }
    }
    return OR_OK;
}

/* src/geometry.cpp:53-67 */
int or_default_positions(size_t n_ant, float spacing, float* pos3) {
    switch (n_ant) {
    case 32: return or_regular_array(4, 8, spacing, pos3);
    case 64: return or_regular_array(8, 8, spacing, pos3);
    case 128: return or_regular_array(8, 16, spacing, pos3);
    case 256: return or_regular_array(16, 16, spacing, pos3);
    default: return OR_INVALID;
    }
}

/* src/geometry.cpp:79-95: f_k = start + (float)k * width, float arithmetic. */
int or_channelized_frequencies(size_t n, float start_hz, float width_hz, float* out) {
    if (n == 0 || !isfinite(start_hz) || !(start_hz > 0.0f) || !isfinite(width_hz)
        || !(width_hz > 0.0f))
        return OR_INVALID;
    for (size_t k = 0; k < n; ++k) {
        const float step = (float)k * width_hz;
        out[k] = start_hz + step;
    }
    return OR_OK;
}

/* src/geometry.cpp:97-106 */
int or_direction_from_lm(float l, float m, float* out3) {
    if (!isfinite(l) || !isfinite(m)) return OR_INVALID;
    const float ll = l * l;
    const float mm = m * m;
    const float t2 = ll + mm;
    if (t2 > 1.0f) return OR_INVALID;
    out3[0] = l;
    out3[1] = m;
    out3[2] = sqrtf(1.0f - t2);
    return OR_OK;
}

/* src/geometry.cpp:108-125 */
int or_default_beam_grid(size_t n_beams, float l_step, float m, float* out3) {
    if (n_beams == 0 || n_beams > 128) return OR_INVALID;
    if (!isfinite(l_step) || !(l_step > 0.0f)) return OR_INVALID;
    const long long centre = (long long)(n_beams / 2);
    for (size_t b = 0; b < n_beams; ++b) {
        const long long offset = (long long)b - centre;
        const int rc = or_direction_from_lm((float)offset * l_step, m, out3 + 3 * b);
        if (rc) return rc;
    }
    return OR_OK;
}

/* src/geometry.cpp:332-366 (double-precision grid, narrowed per cell). */
int or_rectangular_beam_grid(size_t n_ant, float spacing, float design_hz, float* out3) {
    if (!(spacing > 0.0f) || !isfinite(spacing)) return OR_INVALID;
    if (!(design_hz > 0.0f) || !isfinite(design_hz)) return OR_INVALID;
    const size_t rows = n_ant == 32 ? 4 : n_ant == 64 ? 8 : 0;
    const size_t cols = (n_ant == 32 || n_ant == 64) ? 8 : 0;
    if (rows == 0) return OR_INVALID;
    const double wavelength = kSpeedOfLight / (double)design_hz;
    const double dl = wavelength / ((double)cols * spacing);
    const double dm = wavelength / ((double)rows * spacing);
    size_t i = 0;
    for (size_t r = 0; r < rows; ++r) {
        const double m = ((double)r - (double)(rows - 1) / 2.0) * dm;
        for (size_t c = 0; c < cols; ++c, ++i) {
            const double l = ((double)c - (double)(cols - 1) / 2.0) * dl;
            const int rc = or_direction_from_lm((float)l, (float)m, out3 + 3 * i);
            if (rc) return rc;
        }
    }
    return OR_OK;
}

/* --------------------------------------------------------------- weights */
static int check_unit_direction(const float* d) {
    const double n2 = (double)d[0] * d[0] + (double)d[1] * d[1] + (double)d[2] * d[2];
    if (!isfinite(n2) || fabs(n2 - 1.0) > 1.0e-3) return OR_INVALID;
    return OR_OK;
}

static double phase_of(const float* p, const float* d, double k) {
    /* src/weights.cpp:42-50: delay in double from float position/direction. */
    const double delay = (double)p[0] * d[0] + (double)p[1] * d[1] + (double)p[2] * d[2];
    return k * delay;
}

static double wave_number(float f_hz) {
    /* src/weights.cpp:37-40 */
    return kTwoPi * (double)f_hz / kSpeedOfLight;
}

/* src/weights.cpp:11-60 */
int or_generate_weights(size_t B, size_t F, size_t A, const float* pos3,
                        const float* freqs, const float* dirs3, float* w) {
    for (size_t b = 0; b < B; ++b) {
        if (check_unit_direction(dirs3 + 3 * b)) return OR_INVALID;
        for (size_t f = 0; f < F; ++f) {
            if (!isfinite(freqs[f]) || !(freqs[f] > 0.0f)) return OR_INVALID;
            const double k = wave_number(freqs[f]);
            for (size_t a = 0; a < A; ++a) {
                const double ph = phase_of(pos3 + 3 * a, dirs3 + 3 * b, k);
                const size_t i = (b * F + f) * A + a;
                w[2 * i + 0] = (float)cos(ph);
                w[2 * i + 1] = (float)sin(ph);
            }
        }
    }
    return OR_OK;
}

int or_generate_weights_f64(size_t B, size_t F, size_t A, const float* pos3,
                            const float* freqs, const float* dirs3, double* w) {
    for (size_t b = 0; b < B; ++b) {
        for (size_t f = 0; f < F; ++f) {
            const double k = wave_number(freqs[f]);
            for (size_t a = 0; a < A; ++a) {
                const double ph = phase_of(pos3 + 3 * a, dirs3 + 3 * b, k);
                const size_t i = (b * F + f) * A + a;
                w[2 * i + 0] = cos(ph);
                w[2 * i + 1] = sin(ph);
            }
        }
    }
    return OR_OK;
}

/* src/weights.cpp:63-122, layout include/beamformer/indexing.hpp:54-61 */
int or_generate_tiled_weights(size_t B, size_t F, size_t A, const float* pos3,
                              const float* freqs, const float* dirs3, float* w) {
    const size_t tiles = (B + 31) / 32;
    memset(w, 0, sizeof(float) * 2 * F * tiles * A * 32);
    for (size_t f = 0; f < F; ++f) {
        if (!isfinite(freqs[f]) || !(freqs[f] > 0.0f)) return OR_INVALID;
        const double k = wave_number(freqs[f]);
        for (size_t bt = 0; bt < tiles; ++bt)
            for (size_t a = 0; a < A; ++a)
                for (size_t lb = 0; lb < 32; ++lb) {
                    const size_t b = bt * 32 + lb;
                    if (b >= B) continue;
                    if (check_unit_direction(dirs3 + 3 * b)) return OR_INVALID;
                    const double ph = phase_of(pos3 + 3 * a, dirs3 + 3 * b, k);
                    const size_t i = ((f * tiles + bt) * A + a) * 32 + lb;
                    w[2 * i + 0] = (float)cos(ph);
                    w[2 * i + 1] = (float)sin(ph);
                }
    }
    return OR_OK;
}

/* ----------------------------------------------------------- beamformer */
/* One output cell of src/cpu_beamformer.cpp:59-75: float MAC in antenna
 * order, each product and sum rounded separately (no FMA). */
static inline float cell_power(const uint8_t* xrow, const float* wrow, size_t A) {
    float sr = 0.0f, si = 0.0f;
    for (size_t a = 0; a < A; ++a) {
        const float xr = (float)or_decode_nibble(xrow[a]);
        const float xi = (float)or_decode_nibble((unsigned)(xrow[a] >> 4));
        const float wr = wrow[2 * a + 0];
        const float wi = wrow[2 * a + 1];
        const float p1 = wr * xr;
        const float p2 = wi * xi;
        const float re = p1 - p2;
        const float p3 = wr * xi;
        const float p4 = wi * xr;
        const float im = p3 + p4;
        sr = sr + re;
        si = si + im;
    }
    const float a2 = sr * sr;
    const float b2 = si * si;
    return a2 + b2;
}

/* src/cpu_beamformer.cpp:40-79 */
void or_beamform_packed(const uint8_t* x, const float* w, size_t T, size_t F,
                        size_t A, size_t B, float* out) {
    for (size_t t = 0; t < T; ++t)
        for (size_t f = 0; f < F; ++f)
            for (size_t b = 0; b < B; ++b)
                out[(t * F + f) * B + b] =
                    cell_power(x + (t * F + f) * A, w + 2 * ((b * F + f) * A), A);
}

void or_beamform_packed_omp(const uint8_t* x, const float* w, size_t T, size_t F,
                            size_t A, size_t B, float* out) {
    const long long cells = (long long)(T * F);
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < cells; ++c) {
        const size_t t = (size_t)c / F, f = (size_t)c % F;
        for (size_t b = 0; b < B; ++b)
            out[(t * F + f) * B + b] =
                cell_power(x + (t * F + f) * A, w + 2 * ((b * F + f) * A), A);
    }
}

void or_beamform_packed_f64(const uint8_t* x, const float* w, size_t T, size_t F,
                            size_t A, size_t B, double* out) {
    const long long cells = (long long)(T * F);
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < cells; ++c) {
        const size_t t = (size_t)c / F, f = (size_t)c % F;
        const uint8_t* xr = x + (t * F + f) * A;
        for (size_t b = 0; b < B; ++b) {
            const float* wr = w + 2 * ((b * F + f) * A);
            double sr = 0.0, si = 0.0;
            for (size_t a = 0; a < A; ++a) {
                const double vr = or_decode_nibble(xr[a]);
                const double vi = or_decode_nibble((unsigned)(xr[a] >> 4));
                sr += (double)wr[2 * a] * vr - (double)wr[2 * a + 1] * vi;
                si += (double)wr[2 * a] * vi + (double)wr[2 * a + 1] * vr;
            }
            out[(t * F + f) * B + b] = sr * sr + si * si;
        }
    }
}

void or_beamform_packed_f64w(const uint8_t* x, const double* w, size_t T, size_t F,
                             size_t A, size_t B, double* out) {
    const long long cells = (long long)(T * F);
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < cells; ++c) {
        const size_t t = (size_t)c / F, f = (size_t)c % F;
        const uint8_t* xr = x + (t * F + f) * A;
        for (size_t b = 0; b < B; ++b) {
            const double* wr = w + 2 * ((b * F + f) * A);
            double sr = 0.0, si = 0.0;
            for (size_t a = 0; a < A; ++a) {
                const double vr = or_decode_nibble(xr[a]);
                const double vi = or_decode_nibble((unsigned)(xr[a] >> 4));
                sr += wr[2 * a] * vr - wr[2 * a + 1] * vi;
                si += wr[2 * a] * vi + wr[2 * a + 1] * vr;
            }
            out[(t * F + f) * B + b] = sr * sr + si * si;
        }
    }
}

/* ---------------------------------------------------------- integration */
/* src/temporal_integration.cpp:16-26 */
size_t or_integrated_time_count(size_t n_time, size_t spectra) {
    if (spectra == 0 || n_time == 0) return 0;
    return n_time / spectra + (n_time % spectra == 0 ? 0u : 1u);
}

/* src/temporal_integration.cpp:79-110: float sum in increasing time. */
void or_integrate(const float* in, size_t T, size_t F, size_t B, size_t S, float* out) {
    const size_t W = or_integrated_time_count(T, S);
    for (size_t w = 0; w < W; ++w) {
        const size_t t0 = w * S;
        const size_t t1 = (t0 + S < T) ? t0 + S : T;
        for (size_t f = 0; f < F; ++f)
            for (size_t b = 0; b < B; ++b) {
                float s = 0.0f;
                for (size_t t = t0; t < t1; ++t) s = s + in[(t * F + f) * B + b];
                out[(w * F + f) * B + b] = s;
            }
    }
}

void or_integrate_f64(const double* in, size_t T, size_t F, size_t B, size_t S,
                      double* out) {
    const size_t W = or_integrated_time_count(T, S);
    for (size_t w = 0; w < W; ++w) {
        const size_t t0 = w * S;
        const size_t t1 = (t0 + S < T) ? t0 + S : T;
        for (size_t f = 0; f < F; ++f)
            for (size_t b = 0; b < B; ++b) {
                double s = 0.0;
                for (size_t t = t0; t < t1; ++t) s += in[(t * F + f) * B + b];
                out[(w * F + f) * B + b] = s;
            }
    }
}

/* ------------------------------------------------------------ quantiser */
/* src/quantization.cpp:34-59 */
size_t or_quant_param_count(size_t Tw, size_t F, size_t B) {
    return Tw * ((F + 15) / 16) * ((B + 15) / 16);
}

/* src/quantization.cpp:20-30 (lrint under the default round-to-nearest-even). */
static int quantize_value(float v, float offset, float scale) {
    if (!isfinite(v)) return -128;
    if (scale == 0.0f) return 0;
    const float d = v - offset;
    const float q = d / scale;
    long r = lrintf(q);
    if (r < -127) r = -127;
    if (r > 127) r = 127;
    return (int)r;
}

/* src/quantization.cpp:61-161 */
void or_quantize(const float* in, size_t Tw, size_t F, size_t B, int8_t* codes,
                 float* params) {
    const size_t ft = (F + 15) / 16, bt = (B + 15) / 16;
    const float out_min = -127.0f - 0.5f;
    const float out_max = 127.0f + 0.5f;
    const float out_range = out_max - out_min;
    for (size_t t = 0; t < Tw; ++t)
        for (size_t fi = 0; fi < ft; ++fi)
            for (size_t bi = 0; bi < bt; ++bi) {
                float mn = INFINITY, mx = -INFINITY;
                int nonfinite = 0;
                for (size_t lf = 0; lf < 16; ++lf) {
                    const size_t f = fi * 16 + lf;
                    if (f >= F) continue;
                    for (size_t lb = 0; lb < 16; ++lb) {
                        const size_t b = bi * 16 + lb;
                        if (b >= B) continue;
                        const float v = in[(t * F + f) * B + b];
                        if (!isfinite(v)) {
                            nonfinite = 1;
                        } else {
                            mn = v < mn ? v : mn; /* std::min(minimum, value) */
                            mx = mx < v ? v : mx; /* std::max(maximum, value) */
                        }
                    }
                }
                float offset = 0.0f, scale = 0.0f;
                if (!nonfinite && isfinite(mn) && isfinite(mx)) {
                    const float span = mx - mn;
                    scale = span / out_range;
                    const float shift = out_min * scale;
                    offset = mn - shift;
                    if (!isfinite(offset) || !isfinite(scale)) {
                        offset = 0.0f;
                        scale = 0.0f;
                    }
                }
                const size_t p = (t * ft + fi) * bt + bi;
                params[2 * p + 0] = offset;
                params[2 * p + 1] = scale;
                for (size_t lf = 0; lf < 16; ++lf) {
                    const size_t f = fi * 16 + lf;
                    if (f >= F) continue;
                    for (size_t lb = 0; lb < 16; ++lb) {
                        const size_t b = bi * 16 + lb;
                        if (b >= B) continue;
                        const size_t i = (t * F + f) * B + b;
                        codes[i] = (int8_t)quantize_value(in[i], offset, scale);
                    }
                }
            }
}

/* src/quantization.cpp:163-193 */
void or_dequantize(const int8_t* codes, const float* params, size_t Tw, size_t F,
                   size_t B, float* out) {
    const size_t ft = (F + 15) / 16, bt = (B + 15) / 16;
    for (size_t t = 0; t < Tw; ++t)
        for (size_t f = 0; f < F; ++f)
            for (size_t b = 0; b < B; ++b) {
                const size_t i = (t * F + f) * B + b;
                if (codes[i] == -128) {
                    out[i] = NAN;
                    continue;
                }
                const size_t p = (t * ft + f / 16) * bt + b / 16;
                const float prod = params[2 * p + 1] * (float)codes[i];
                out[i] = params[2 * p + 0] + prod;
            }
}

/* --------------------------------------------------------------- mt19937 */
/* The standard 32-bit Mersenne Twister (std::mt19937), restated from its
 * published definition (Matsumoto & Nishimura 1998; C++ [rand.eng.mers]). */
void or_mt_seed(or_mt19937* g, uint32_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->index = 624;
}

uint32_t or_mt_next(or_mt19937* g) {
    if (g->index >= 624) {
        for (int i = 0; i < 624; ++i) {
            const uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            g->mt[i] = v;
        }
        g->index = 0;
    }
    uint32_t y = g->mt[g->index++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* ------------------------------------------------------------ synthetic */
/* src/synthetic_data.cpp:107-119 (values in [-7, 7]; -8 never occurs). */
void or_make_noise(size_t count, uint32_t seed, uint8_t* out) {
    or_mt19937 g;
    or_mt_seed(&g, seed);
    for (size_t i = 0; i < count; ++i) {
        const uint32_t bits = or_mt_next(&g);
        const int re = (int)(bits & 0xFFFFu) % 15 - 7;
        const int im = (int)((bits >> 16) & 0xFFFFu) % 15 - 7;
        or_pack(re, im, out + i);
    }
}

/* src/synthetic_data.cpp:71-105 */
int or_make_point_source(size_t T, size_t F, size_t A, const float* pos3,
                         const float* freqs, const float* dir3, float amplitude,
                         uint8_t* out) {
    if (!isfinite(amplitude) || !(amplitude > 0.0f) || amplitude > 7.0f) return OR_INVALID;
    if (check_unit_direction(dir3)) return OR_INVALID;
    uint8_t* spec = (uint8_t*)malloc(F * A);
    for (size_t f = 0; f < F; ++f) {
        const double k = wave_number(freqs[f]);
        for (size_t a = 0; a < A; ++a) {
            const double ph = phase_of(pos3 + 3 * a, dir3, k);
            const int re = (int)lround((double)amplitude * cos(ph));
            const int im = (int)lround(-(double)amplitude * sin(ph));
            or_pack(re, im, spec + f * A + a);
        }
    }
    for (size_t t = 0; t < T; ++t) memcpy(out + t * F * A, spec, F * A);
    free(spec);
    return OR_OK;
}

/* splitmix64 finaliser over (key, seed): the counter hash used by the
 * product's device generators for seeded synthetic int4 data. */
uint64_t or_hash64(uint64_t key, uint64_t seed) {
    uint64_t z = key + (seed + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Uniform [-8, 7] for both components: one hash byte per sample. */
void or_make_uniform_int4(size_t count, uint64_t seed, uint8_t* out) {
    for (size_t i = 0; i < count; ++i) out[i] = (uint8_t)(or_hash64(i, seed) & 0xFFu);
}

/* --------------------------------------------------------------- tracker */
/* src/beam_tracker.cpp:158-164 */
size_t or_tracker_window_count(size_t n_time, size_t spectra) {
    if (spectra == 0) return 0;
    return (n_time + spectra - 1) / spectra;
}

/* src/beam_tracker.cpp:166-179: l = l0 + (float)t * rate in float. */
int or_tracker_direction(float l0, float m0, float rate_l, float rate_m, size_t t,
                         float* out3) {
    const float tf = (float)t;
    const float dl = tf * rate_l;
    const float dm = tf * rate_m;
    return or_direction_from_lm(l0 + dl, m0 + dm, out3);
}

static int tracker_directions(size_t T, size_t S, float l0, float m0, float rl,
                              float rm, float* dirs) {
    const size_t W = or_tracker_window_count(T, S);
    for (size_t w = 0; w < W; ++w) {
        /* src/beam_tracker.cpp:181-190: direction at the window's first sample */
        const int rc = or_tracker_direction(l0, m0, rl, rm, w * S, dirs + 3 * w);
        if (rc) return rc;
    }
    return OR_OK;
}

/* src/beam_tracker.cpp:204-260 with the hard-coded geometry of :218-219
 * (default_positions(A) at 0.6 m, channelized_frequencies at 300 MHz +
 * k * 300 kHz). */
int or_beam_tracker(const uint8_t* x, size_t T, size_t F, size_t A, float l0, float m0,
                    float rl, float rm, size_t S, float* out) {
    if (or_validate_dimensions(T, F, A, 1) || S == 0) return OR_INVALID;
    float pos[256 * 3];
    float* freqs = (float*)malloc(sizeof(float) * F);
    or_default_positions(A, 0.6f, pos);
    or_channelized_frequencies(F, 300000000.0f, 300000.0f, freqs);
    const size_t W = or_tracker_window_count(T, S);
    float* dirs = (float*)malloc(sizeof(float) * 3 * W);
    int rc = tracker_directions(T, S, l0, m0, rl, rm, dirs);
    float* wgt = (float*)malloc(sizeof(float) * 2 * F * A);
    for (size_t w = 0; rc == OR_OK && w < W; ++w) {
        or_generate_weights(1, F, A, pos, freqs, dirs + 3 * w, wgt);
        const size_t t1 = (w * S + S < T) ? w * S + S : T;
        for (size_t t = w * S; t < t1; ++t)
            for (size_t f = 0; f < F; ++f)
                out[t * F + f] = cell_power(x + (t * F + f) * A, wgt + 2 * f * A, A);
    }
    free(wgt);
    free(dirs);
    free(freqs);
    return rc;
}

/* src/beam_tracker_opt_v2.cpp:143-311: pass 1 weights per (w, f) in
 * parallel, pass 2 cells per (w, t, f) in parallel; per-cell arithmetic is
 * the naive tracker's, so the output is bit-identical to or_beam_tracker. */
int or_beam_tracker_omp(const uint8_t* x, size_t T, size_t F, size_t A, float l0,
                        float m0, float rl, float rm, size_t S, float* out) {
    if (or_validate_dimensions(T, F, A, 1) || S == 0) return OR_INVALID;
    float pos[256 * 3];
    float* freqs = (float*)malloc(sizeof(float) * F);
    or_default_positions(A, 0.6f, pos);
    or_channelized_frequencies(F, 300000000.0f, 300000.0f, freqs);
    const size_t W = or_tracker_window_count(T, S);
    float* dirs = (float*)malloc(sizeof(float) * 3 * W);
    const int rc = tracker_directions(T, S, l0, m0, rl, rm, dirs);
    if (rc) {
        free(dirs);
        free(freqs);
        return rc;
    }
    double* kf = (double*)malloc(sizeof(double) * F);
    for (size_t f = 0; f < F; ++f) kf[f] = wave_number(freqs[f]);
    float* wgt = (float*)malloc(sizeof(float) * 2 * W * F * A);
#pragma omp parallel
    {
#pragma omp for collapse(2) schedule(static)
        for (long long w = 0; w < (long long)W; ++w)
            for (long long f = 0; f < (long long)F; ++f) {
                const size_t base = ((size_t)w * F + (size_t)f) * A;
                for (size_t a = 0; a < A; ++a) {
                    const double ph = phase_of(pos + 3 * a, dirs + 3 * w, kf[f]);
                    wgt[2 * (base + a) + 0] = (float)cos(ph);
                    wgt[2 * (base + a) + 1] = (float)sin(ph);
                }
            }
#pragma omp for collapse(3) schedule(static)
        for (long long w = 0; w < (long long)W; ++w)
            for (long long t = 0; t < (long long)S; ++t)
                for (long long f = 0; f < (long long)F; ++f) {
                    const size_t time = (size_t)w * S + (size_t)t;
                    if (time >= T) continue;
                    out[time * F + (size_t)f] =
                        cell_power(x + (time * F + (size_t)f) * A,
                                   wgt + 2 * (((size_t)w * F + (size_t)f) * A), A);
                }
    }
    free(wgt);
    free(kf);
    free(dirs);
    free(freqs);
    return OR_OK;
}

/* src/beam_tracker.cpp:262-306 */
int or_make_moving_point_source(size_t T, size_t F, size_t A, const float* pos3,
                                const float* freqs, float l0, float m0, float rl,
                                float rm, float amplitude, uint8_t* out) {
    if (!isfinite(amplitude) || !(amplitude > 0.0f) || amplitude > 7.0f) return OR_INVALID;
    for (size_t t = 0; t < T; ++t) {
        float d[3];
        const int rc = or_tracker_direction(l0, m0, rl, rm, t, d);
        if (rc) return rc;
        for (size_t f = 0; f < F; ++f) {
            const double k = wave_number(freqs[f]);
            for (size_t a = 0; a < A; ++a) {
                const double ph = phase_of(pos3 + 3 * a, d, k);
                const int re = (int)lround((double)amplitude * cos(ph));
                const int im = (int)lround(-(double)amplitude * sin(ph));
                or_pack(re, im, out + (t * F + f) * A + a);
            }
        }
    }
    return OR_OK;
}

static void window_weights(size_t B, size_t F, size_t A, const float* pos3,
                           const float* freqs, const float* dirs3, float* wgt,
                           double* wgt64) {
    for (size_t b = 0; b < B; ++b)
        for (size_t f = 0; f < F; ++f) {
            const double k = wave_number(freqs[f]);
            for (size_t a = 0; a < A; ++a) {
                const double ph = phase_of(pos3 + 3 * a, dirs3 + 3 * b, k);
                const size_t i = (b * F + f) * A + a;
                wgt[2 * i] = (float)cos(ph);
                wgt[2 * i + 1] = (float)sin(ph);
                if (wgt64) {
                    wgt64[2 * i] = cos(ph);
                    wgt64[2 * i + 1] = sin(ph);
                }
            }
        }
}

int or_multibeam_tracker(const uint8_t* x, size_t T, size_t F, size_t A, size_t B,
                         size_t S, const float* pos3, const float* freqs,
                         const float* dirs3, float* out, double* out64) {
    if (S == 0 || B == 0 || A == 0 || F == 0) return OR_INVALID;
    const size_t W = or_tracker_window_count(T, S);
    float* wgt = (float*)malloc(sizeof(float) * 2 * B * F * A);
    double* wgt64 = out64 ? (double*)malloc(sizeof(double) * 2 * B * F * A) : NULL;
    for (size_t w = 0; w < W; ++w) {
        window_weights(B, F, A, pos3, freqs, dirs3 + 3 * B * w, wgt, wgt64);
        const size_t t0 = w * S, t1 = (t0 + S < T) ? t0 + S : T;
#pragma omp parallel for schedule(static)
        for (long long fb = 0; fb < (long long)(F * B); ++fb) {
            const size_t f = (size_t)fb / B, b = (size_t)fb % B;
            const float* wr = wgt + 2 * ((b * F + f) * A);
            float s = 0.0f;
            double s64 = 0.0;
            for (size_t t = t0; t < t1; ++t) {
                const uint8_t* xr = x + (t * F + f) * A;
                s = s + cell_power(xr, wr, A);
                if (wgt64) {
                    const double* w6 = wgt64 + 2 * ((b * F + f) * A);
                    double sr = 0.0, si = 0.0;
                    for (size_t a = 0; a < A; ++a) {
                        const double vr = or_decode_nibble(xr[a]);
                        const double vi = or_decode_nibble((unsigned)(xr[a] >> 4));
                        sr += w6[2 * a] * vr - w6[2 * a + 1] * vi;
                        si += w6[2 * a] * vi + w6[2 * a + 1] * vr;
                    }
                    s64 += sr * sr + si * si;
                }
            }
            out[(w * F + f) * B + b] = s;
            if (out64) out64[(w * F + f) * B + b] = s64;
        }
    }
    free(wgt);
    free(wgt64);
    return OR_OK;
}

int or_multibeam_tracker_spectra(const uint8_t* x, size_t T, size_t F, size_t A,
                                 size_t B, size_t S, const float* pos3,
                                 const float* freqs, const float* dirs3, float* out,
                                 double* out64) {
    if (S == 0 || B == 0 || A == 0 || F == 0) return OR_INVALID;
    const size_t W = or_tracker_window_count(T, S);
    float* wgt = (float*)malloc(sizeof(float) * 2 * B * F * A);
    double* wgt64 = out64 ? (double*)malloc(sizeof(double) * 2 * B * F * A) : NULL;
    for (size_t w = 0; w < W; ++w) {
        window_weights(B, F, A, pos3, freqs, dirs3 + 3 * B * w, wgt, wgt64);
        const size_t t0 = w * S, t1 = (t0 + S < T) ? t0 + S : T;
#pragma omp parallel for schedule(static)
        for (long long tf = 0; tf < (long long)((t1 - t0) * F); ++tf) {
            const size_t t = t0 + (size_t)tf / F, f = (size_t)tf % F;
            const uint8_t* xr = x + (t * F + f) * A;
            for (size_t b = 0; b < B; ++b) {
                out[(t * F + f) * B + b] = cell_power(xr, wgt + 2 * ((b * F + f) * A), A);
                if (wgt64) {
                    const double* w6 = wgt64 + 2 * ((b * F + f) * A);
                    double sr = 0.0, si = 0.0;
                    for (size_t a = 0; a < A; ++a) {
                        const double vr = or_decode_nibble(xr[a]);
                        const double vi = or_decode_nibble((unsigned)(xr[a] >> 4));
                        sr += w6[2 * a] * vr - w6[2 * a + 1] * vi;
                        si += w6[2 * a] * vi + w6[2 * a + 1] * vr;
                    }
                    out64[(t * F + f) * B + b] = sr * sr + si * si;
                }
            }
        }
    }
    free(wgt);
    free(wgt64);
    return OR_OK;
}

/* -------------------------------------------------------------------- DOA */
void or_covariance_i64(const uint8_t* x, size_t T, size_t F, size_t A, size_t t0,
                       size_t K, size_t f, int64_t* out) {
    (void)T;
    memset(out, 0, sizeof(int64_t) * 2 * A * A);
    for (size_t k = 0; k < K; ++k) {
        const uint8_t* row = x + ((t0 + k) * F + f) * A;
        for (size_t r = 0; r < A; ++r) {
            const int64_t xr = or_decode_nibble(row[r]);
            const int64_t xi = or_decode_nibble((unsigned)(row[r] >> 4));
            for (size_t c = 0; c < A; ++c) {
                const int64_t yr = or_decode_nibble(row[c]);
                const int64_t yi = or_decode_nibble((unsigned)(row[c] >> 4));
                /* x_r * conj(x_c) */
                out[2 * (r * A + c) + 0] += xr * yr + xi * yi;
                out[2 * (r * A + c) + 1] += xi * yr - xr * yi;
            }
        }
    }
}

void or_smooth_forward_i64(const int64_t* R, size_t A, size_t M, int64_t* out) {
    const size_t L = A - M + 1;
    memset(out, 0, sizeof(int64_t) * 2 * M * M);
    for (size_t l = 0; l < L; ++l)
        for (size_t r = 0; r < M; ++r)
            for (size_t c = 0; c < M; ++c) {
                out[2 * (r * M + c) + 0] += R[2 * ((l + r) * A + (l + c)) + 0];
                out[2 * (r * M + c) + 1] += R[2 * ((l + r) * A + (l + c)) + 1];
            }
}

void or_diagonal_load_f64(double* R, size_t M, double load) {
    if (load <= 0.0 || M == 0) return;
    double tr = 0.0;
    for (size_t i = 0; i < M; ++i) tr += R[2 * (i * M + i)];
    const double eps = load * tr / (double)M;
    for (size_t i = 0; i < M; ++i) R[2 * (i * M + i)] += eps;
}

/* Lower Cholesky L L^H of Hermitian R, row-major complex (re, im) pairs;
 * the column-by-column recurrence of src/cpu_opt_beam_tracker.cpp:398-421. */
int or_cholesky_f64(double* R, size_t M) {
    for (size_t i = 0; i < M; ++i) {
        for (size_t j = 0; j <= i; ++j) {
            double sr = R[2 * (i * M + j)], si = R[2 * (i * M + j) + 1];
            for (size_t k = 0; k < j; ++k) {
                const double ar = R[2 * (i * M + k)], ai = R[2 * (i * M + k) + 1];
                const double br = R[2 * (j * M + k)], bi = R[2 * (j * M + k) + 1];
                /* sum -= L[i][k] * conj(L[j][k]) */
                sr -= ar * br + ai * bi;
                si -= ai * br - ar * bi;
            }
            if (i == j) {
                if (!(sr > 0.0)) return 0;
                R[2 * (i * M + j)] = sqrt(sr);
                R[2 * (i * M + j) + 1] = 0.0;
            } else {
                const double inv = 1.0 / R[2 * (j * M + j)];
                R[2 * (i * M + j)] = sr * inv;
                R[2 * (i * M + j) + 1] = si * inv;
            }
        }
    }
    return 1;
}

double or_capon_from_factor_f64(const double* L, const double* a, size_t M) {
    double* v = (double*)malloc(sizeof(double) * 2 * M);
    double nrm = 0.0;
    for (size_t i = 0; i < M; ++i) {
        double sr = a[2 * i], si = a[2 * i + 1];
        for (size_t k = 0; k < i; ++k) {
            const double lr = L[2 * (i * M + k)], li = L[2 * (i * M + k) + 1];
            sr -= lr * v[2 * k] - li * v[2 * k + 1];
            si -= lr * v[2 * k + 1] + li * v[2 * k];
        }
        const double d = L[2 * (i * M + i)];
        v[2 * i] = sr / d;
        v[2 * i + 1] = si / d;
        nrm += v[2 * i] * v[2 * i] + v[2 * i + 1] * v[2 * i + 1];
    }
    free(v);
    return nrm > 0.0 ? 1.0 / nrm : 0.0;
}

double or_bartlett_f64(const double* R, const double* a, size_t M) {
    double acc = 0.0;
    for (size_t r = 0; r < M; ++r) {
        double rr = 0.0, ri = 0.0;
        for (size_t c = 0; c < M; ++c) {
            const double xr = R[2 * (r * M + c)], xi = R[2 * (r * M + c) + 1];
            rr += xr * a[2 * c] - xi * a[2 * c + 1];
            ri += xr * a[2 * c + 1] + xi * a[2 * c];
        }
        /* conj(a_r) * row_sum */
        acc += a[2 * r] * rr + a[2 * r + 1] * ri;
    }
    return acc;
}

double or_quadratic_peak_offset(double pm, double pc, double pp) {
    const double denom = pm - 2.0 * pc + pp;
    if (!(denom < 0.0)) return 0.0;
    double d = 0.5 * (pm - pp) / denom;
    if (d > 0.5) d = 0.5;
    if (d < -0.5) d = -0.5;
    return d;
}

size_t or_doa_peaks(const double* P, const double* theta, size_t n, size_t max_peaks,
                    double* out) {
    size_t count = 0;
    for (size_t i = 0; i < 2 * max_peaks; ++i) out[i] = 0.0;
    for (size_t q = 1; q + 1 < n; ++q) {
        if (!(P[q] > P[q - 1] && P[q] >= P[q + 1])) continue;
        /* slot: after every kept peak at least as strong (earlier q wins ties) */
        size_t slot = 0;
        while (slot < count && out[2 * slot + 1] >= P[q]) ++slot;
        if (slot >= max_peaks) continue;
        const size_t last = count < max_peaks ? count : max_peaks - 1;
        for (size_t s = last; s > slot; --s) {
            out[2 * s] = out[2 * (s - 1)];
            out[2 * s + 1] = out[2 * (s - 1) + 1];
        }
        const double d = or_quadratic_peak_offset(P[q - 1], P[q], P[q + 1]);
        out[2 * slot] = theta[q] + d * (theta[q + 1] - theta[q]);
        out[2 * slot + 1] = P[q];
        if (count < max_peaks) ++count;
    }
    return count;
}

/* The reference loads R_block before the fold only for Capon
 * (src/cpu_opt_beam_tracker.cpp:1879-1892), so its Capon recursion is
 * R_L <- a R_L + (1 - a)(R_block + eps I) and its Bartlett recursion is the
 * unloaded R <- a R + (1 - a) R_block.  Since the load only touches the
 * diagonal, R_L = R + e I with e <- a e + (1 - a) eps: the state carries
 * the unloaded R and the scalar e per channel, Bartlett scans R and Capon
 * factors R + e I. */
int or_doa_update(const uint8_t* x, size_t T, size_t F, size_t A, size_t t0, size_t K,
                  size_t M, double alpha, double load, int first, const float* freqs,
                  const float* xpos, size_t n_theta, const double* theta,
                  double* R_state, double* load_state, double* capon, double* bartlett,
                  int* chol_ok) {
    if (M == 0 || M > A || K == 0 || t0 + K > T) return OR_INVALID;
    const size_t L = A - M + 1;
    int rc = OR_OK;
#pragma omp parallel for schedule(dynamic)
    for (long long fi = 0; fi < (long long)F; ++fi) {
        const size_t f = (size_t)fi;
        int64_t* Rfull = (int64_t*)malloc(sizeof(int64_t) * 2 * A * A);
        int64_t* Rs = (int64_t*)malloc(sizeof(int64_t) * 2 * M * M);
        double* Rb = (double*)malloc(sizeof(double) * 2 * M * M);
        double* Lf = (double*)malloc(sizeof(double) * 2 * M * M);
        double* a = (double*)malloc(sizeof(double) * 2 * M);
        or_covariance_i64(x, T, F, A, t0, K, f, Rfull);
        or_smooth_forward_i64(Rfull, A, M, Rs);
        const double scale = 1.0 / ((double)K * (double)L);
        for (size_t i = 0; i < 2 * M * M; ++i) Rb[i] = (double)Rs[i] * scale;
        double tr = 0.0;
        for (size_t i = 0; i < M; ++i) tr += Rb[2 * (i * M + i)];
        const double eps = load > 0.0 ? load * tr / (double)M : 0.0;
        double* R = R_state + 2 * M * M * f;
        for (size_t i = 0; i < 2 * M * M; ++i)
            R[i] = first ? Rb[i] : alpha * R[i] + (1.0 - alpha) * Rb[i];
        load_state[f] = first ? eps : alpha * load_state[f] + (1.0 - alpha) * eps;
        memcpy(Lf, R, sizeof(double) * 2 * M * M);
        for (size_t i = 0; i < M; ++i) Lf[2 * (i * M + i)] += load_state[f];
        const int ok = or_cholesky_f64(Lf, M);
        chol_ok[f] = ok;
        const double k = wave_number(freqs[f]);
        for (size_t q = 0; q < n_theta; ++q) {
            const double s = sin(theta[q]);
            for (size_t m = 0; m < M; ++m) {
                const double ph = k * (double)xpos[m] * s;
                a[2 * m] = cos(ph);
                a[2 * m + 1] = -sin(ph);
            }
            bartlett[f * n_theta + q] = or_bartlett_f64(R, a, M);
            capon[f * n_theta + q] = ok ? or_capon_from_factor_f64(Lf, a, M) : 0.0;
        }
        free(Rfull);
        free(Rs);
        free(Rb);
        free(Lf);
        free(a);
    }
    return rc;
}

/* ---------------------------------------------------------------- metrics */
double or_max_rel_err_floored(const float* got, const double* ref, size_t n, double floor) {
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double den = fabs(ref[i]);
        if (den < floor) den = floor;
        const double e = fabs((double)got[i] - ref[i]) / den;
        if (e > m || isnan(e)) m = e;
    }
    return m;
}
