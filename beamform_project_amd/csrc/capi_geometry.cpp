// capi_geometry.cpp -- extern "C" array geometry, frequency plans, beam
// grids, their text loaders and the steering-weight generators: the
// producers of the beamformer's weight inputs (src/geometry.cpp:33-451,
// src/weights.cpp:11-122).  Host functions reproduce the reference's float
// and double arithmetic step for step (the values are the contract: the
// same directions and weights feed the same kernels); the weight
// generators also have device forms that write the weights straight into
// HBM (k_weights.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "bf_common.hpp"

namespace bfhip {

void launch_generate_weights(hipStream_t stream, const bf_dims& d, const float* pos,
                             const float* freqs, const float* dirs, float* out, bool tiled);

namespace {

constexpr size_t kTiledBeamTile = 32;  // indexing.hpp:52-61

void require_ptr(const void* p, const char* what) {
    if (p == nullptr) invalid(std::string(what) + " must not be null");
}

// direction_from_lm (src/geometry.cpp:97-106).
void direction_from_lm(float l, float m, float* out) {
    if (!std::isfinite(l) || !std::isfinite(m))
        invalid("direction cosines l and m must be finite");
    const float t2 = l * l + m * m;
    if (t2 > 1.0f) invalid("direction cosines must satisfy l*l + m*m <= 1");
    out[0] = l;
    out[1] = m;
    out[2] = std::sqrt(1.0f - t2);
}

// regular_array (src/geometry.cpp:33-51): x = column * spacing, y = row * spacing.
void regular_array(size_t rows, size_t cols, float spacing, float* out) {
    if (rows == 0 || cols == 0) invalid("array rows and columns must be positive");
    if (spacing <= 0.0f) invalid("antenna spacing must be positive");
    for (size_t r = 0; r < rows; ++r)
        for (size_t c = 0; c < cols; ++c) {
            float* p = out + 3 * (r * cols + c);
            p[0] = (float)c * spacing;
            p[1] = (float)r * spacing;
            p[2] = 0.0f;
        }
}

std::vector<long long> centered_range(size_t count) {
    if (count == 0) invalid("centered integer range count must be positive");
    std::vector<long long> v(count);
    const long long start = -(long long)(count / 2);
    for (size_t i = 0; i < count; ++i) v[i] = start + (long long)i;
    return v;
}

struct Selection {
    std::vector<float> centers;  // [n][2]
    size_t n_u = 0, n_v = 0, n_bank_u = 0, n_bank_v = 0;
};

// select_fft_beam_centers_rectangular (src/geometry.cpp:139-215): the
// n_beams bins of the 2m x 2n zero-padded FFT bank nearest the origin,
// ordered by (radius^2, u, v).
Selection select_fft(float du, float dv, size_t n_beams, size_t m_side, size_t n_side) {
    if (!std::isfinite(du) || du <= 0.0f || !std::isfinite(dv) || dv <= 0.0f)
        invalid("FFT beam spacing must be positive and finite");
    if (m_side == 0 || n_side == 0) invalid("FFT array sides must be positive");
    if (n_beams == 0) invalid("FFT beam count must be positive");
    Selection s;
    s.n_bank_u = 2 * m_side;
    s.n_bank_v = 2 * n_side;
    if (n_beams > s.n_bank_u * s.n_bank_v)
        invalid("n_beams exceeds the zero-padded rectangular FFT bin bank");
    s.n_u = (size_t)std::ceil(std::sqrt((double)n_beams));
    s.n_v = ceil_div(n_beams, s.n_u);
    if (s.n_u > s.n_bank_u) {
        s.n_u = s.n_bank_u;
        s.n_v = ceil_div(n_beams, s.n_u);
    }
    if (s.n_v > s.n_bank_v) {
        s.n_v = s.n_bank_v;
        s.n_u = ceil_div(n_beams, s.n_v);
    }
    if (s.n_u > s.n_bank_u || s.n_u * s.n_v < n_beams)
        fail(BF_ERR_LOGIC, "failed to select a sufficient FFT beam window");
    struct Ranked {
        double r2;
        float u, v;
    };
    std::vector<Ranked> ranked;
    ranked.reserve(s.n_u * s.n_v);
    for (long long i : centered_range(s.n_u))
        for (long long j : centered_range(s.n_v)) {
            const float u = (float)i * du;
            const float v = (float)j * dv;
            ranked.push_back({(double)u * u + (double)v * v, u, v});
        }
    std::sort(ranked.begin(), ranked.end(), [](const Ranked& a, const Ranked& b) {
        return std::tie(a.r2, a.u, a.v) < std::tie(b.r2, b.u, b.v);
    });
    s.centers.resize(2 * n_beams);
    for (size_t k = 0; k < n_beams; ++k) {
        s.centers[2 * k] = ranked[k].u;
        s.centers[2 * k + 1] = ranked[k].v;
    }
    return s;
}

// generate_hex_targets_cropped_fov (src/geometry.cpp:275-330).
std::vector<float> hex_targets(size_t n_beams, float u_max, float v_max) {
    if (n_beams == 0) return {};
    if (!std::isfinite(u_max) || u_max <= 0.0f || !std::isfinite(v_max) || v_max <= 0.0f)
        invalid("hex FoV axes must be positive and finite");
    double a = std::sqrt(kTwoPi / (std::sqrt(3.0) * (double)n_beams));
    std::vector<std::pair<float, float>> pts;
    for (int attempt = 0; attempt < 10; ++attempt) {
        pts.clear();
        const double dy = std::sqrt(3.0) * 0.5 * a;
        const long long jm = (long long)std::ceil(1.0 / std::max(dy, 1.0e-12)) + 2;
        const long long im = (long long)std::ceil(1.0 / std::max(a, 1.0e-12)) + 2;
        for (long long j = -jm; j <= jm; ++j) {
            const double y = (double)j * dy;
            const double shift = j % 2 != 0 ? 0.5 * a : 0.0;
            for (long long i = -im; i <= im; ++i) {
                const double x = (double)i * a + shift;
                if (x * x + y * y <= 1.0) pts.emplace_back((float)(x * u_max), (float)(y * v_max));
            }
        }
        if (pts.size() >= n_beams) break;
        a *= 0.9;
    }
    if (pts.empty()) return {0.0f, 0.0f};
    std::sort(pts.begin(), pts.end(), [](const auto& p, const auto& q) {
        const double rp = (double)p.first * p.first + (double)p.second * p.second;
        const double rq = (double)q.first * q.first + (double)q.second * q.second;
        return std::tie(rp, p.first, p.second) < std::tie(rq, q.first, q.second);
    });
    if (pts.size() > n_beams) pts.resize(n_beams);
    std::vector<float> out(2 * pts.size());
    for (size_t k = 0; k < pts.size(); ++k) {
        out[2 * k] = pts[k].first;
        out[2 * k + 1] = pts[k].second;
    }
    return out;
}

// Text loaders (src/geometry.cpp:17-31, 368-451): '#' starts a comment,
// commas separate like blanks, blank lines are skipped.
std::string data_part(std::string line) {
    const size_t c = line.find('#');
    if (c != std::string::npos) line.erase(c);
    std::replace(line.begin(), line.end(), ',', ' ');
    return line;
}

[[noreturn]] void parse_error(const std::string& path, size_t line, const std::string& why) {
    fail(BF_ERR_RUNTIME, path + ":" + std::to_string(line) + ": " + why);
}

std::vector<float> load_vec3(const std::string& path, size_t expected) {
    std::ifstream in(path);
    if (!in) fail(BF_ERR_RUNTIME, "cannot open positions file: " + path);
    std::vector<float> out;
    std::string line;
    size_t n = 0;
    while (std::getline(in, line)) {
        ++n;
        std::istringstream v(data_part(std::move(line)));
        float p[3];
        if (!(v >> p[0])) continue;
        if (!(v >> p[1] >> p[2])) parse_error(path, n, "expected three coordinates");
        std::string extra;
        if (v >> extra) parse_error(path, n, "unexpected value after third coordinate");
        out.insert(out.end(), p, p + 3);
    }
    if (out.size() / 3 != expected)
        fail(BF_ERR_RUNTIME, "positions file contains " + std::to_string(out.size() / 3) +
                                 " rows; expected " + std::to_string(expected));
    return out;
}

void check_weight_inputs(const bf_dims& d, const float* pos, const float* freqs,
                         const float* dirs) {
    validate_dimensions(d);
    require_ptr(pos, "positions");
    require_ptr(freqs, "frequencies");
    require_ptr(dirs, "directions");
    for (size_t b = 0; b < d.n_beams; ++b) {
        const float* q = dirs + 3 * b;
        const double n2 = (double)q[0] * q[0] + (double)q[1] * q[1] + (double)q[2] * q[2];
        if (!std::isfinite(n2) || std::fabs(n2 - 1.0) > 1.0e-3)
            invalid("beam directions must be finite unit vectors");
    }
    for (size_t f = 0; f < d.n_freq; ++f)
        if (!std::isfinite(freqs[f]) || freqs[f] <= 0.0f)
            invalid("frequencies must be positive and finite");
}

// One weight: (float)cos, (float)sin of the FP64 phase (src/weights.cpp:39-55).
void weight(const float* p, const float* q, double k, float* w) {
    const double delay = (double)p[0] * q[0] + (double)p[1] * q[1] + (double)p[2] * q[2];
    const double ph = k * delay;
    w[0] = (float)std::cos(ph);
    w[1] = (float)std::sin(ph);
}

size_t tiled_count(const bf_dims& d) {
    return d.n_freq * ceil_div(d.n_beams, kTiledBeamTile) * d.n_ant * kTiledBeamTile;
}

}  // namespace
}  // namespace bfhip

using namespace bfhip;

extern "C" {

bf_status bf_regular_array(size_t rows, size_t columns, float spacing_m, float* out) {
    return guarded([&] {
        require_ptr(out, "out");
        regular_array(rows, columns, spacing_m, out);
    });
}

bf_status bf_default_positions(size_t n_ant, float spacing_m, float* out) {
    return guarded([&] {
        require_ptr(out, "out");
        // src/geometry.cpp:53-67
        switch (n_ant) {
        case 32: regular_array(4, 8, spacing_m, out); break;
        case 64: regular_array(8, 8, spacing_m, out); break;
        case 128: regular_array(8, 16, spacing_m, out); break;
        case 256: regular_array(16, 16, spacing_m, out); break;
        default:
            invalid("default geometry is available only for 32, 64, 128, or 256 antennas");
        }
    });
}

bf_status bf_constant_frequencies(size_t n_freq, float frequency_hz, float* out) {
    return guarded([&] {
        require_ptr(out, "out");
        if (n_freq == 0) invalid("frequency count must be positive");
        if (frequency_hz <= 0.0f) invalid("frequency must be positive");
        for (size_t f = 0; f < n_freq; ++f) out[f] = frequency_hz;
    });
}

bf_status bf_channelized_frequencies(size_t n_freq, float start_hz, float channel_width_hz,
                                     float* out) {
    return guarded([&] {
        require_ptr(out, "out");
        if (n_freq == 0) invalid("frequency count must be positive");
        if (!std::isfinite(start_hz) || start_hz <= 0.0f || !std::isfinite(channel_width_hz) ||
            channel_width_hz <= 0.0f)
            invalid("frequency start and channel width must be positive");
        for (size_t f = 0; f < n_freq; ++f) out[f] = start_hz + (float)f * channel_width_hz;
    });
}

bf_status bf_direction_from_lm(float l, float m, float* out) {
    return guarded([&] {
        require_ptr(out, "out");
        direction_from_lm(l, m, out);
    });
}

bf_status bf_default_beam_grid(size_t n_beams, float l_step, float m, float* out) {
    return guarded([&] {
        require_ptr(out, "out");
        // src/geometry.cpp:108-125
        if (n_beams == 0 || n_beams > kMaximumBeams)
            invalid("beam grid size must be between 1 and 128");
        if (!std::isfinite(l_step) || l_step <= 0.0f)
            invalid("beam l step must be positive and finite");
        const long long centre = (long long)(n_beams / 2);
        for (size_t b = 0; b < n_beams; ++b)
            direction_from_lm((float)((long long)b - centre) * l_step, m, out + 3 * b);
    });
}

bf_status bf_centered_integer_range(size_t count, long long* out) {
    return guarded([&] {
        require_ptr(out, "out");
        const std::vector<long long> v = centered_range(count);
        std::copy(v.begin(), v.end(), out);
    });
}

bf_status bf_select_fft_beam_centers_rectangular(float du_fft_u, float dv_fft_v,
                                                 size_t n_beams, size_t m_side, size_t n_side,
                                                 float* centers, size_t* shape) {
    return guarded([&] {
        require_ptr(centers, "centers");
        require_ptr(shape, "shape");
        const Selection s = select_fft(du_fft_u, dv_fft_v, n_beams, m_side, n_side);
        std::copy(s.centers.begin(), s.centers.end(), centers);
        shape[0] = s.n_u;
        shape[1] = s.n_v;
        shape[2] = s.n_bank_u;
        shape[3] = s.n_bank_v;
    });
}

bf_status bf_fft_beam_grid(size_t n_ant, size_t n_beams, float spacing_m,
                           float design_frequency_hz, float* out) {
    return guarded([&] {
        require_ptr(out, "out");
        // src/geometry.cpp:217-249
        if (!std::isfinite(spacing_m) || spacing_m <= 0.0f ||
            !std::isfinite(design_frequency_hz) || design_frequency_hz <= 0.0f)
            invalid("FFT beam-grid spacing and design frequency must be positive");
        const size_t n_side = n_ant == 32 ? 4 : n_ant == 64 ? 8 : 0;
        const size_t m_side = (n_ant == 32 || n_ant == 64) ? 8 : 0;
        if (m_side == 0) invalid("FFT beam grid is available only for 32 or 64 antennas");
        const float lambda = (float)(kSpeedOfLight / (double)design_frequency_hz);
        const float du = lambda / (2.0f * (float)m_side * spacing_m);
        const float dv = lambda / (2.0f * (float)n_side * spacing_m);
        const Selection s = select_fft(du, dv, n_beams, m_side, n_side);
        for (size_t b = 0; b < n_beams; ++b)
            direction_from_lm(s.centers[2 * b], s.centers[2 * b + 1], out + 3 * b);
    });
}

bf_status bf_estimate_nbeams_hex_formula(float spacing_m, size_t m_side, size_t n_side,
                                         float bw_e_deg, float bw_h_deg, float wavelength_m,
                                         size_t* out) {
    return guarded([&] {
        require_ptr(out, "out");
        // src/geometry.cpp:251-273
        if (!std::isfinite(spacing_m) || spacing_m <= 0.0f || m_side == 0 || n_side == 0 ||
            !std::isfinite(bw_e_deg) || bw_e_deg <= 0.0f || !std::isfinite(bw_h_deg) ||
            bw_h_deg <= 0.0f || !std::isfinite(wavelength_m) || wavelength_m <= 0.0f)
            invalid("hex beam estimate inputs must be positive");
        const double eta = kTwoPi / (4.0 * std::sqrt(3.0));
        const double du = (double)spacing_m * (double)(m_side - 1);
        const double dv = (double)spacing_m * (double)(n_side - 1);
        const double rad = kTwoPi / 360.0;
        const double est = eta * 4.0 * du * dv * std::sin((double)bw_e_deg * rad / 2.0) *
                           std::sin((double)bw_h_deg * rad / 2.0) /
                           ((double)wavelength_m * wavelength_m);
        *out = std::max<size_t>(1, (size_t)std::ceil(est));
    });
}

bf_status bf_generate_hex_targets_cropped_fov(size_t n_beams, float u_max, float v_max,
                                              float* out, size_t* n_out) {
    return guarded([&] {
        require_ptr(n_out, "n_out");
        const std::vector<float> v = hex_targets(n_beams, u_max, v_max);
        *n_out = v.size() / 2;
        if (!v.empty()) {
            require_ptr(out, "out");
            std::copy(v.begin(), v.end(), out);
        }
    });
}

bf_status bf_rectangular_beam_grid(size_t n_ant, float spacing_m, float design_frequency_hz,
                                   float* out) {
    return guarded([&] {
        require_ptr(out, "out");
        // src/geometry.cpp:332-366
        if (spacing_m <= 0.0f || !std::isfinite(spacing_m))
            invalid("antenna spacing must be positive and finite");
        if (design_frequency_hz <= 0.0f || !std::isfinite(design_frequency_hz))
            invalid("beam-grid design frequency must be positive and finite");
        const size_t rows = n_ant == 32 ? 4 : n_ant == 64 ? 8 : 0;
        const size_t cols = (n_ant == 32 || n_ant == 64) ? 8 : 0;
        if (rows == 0) invalid("rectangular beam grid is available only for 32 or 64 beams");
        const double lambda = kSpeedOfLight / (double)design_frequency_hz;
        const double dl = lambda / ((double)cols * spacing_m);
        const double dm = lambda / ((double)rows * spacing_m);
        for (size_t r = 0; r < rows; ++r) {
            const double m = ((double)r - (double)(rows - 1) / 2.0) * dm;
            for (size_t c = 0; c < cols; ++c) {
                const double l = ((double)c - (double)(cols - 1) / 2.0) * dl;
                direction_from_lm((float)l, (float)m, out + 3 * (r * cols + c));
            }
        }
    });
}

bf_status bf_load_positions(const char* path, size_t expected_count, float* out) {
    return guarded([&] {
        require_ptr(path, "path");
        require_ptr(out, "out");
        const std::vector<float> v = load_vec3(path, expected_count);
        std::copy(v.begin(), v.end(), out);
    });
}

bf_status bf_load_frequencies(const char* path, size_t expected_count, float* out) {
    return guarded([&] {
        require_ptr(path, "path");
        require_ptr(out, "out");
        const std::string p(path);
        std::ifstream in(p);
        if (!in) fail(BF_ERR_RUNTIME, "cannot open frequency file: " + p);
        std::vector<float> fr;
        std::string line;
        size_t n = 0;
        while (std::getline(in, line)) {
            ++n;
            std::istringstream v(data_part(std::move(line)));
            float f = 0.0f;
            if (!(v >> f)) continue;
            if (f <= 0.0f) parse_error(p, n, "frequency must be positive");
            std::string extra;
            if (v >> extra) parse_error(p, n, "expected one frequency per line");
            fr.push_back(f);
        }
        if (fr.size() != expected_count)
            fail(BF_ERR_RUNTIME, "frequency file contains " + std::to_string(fr.size()) +
                                     " rows; expected " + std::to_string(expected_count));
        std::copy(fr.begin(), fr.end(), out);
    });
}

bf_status bf_load_directions(const char* path, size_t expected_count, float* out) {
    return guarded([&] {
        require_ptr(path, "path");
        require_ptr(out, "out");
        const std::vector<float> v = load_vec3(path, expected_count);
        for (size_t b = 0; b < expected_count; ++b) {
            const float* q = &v[3 * b];
            const float n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
            if (!std::isfinite(n2) || std::fabs(n2 - 1.0f) > 1.0e-3f)
                fail(BF_ERR_RUNTIME, "beam directions must be finite unit vectors");
        }
        std::copy(v.begin(), v.end(), out);
    });
}

bf_status bf_generate_weights(const bf_dims* dims, const float* positions,
                              const float* frequencies, const float* directions, float* out,
                              size_t n_complex) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        check_weight_inputs(d, positions, frequencies, directions);
        require_ptr(out, "out");
        if (n_complex < d.n_beams * d.n_freq * d.n_ant)
            invalid("weight output is smaller than dimensions");
        for (size_t b = 0; b < d.n_beams; ++b)
            for (size_t f = 0; f < d.n_freq; ++f) {
                const double k = kTwoPi * (double)frequencies[f] / kSpeedOfLight;
                for (size_t a = 0; a < d.n_ant; ++a)
                    weight(positions + 3 * a, directions + 3 * b, k,
                           out + 2 * ((b * d.n_freq + f) * d.n_ant + a));
            }
    });
}

bf_status bf_generate_tiled_weights(const bf_dims* dims, const float* positions,
                                    const float* frequencies, const float* directions,
                                    float* out, size_t n_complex) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        check_weight_inputs(d, positions, frequencies, directions);
        require_ptr(out, "out");
        const size_t n = tiled_count(d);
        if (n_complex < n) invalid("tiled weight output is smaller than dimensions");
        std::memset(out, 0, 2 * n * sizeof(float));
        const size_t tiles = ceil_div(d.n_beams, kTiledBeamTile);
        for (size_t f = 0; f < d.n_freq; ++f) {
            const double k = kTwoPi * (double)frequencies[f] / kSpeedOfLight;
            for (size_t bt = 0; bt < tiles; ++bt)
                for (size_t a = 0; a < d.n_ant; ++a)
                    for (size_t lb = 0; lb < kTiledBeamTile; ++lb) {
                        const size_t b = bt * kTiledBeamTile + lb;
                        if (b >= d.n_beams) continue;
                        weight(positions + 3 * a, directions + 3 * b, k,
                               out + 2 * (((f * tiles + bt) * d.n_ant + a) * kTiledBeamTile + lb));
                    }
        }
    });
}

bf_status bf_generate_weights_device(void* stream, const bf_dims* dims,
                                     const float* d_positions, const float* d_frequencies,
                                     const float* d_directions, float* d_weights, int tiled) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        require_ptr(d_positions, "positions");
        require_ptr(d_frequencies, "frequencies");
        require_ptr(d_directions, "directions");
        require_ptr(d_weights, "weights");
        launch_generate_weights(static_cast<hipStream_t>(stream), d, d_positions, d_frequencies,
                                d_directions, d_weights, tiled != 0);
    });
}

}  // extern "C"
