// amd_geometry.hpp -- header-only C++ layer over the geometry, beam-grid and
// weight functions of bf_capi.h, with the reference's names, types and
// exceptions (include/beamformer/geometry.hpp, weights.hpp).  Values equal
// the reference's bit for bit (the host forms); generate_weights_device
// writes the same weights into device memory.
#pragma once

#include <array>
#include <cstddef>
#include <filesystem>
#include <vector>

#include "amd.hpp"

namespace beamformer {
namespace amd {

using Vec3 = std::array<float, 3>;
using Vec2 = std::array<float, 2>;

inline constexpr float default_spacing_m = 0.6F;                     // config.hpp:27
inline constexpr float beam_grid_design_frequency_hz = 400'000'000.0F;  // config.hpp:26
inline constexpr std::size_t tiled_weight_beam_tile = 32;              // indexing.hpp:52

struct FftBeamSelection {
    std::vector<Vec2> centers;
    std::size_t n_u = 0;
    std::size_t n_v = 0;
    std::size_t n_bank_u = 0;
    std::size_t n_bank_v = 0;
};

namespace detail {
template <typename V>
inline std::vector<V> sized(std::size_t n) {
    return std::vector<V>(n);
}
template <typename V>
inline float* flat(std::vector<V>& v) {
    return v.empty() ? nullptr : v.front().data();
}
template <typename V>
inline const float* flat(const std::vector<V>& v) {
    return v.empty() ? nullptr : v.front().data();
}
}  // namespace detail

inline std::vector<Vec3> regular_array(std::size_t rows, std::size_t columns,
                                       float spacing_m = default_spacing_m) {
    auto out = detail::sized<Vec3>(rows * columns);
    check(bf_regular_array(rows, columns, spacing_m, detail::flat(out)));
    return out;
}

inline std::vector<Vec3> default_positions(std::size_t n_ant,
                                           float spacing_m = default_spacing_m) {
    std::vector<Vec3> out(n_ant ? n_ant : 1);
    check(bf_default_positions(n_ant, spacing_m, detail::flat(out)));
    return out;
}

inline std::vector<float> constant_frequencies(
    std::size_t n_freq = 336, float frequency_hz = beam_grid_design_frequency_hz) {
    std::vector<float> out(n_freq ? n_freq : 1);
    check(bf_constant_frequencies(n_freq, frequency_hz, out.data()));
    return out;
}

inline std::vector<float> channelized_frequencies(std::size_t n_freq = 336,
                                                  float start_hz = 300'000'000.0F,
                                                  float channel_width_hz = 300'000.0F) {
    std::vector<float> out(n_freq ? n_freq : 1);
    check(bf_channelized_frequencies(n_freq, start_hz, channel_width_hz, out.data()));
    return out;
}

inline Vec3 direction_from_lm(float l, float m) {
    Vec3 out{};
    check(bf_direction_from_lm(l, m, out.data()));
    return out;
}

inline std::vector<Vec3> default_beam_grid(std::size_t n_beams, float l_step = 0.02F,
                                           float m = 0.0F) {
    std::vector<Vec3> out(n_beams ? n_beams : 1);
    check(bf_default_beam_grid(n_beams, l_step, m, detail::flat(out)));
    return out;
}

inline std::vector<long long> centered_integer_range(std::size_t count) {
    std::vector<long long> out(count ? count : 1);
    check(bf_centered_integer_range(count, out.data()));
    return out;
}

inline FftBeamSelection select_fft_beam_centers_rectangular(float du_fft_u, float dv_fft_v,
                                                            std::size_t n_beams,
                                                            std::size_t m_side,
                                                            std::size_t n_side) {
    FftBeamSelection s;
    s.centers.resize(n_beams ? n_beams : 1);
    std::size_t shape[4] = {0, 0, 0, 0};
    check(bf_select_fft_beam_centers_rectangular(du_fft_u, dv_fft_v, n_beams, m_side, n_side,
                                                 detail::flat(s.centers), shape));
    s.n_u = shape[0];
    s.n_v = shape[1];
    s.n_bank_u = shape[2];
    s.n_bank_v = shape[3];
    return s;
}

inline std::vector<Vec3> fft_beam_grid(std::size_t n_ant, std::size_t n_beams,
                                       float spacing_m = default_spacing_m,
                                       float design_frequency_hz = beam_grid_design_frequency_hz) {
    std::vector<Vec3> out(n_beams ? n_beams : 1);
    check(bf_fft_beam_grid(n_ant, n_beams, spacing_m, design_frequency_hz, detail::flat(out)));
    return out;
}

inline std::size_t estimate_nbeams_hex_formula(float spacing_m, std::size_t m_side,
                                               std::size_t n_side, float bw_e_deg,
                                               float bw_h_deg, float wavelength_m) {
    std::size_t out = 0;
    check(bf_estimate_nbeams_hex_formula(spacing_m, m_side, n_side, bw_e_deg, bw_h_deg,
                                         wavelength_m, &out));
    return out;
}

inline std::vector<Vec2> generate_hex_targets_cropped_fov(std::size_t n_beams, float u_max,
                                                          float v_max) {
    std::vector<Vec2> out(n_beams ? n_beams : 1);
    std::size_t n = 0;
    check(bf_generate_hex_targets_cropped_fov(n_beams, u_max, v_max, detail::flat(out), &n));
    out.resize(n);
    return out;
}

inline std::vector<Vec3> rectangular_beam_grid(
    std::size_t n_ant, float spacing_m = default_spacing_m,
    float design_frequency_hz = beam_grid_design_frequency_hz) {
    std::vector<Vec3> out(n_ant ? n_ant : 1);
    check(bf_rectangular_beam_grid(n_ant, spacing_m, design_frequency_hz, detail::flat(out)));
    return out;
}

inline std::vector<Vec3> load_positions(const std::filesystem::path& path,
                                        std::size_t expected_count) {
    std::vector<Vec3> out(expected_count ? expected_count : 1);
    check(bf_load_positions(path.c_str(), expected_count, detail::flat(out)));
    out.resize(expected_count);
    return out;
}

inline std::vector<float> load_frequencies(const std::filesystem::path& path,
                                           std::size_t expected_count) {
    std::vector<float> out(expected_count ? expected_count : 1);
    check(bf_load_frequencies(path.c_str(), expected_count, out.data()));
    out.resize(expected_count);
    return out;
}

inline std::vector<Vec3> load_directions(const std::filesystem::path& path,
                                         std::size_t expected_count) {
    std::vector<Vec3> out(expected_count ? expected_count : 1);
    check(bf_load_directions(path.c_str(), expected_count, detail::flat(out)));
    out.resize(expected_count);
    return out;
}

// Interleaved complex f32, [B][F][A] (generate_weights) and
// [F][ceil(B/32)][A][32] (generate_tiled_weights).
inline std::vector<float> generate_weights(const Dimensions& dims,
                                           const std::vector<Vec3>& positions_m,
                                           const std::vector<float>& frequencies_hz,
                                           const std::vector<Vec3>& beam_directions) {
    if (positions_m.size() != dims.n_ant)
        throw std::invalid_argument("position count must match n_ant");
    if (frequencies_hz.size() != dims.n_freq)
        throw std::invalid_argument("frequency count must match n_freq");
    if (beam_directions.size() != dims.n_beams)
        throw std::invalid_argument("beam direction count must match n_beams");
    const bf_dims c = dims.c();
    std::vector<float> out(2 * dims.n_beams * dims.n_freq * dims.n_ant);
    check(bf_generate_weights(&c, detail::flat(positions_m), frequencies_hz.data(),
                              detail::flat(beam_directions), out.data(), out.size() / 2));
    return out;
}

inline std::vector<float> generate_tiled_weights(const Dimensions& dims,
                                                 const std::vector<Vec3>& positions_m,
                                                 const std::vector<float>& frequencies_hz,
                                                 const std::vector<Vec3>& beam_directions) {
    if (positions_m.size() != dims.n_ant)
        throw std::invalid_argument("position count must match n_ant");
    if (frequencies_hz.size() != dims.n_freq)
        throw std::invalid_argument("frequency count must match n_freq");
    if (beam_directions.size() != dims.n_beams)
        throw std::invalid_argument("beam direction count must match n_beams");
    const bf_dims c = dims.c();
    const std::size_t tiles = (dims.n_beams + tiled_weight_beam_tile - 1) / tiled_weight_beam_tile;
    std::vector<float> out(2 * dims.n_freq * tiles * dims.n_ant * tiled_weight_beam_tile);
    check(bf_generate_tiled_weights(&c, detail::flat(positions_m), frequencies_hz.data(),
                                    detail::flat(beam_directions), out.data(), out.size() / 2));
    return out;
}

// Device arrays: positions [A][3], frequencies [F], directions [B][3].
inline void generate_weights_device(const Dimensions& dims, const float* d_positions,
                                    const float* d_frequencies, const float* d_directions,
                                    float* d_weights, bool tiled = false,
                                    void* stream = nullptr) {
    const bf_dims c = dims.c();
    check(bf_generate_weights_device(stream, &c, d_positions, d_frequencies, d_directions,
                                     d_weights, tiled ? 1 : 0));
}

}  // namespace amd
}  // namespace beamformer
