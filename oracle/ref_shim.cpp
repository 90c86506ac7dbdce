// ref_shim.cpp -- C entry points over the reference CPU core, compiled
// together with the reference's own sources (see oracle/Makefile) into
// oracle/_ref/libbfref.so.  TEST INFRASTRUCTURE ONLY: used to generate the
// golden fixtures under tests/golden/ and, where present, as the timed CPU
// baseline.  Every function catches the reference's exceptions and returns
// -1 (invalid_argument), -2 (out_of_range) or -3 (other).
#include "beamformer/beam_tracker.hpp"
#include "beamformer/beam_tracker_opt_v2.hpp"
#include "beamformer/cpu_beamformer.hpp"
#include "beamformer/cpu_opt_beam_tracker.hpp"
#include "beamformer/geometry.hpp"
#include "beamformer/io.hpp"
#include "beamformer/quantization.hpp"
#include "beamformer/synthetic_data.hpp"
#include "beamformer/temporal_integration.hpp"
#include "beamformer/weights.hpp"

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

using namespace beamformer;

namespace {

template <typename Fn>
int guarded(Fn&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::invalid_argument&) {
        return -1;
    } catch (const std::out_of_range&) {
        return -2;
    } catch (...) {
        return -3;
    }
}

std::vector<Vec3> vec3s(const float* p, std::size_t n) {
    std::vector<Vec3> v(n);
    for (std::size_t i = 0; i < n; ++i) v[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    return v;
}

void put_vec3s(const std::vector<Vec3>& v, float* out) {
    for (std::size_t i = 0; i < v.size(); ++i) {
        out[3 * i] = v[i][0];
        out[3 * i + 1] = v[i][1];
        out[3 * i + 2] = v[i][2];
    }
}

}  // namespace

extern "C" {

int ref_default_positions(std::size_t n_ant, float spacing, float* out) {
    return guarded([&] { put_vec3s(default_positions(n_ant, spacing), out); });
}

int ref_channelized_frequencies(std::size_t n, float start, float width, float* out) {
    return guarded([&] {
        const auto f = channelized_frequencies(n, start, width);
        std::memcpy(out, f.data(), n * sizeof(float));
    });
}

int ref_direction_from_lm(float l, float m, float* out) {
    return guarded([&] { put_vec3s({direction_from_lm(l, m)}, out); });
}

int ref_default_beam_grid(std::size_t n, float step, float m, float* out) {
    return guarded([&] { put_vec3s(default_beam_grid(n, step, m), out); });
}

int ref_rectangular_beam_grid(std::size_t n_ant, float spacing, float design, float* out) {
    return guarded([&] { put_vec3s(rectangular_beam_grid(n_ant, spacing, design), out); });
}

int ref_fft_beam_grid(std::size_t n_ant, std::size_t n_beams, float* out) {
    return guarded([&] { put_vec3s(fft_beam_grid(n_ant, n_beams), out); });
}

int ref_make_loss_mask(std::size_t T, std::size_t F, std::size_t A, std::uint32_t seed,
                       float p, std::uint8_t* out) {
    return guarded([&] {
        const auto m = make_loss_mask(Dimensions{T, F, A, 1}, seed, p);
        std::memcpy(out, m.data(), m.size());
    });
}

int ref_hex_targets(std::size_t n, float u_max, float v_max, float* out, std::size_t* n_out) {
    return guarded([&] {
        const auto v = generate_hex_targets_cropped_fov(n, u_max, v_max);
        *n_out = v.size();
        for (std::size_t i = 0; i < v.size(); ++i) {
            out[2 * i] = v[i][0];
            out[2 * i + 1] = v[i][1];
        }
    });
}

int ref_estimate_hex(float spacing, std::size_t m_side, std::size_t n_side, float bw_e,
                     float bw_h, float wavelength, std::size_t* out) {
    return guarded([&] {
        *out = estimate_nbeams_hex_formula(spacing, m_side, n_side, bw_e, bw_h, wavelength);
    });
}

int ref_select_fft(float du, float dv, std::size_t n, std::size_t m_side, std::size_t n_side,
                   float* centers, std::size_t* shape) {
    return guarded([&] {
        const auto s = select_fft_beam_centers_rectangular(du, dv, n, m_side, n_side);
        for (std::size_t i = 0; i < s.centers.size(); ++i) {
            centers[2 * i] = s.centers[i][0];
            centers[2 * i + 1] = s.centers[i][1];
        }
        shape[0] = s.n_u;
        shape[1] = s.n_v;
        shape[2] = s.n_bank_u;
        shape[3] = s.n_bank_v;
    });
}

int ref_fft_beam_grid_at(std::size_t n_ant, std::size_t n_beams, float spacing, float design,
                         float* out) {
    return guarded([&] { put_vec3s(fft_beam_grid(n_ant, n_beams, spacing, design), out); });
}

int ref_generate_weights(std::size_t B, std::size_t F, std::size_t A, const float* pos,
                         const float* freqs, const float* dirs, float* out) {
    return guarded([&] {
        const Dimensions d{1, F, A, B};
        const auto w = generate_weights(d, vec3s(pos, A),
                                        std::vector<float>(freqs, freqs + F), vec3s(dirs, B));
        std::memcpy(out, w.data(), w.size() * sizeof(ComplexFloat));
    });
}

int ref_generate_tiled_weights(std::size_t B, std::size_t F, std::size_t A,
                               const float* pos, const float* freqs, const float* dirs,
                               float* out) {
    return guarded([&] {
        const Dimensions d{1, F, A, B};
        const auto w = generate_tiled_weights(
            d, vec3s(pos, A), std::vector<float>(freqs, freqs + F), vec3s(dirs, B));
        std::memcpy(out, w.data(), w.size() * sizeof(ComplexFloat));
    });
}

int ref_beamform_packed(const std::uint8_t* x, const float* w, std::size_t T,
                        std::size_t F, std::size_t A, std::size_t B, float* out) {
    return guarded([&] {
        const Dimensions d{T, F, A, B};
        PackedVoltage p(x, x + T * F * A);
        Weights wv(B * F * A);
        std::memcpy(wv.data(), w, wv.size() * sizeof(ComplexFloat));
        Intensities o(T * F * B);
        cpu_beamform_packed_intensity_into(p, wv, d, o);
        std::memcpy(out, o.data(), o.size() * sizeof(float));
    });
}

int ref_integrate(const float* in, std::size_t T, std::size_t F, std::size_t A,
                  std::size_t B, std::size_t S, float* out) {
    return guarded([&] {
        const Dimensions d{T, F, A, B};
        Intensities i(in, in + T * F * B);
        const auto o = cpu_integrate_intensity(i, d, S);
        std::memcpy(out, o.data(), o.size() * sizeof(float));
    });
}

int ref_quantize(const float* in, std::size_t Tw, std::size_t F, std::size_t A,
                 std::size_t B, std::int8_t* codes, float* params) {
    return guarded([&] {
        const Dimensions d{Tw, F, A, B};
        IntegratedIntensities i(in, in + Tw * F * B);
        const auto q = cpu_quantize_integrated_intensity(i, d);
        std::memcpy(codes, q.codes.data(), q.codes.size());
        std::memcpy(params, q.parameters.data(),
                    q.parameters.size() * sizeof(Int8QuantizationParameters));
    });
}

int ref_make_noise(std::size_t T, std::size_t F, std::size_t A, std::uint32_t seed,
                   std::uint8_t* out) {
    return guarded([&] {
        const auto v = make_noise(Dimensions{T, F, A, 1}, seed);
        std::memcpy(out, v.data(), v.size());
    });
}

int ref_make_point_source(std::size_t T, std::size_t F, std::size_t A, const float* pos,
                          const float* freqs, const float* dir, float amp,
                          std::uint8_t* out) {
    return guarded([&] {
        const auto v = make_point_source(Dimensions{T, F, A, 1}, vec3s(pos, A),
                                         std::vector<float>(freqs, freqs + F),
                                         Vec3{dir[0], dir[1], dir[2]}, amp);
        std::memcpy(out, v.data(), v.size());
    });
}

int ref_make_moving_point_source(std::size_t T, std::size_t F, std::size_t A,
                                 const float* pos, const float* freqs, float l0, float m0,
                                 float rl, float rm, float amp, std::uint8_t* out) {
    return guarded([&] {
        TrackerTrajectoryConfig tr{direction_from_lm(l0, m0), {rl, rm}};
        const auto v = beam_tracker_make_moving_point_source(
            Dimensions{T, F, A, 1}, vec3s(pos, A), std::vector<float>(freqs, freqs + F),
            tr, amp);
        std::memcpy(out, v.data(), v.size());
    });
}

int ref_beam_tracker(const std::uint8_t* x, std::size_t T, std::size_t F, std::size_t A,
                     float l0, float m0, float rl, float rm, std::size_t S, float* out) {
    return guarded([&] {
        TrackerConfig tc;
        tc.trajectory.direction_start = direction_from_lm(l0, m0);
        tc.trajectory.direction_rate_per_sample = {rl, rm};
        tc.integration_spectra = S;
        const Dimensions d{T, F, A, 1};
        PackedVoltage p(x, x + T * F * A);
        Intensities o(T * F);
        beam_tracker_cpu_packed_intensity_into(p, d, tc, o);
        std::memcpy(out, o.data(), o.size() * sizeof(float));
    });
}

int ref_beam_tracker_opt_v2(const std::uint8_t* x, std::size_t T, std::size_t F,
                            std::size_t A, float l0, float m0, float rl, float rm,
                            std::size_t S, float* out) {
    return guarded([&] {
        TrackerConfig tc;
        tc.trajectory.direction_start = direction_from_lm(l0, m0);
        tc.trajectory.direction_rate_per_sample = {rl, rm};
        tc.integration_spectra = S;
        const Dimensions d{T, F, A, 1};
        PackedVoltage p(x, x + T * F * A);
        Intensities o(T * F);
        beam_tracker_opt_v2_cpu_packed_intensity_into(p, d, tc, o);
        std::memcpy(out, o.data(), o.size() * sizeof(float));
    });
}

int ref_tracker_direction(float l0, float m0, float rl, float rm, std::size_t t,
                          float* out) {
    return guarded([&] {
        TrackerTrajectoryConfig tr{direction_from_lm(l0, m0), {rl, rm}};
        put_vec3s({tracker_direction(tr, t)}, out);
    });
}

// ---- on-disk formats (src/io.cpp): the reference writes and reads files
// so tests can compare them byte for byte with the library's.
int ref_write_quantized_metadata(const char* path, std::size_t T, std::size_t F, std::size_t A,
                                 std::size_t B, std::size_t input_n_time, std::size_t S,
                                 std::size_t shard_id, const char* parameters_file) {
    return guarded([&] {
        const QuantizedOutputMetadata m{Dimensions{T, F, A, B}, input_n_time,
                                        TemporalIntegrationConfig{S},
                                        default_shard_descriptors()[shard_id],
                                        parameters_file};
        write_quantized_output_metadata(path, m);
    });
}

// Reads the metadata at `path`; out = {n_time, n_freq, n_ant, n_beams,
// input_n_time, integration_spectra, shard_id, absolute_frequency_start,
// loss_mask_id, timestamp_start, timestamp_step}.
int ref_read_quantized_metadata(const char* path, std::uint64_t* out) {
    return guarded([&] {
        const auto m = read_quantized_output_metadata(path);
        const std::uint64_t v[11] = {m.output_dims.n_time, m.output_dims.n_freq,
                                     m.output_dims.n_ant, m.output_dims.n_beams,
                                     m.input_n_time, m.temporal_integration.integration_spectra,
                                     m.shard.shard_id, m.shard.absolute_frequency_start,
                                     m.shard.loss_mask_id, m.shard.timestamp_start,
                                     m.shard.timestamp_step};
        std::memcpy(out, v, sizeof(v));
    });
}

int ref_write_packed_shard(const char* payload_path, const char* metadata_path,
                           const char* mask_path, const std::uint8_t* payload,
                           const std::uint8_t* mask, std::size_t T, std::size_t F,
                           std::size_t A, std::size_t B, std::size_t shard_id,
                           std::uint64_t timestamp_start, std::uint64_t timestamp_step) {
    return guarded([&] {
        const Dimensions d{T, F, A, B};
        PackedShard sh;
        sh.descriptor = default_shard_descriptors()[shard_id];
        sh.descriptor.timestamp_start = timestamp_start;
        sh.descriptor.timestamp_step = timestamp_step;
        sh.payload.assign(payload, payload + T * F * A);
        sh.loss_mask.assign(mask, mask + T * F);
        write_packed_shard(payload_path, metadata_path, mask_path, sh, d);
    });
}

int ref_read_packed_shard(const char* payload_path, const char* metadata_path,
                          const char* mask_path, std::size_t T, std::size_t F, std::size_t A,
                          std::size_t B, std::uint8_t* payload, std::uint8_t* mask,
                          std::uint64_t* descriptor8) {
    return guarded([&] {
        const auto sh = read_packed_shard(payload_path, metadata_path, mask_path,
                                          Dimensions{T, F, A, B});
        std::memcpy(payload, sh.payload.data(), sh.payload.size());
        std::memcpy(mask, sh.loss_mask.data(), sh.loss_mask.size());
        const auto& s = sh.descriptor;
        const std::uint64_t v[8] = {s.shard_id, s.shard_count, s.local_frequency_count,
                                    s.absolute_frequency_start, s.timestamp_start,
                                    s.timestamp_step, s.loss_mask_id,
                                    s.loss_mask_independent ? 1u : 0u};
        std::memcpy(descriptor8, v, sizeof(v));
    });
}

int ref_write_tiled_weights_file(const char* path, const float* w, std::size_t T, std::size_t F,
                                 std::size_t A, std::size_t B) {
    return guarded([&] {
        const Dimensions d{T, F, A, B};
        const std::size_t n = tiled_weight_count(d);
        TiledWeights tw(n);
        std::memcpy(tw.data(), w, n * sizeof(ComplexFloat));
        write_tiled_weights(path, tw, d);
    });
}

// n_calls consecutive run_into calls of one CpuOptBeamTracker over
// consecutive T-spectrum chunks of x (the covariance recursion carries
// over), each seeded with the same prior; out [n_calls][T][F], dirs
// [n_calls][W][3].
int ref_cpu_opt_tracker(const std::uint8_t* x, std::size_t n_calls, std::size_t T, std::size_t F,
                        std::size_t A, const float* pos, const float* freqs, const float* start,
                        const float* rate, std::size_t S, int estimator, std::size_t G,
                        std::size_t levels, float fov_l, float fov_m, std::size_t p_sub,
                        float lambda, float load, int interp, float* out, float* dirs) {
    return guarded([&] {
        const Dimensions dims{T, F, A, 1};
        CpuOptTrackerConfig c;
        c.estimator = estimator ? TrackerEstimator::Capon : TrackerEstimator::Bartlett;
        c.coarse_grid_resolution = G;
        c.refinement_levels = levels;
        c.search_fov_l = fov_l;
        c.search_fov_m = fov_m;
        c.spatial_smoothing_subarray_size = p_sub;
        c.forgetting_factor = lambda;
        c.diagonal_load = load;
        c.enable_quadratic_peak_interp = interp != 0;
        c.integration_spectra = S;
        CpuOptBeamTracker tracker(vec3s(pos, A), std::vector<float>(freqs, freqs + F), dims, c);
        TrackerTrajectoryConfig prior;
        prior.direction_start = Vec3{start[0], start[1], start[2]};
        prior.direction_rate_per_sample = {rate[0], rate[1]};
        const std::size_t W = tracker_window_count(T, S);
        const std::size_t chunk = T * F * A;
        Intensities intensity(T * F);
        for (std::size_t k = 0; k < n_calls; ++k) {
            PackedVoltage packed(x + k * chunk, x + (k + 1) * chunk);
            tracker.seed_trajectory(prior);
            tracker.run_into(packed, intensity);
            std::memcpy(out + k * T * F, intensity.data(), T * F * sizeof(float));
            for (std::size_t w = 0; w < W; ++w) {
                const Vec3 d = tracker.window_direction(w);
                for (int i = 0; i < 3; ++i) dirs[(k * W + w) * 3 + i] = d[i];
            }
        }
    });
}

}  // extern "C"
