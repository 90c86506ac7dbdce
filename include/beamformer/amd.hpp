// amd.hpp -- header-only C++ layer over bf_capi.h with the reference's
// names, argument meaning and exception types (no HIP headers needed).
//
// Reference surfaces mirrored here:
//   Dimensions, validate_dimensions        include/beamformer/config.hpp:74-94
//   integrated_time_count                  include/beamformer/temporal_integration.hpp
//   quantization_parameter_count           include/beamformer/quantization.hpp
//   launch_packed_beamformer,
//   launch_packed_integrated_beamformer,
//   launch_quantize_integrated_intensity   include/beamformer/cuda_stage_api.hpp:14-42
//   cuda_beamform_packed_*_intensity       include/beamformer/cuda_beamformer.hpp:98-123
//   TrackerConfig, cuda_beam_tracker_v5_*  include/beamformer/beam_tracker.hpp:42-57,
//                                          include/beamformer/cuda_beam_tracker_v5.hpp:69-85
// Status codes become std::invalid_argument, std::out_of_range,
// std::overflow_error, std::logic_error or std::runtime_error carrying the
// library's message, as the reference throws them.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../bf_capi.h"

namespace beamformer {
namespace amd {

inline void check(bf_status s) {
    if (s == BF_OK) return;
    const char* m = bf_last_error_message();
    const std::string msg = m ? m : "";
    switch (s) {
    case BF_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case BF_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
    case BF_ERR_OVERFLOW: throw std::overflow_error(msg);
    case BF_ERR_LOGIC: throw std::logic_error(msg);
    default: throw std::runtime_error(msg);
    }
}

struct Dimensions {
    std::size_t n_time = 0;
    std::size_t n_freq = 0;
    std::size_t n_ant = 0;
    std::size_t n_beams = 0;
    bf_dims c() const { return bf_dims{n_time, n_freq, n_ant, n_beams}; }
};

// CudaBeamformerKernel{Direct, Tiled} keep their meaning (canonical or
// tiled weights, FP32); the MFMA kernels and Auto are additions.
enum class Kernel : int {
    Direct = BF_KERNEL_DIRECT,
    Tiled = BF_KERNEL_TILED,
    Fp32 = BF_KERNEL_FP32,
    SplitF16 = BF_KERNEL_SPLIT_F16,
    F16 = BF_KERNEL_F16,
    Auto = BF_KERNEL_AUTO,
};

struct TemporalIntegrationConfig {
    std::size_t integration_spectra = 0;
};

inline void validate_dimensions(const Dimensions& d) {
    const bf_dims c = d.c();
    check(bf_validate_dimensions(&c));
}

inline std::size_t integrated_time_count(std::size_t n_time, std::size_t spectra) {
    std::size_t out = 0;
    check(bf_integrated_time_count(n_time, spectra, &out));
    return out;
}

inline std::size_t quantization_parameter_count(const Dimensions& integrated) {
    const bf_dims c = integrated.c();
    std::size_t out = 0;
    check(bf_quantization_parameter_count(&c, &out));
    return out;
}

// Caller-owned device buffers; asynchronous on `stream`; no allocation.
inline void launch_packed_beamformer(Kernel kernel, void* stream, const std::uint8_t* packed,
                                     const float* weights, float* intensity,
                                     const Dimensions& d) {
    const bf_dims c = d.c();
    check(bf_launch_packed_beamformer(static_cast<int>(kernel), stream, packed, weights,
                                      intensity, &c));
}

inline void launch_packed_integrated_beamformer(Kernel kernel, void* stream,
                                                const std::uint8_t* packed,
                                                const float* weights, float* integrated,
                                                const Dimensions& d,
                                                TemporalIntegrationConfig cfg) {
    const bf_dims c = d.c();
    check(bf_launch_packed_integrated_beamformer(static_cast<int>(kernel), stream, packed,
                                                 weights, integrated, &c,
                                                 cfg.integration_spectra));
}

inline void launch_quantize_integrated_intensity(void* stream, const float* integrated,
                                                 std::int8_t* codes, float* parameters,
                                                 const Dimensions& integrated_dims) {
    const bf_dims c = integrated_dims.c();
    check(bf_launch_quantize_integrated_intensity(stream, integrated, codes, parameters, &c));
}

// One-shot host API: uploads, runs and downloads (synchronous).  Weights
// are complex f32 interleaved (re, im), [B][F][A] or tiled.
inline std::vector<float> beamform_packed_intensity(const std::vector<std::uint8_t>& packed,
                                                    const std::vector<float>& weights,
                                                    const Dimensions& d,
                                                    Kernel kernel = Kernel::Direct,
                                                    bf_timings* timings = nullptr) {
    const bf_dims c = d.c();
    std::vector<float> out(d.n_time * d.n_freq * d.n_beams);
    check(bf_beamform_packed_intensity(packed.data(), packed.size(), weights.data(),
                                       weights.size() / 2, &c, static_cast<int>(kernel),
                                       out.data(), out.size(), timings));
    return out;
}

inline std::vector<float> beamform_packed_integrated_intensity(
    const std::vector<std::uint8_t>& packed, const std::vector<float>& weights,
    const Dimensions& d, TemporalIntegrationConfig cfg, Kernel kernel = Kernel::Tiled,
    bf_timings* timings = nullptr) {
    const bf_dims c = d.c();
    std::vector<float> out(integrated_time_count(d.n_time, cfg.integration_spectra) *
                           d.n_freq * d.n_beams);
    check(bf_beamform_packed_integrated_intensity(
        packed.data(), packed.size(), weights.data(), weights.size() / 2, &c,
        cfg.integration_spectra, static_cast<int>(kernel), out.data(), out.size(), timings));
    return out;
}

// cuda_beamform_packed_quantized_integrated_intensity (cuda_beamformer.hpp:
// 113-123): int8 codes [w][f][b] and {offset, scale} pairs per 1x16x16 chunk.
struct QuantizedIntegratedIntensity {
    std::vector<std::int8_t> codes;
    std::vector<float> parameters;  // {offset, scale} pairs
};

inline QuantizedIntegratedIntensity beamform_packed_quantized_integrated_intensity(
    const std::vector<std::uint8_t>& packed, const std::vector<float>& weights,
    const Dimensions& d, TemporalIntegrationConfig cfg, Kernel kernel = Kernel::Tiled,
    bf_timings* timings = nullptr) {
    const bf_dims c = d.c();
    const Dimensions out_dims{integrated_time_count(d.n_time, cfg.integration_spectra),
                              d.n_freq, d.n_ant, d.n_beams};
    QuantizedIntegratedIntensity out;
    out.codes.resize(out_dims.n_time * d.n_freq * d.n_beams);
    out.parameters.resize(2 * quantization_parameter_count(out_dims));
    check(bf_beamform_packed_quantized_integrated_intensity(
        packed.data(), packed.size(), weights.data(), weights.size() / 2, &c,
        cfg.integration_spectra, static_cast<int>(kernel), out.codes.data(), out.codes.size(),
        out.parameters.data(), out.parameters.size() / 2, timings));
    return out;
}

// Reference TrackerConfig: direction at sample 0 and linear drift per sample.
struct TrackerConfig {
    float direction_start[3] = {0.0f, 0.0f, 1.0f};
    float direction_rate_per_sample[2] = {0.0f, 0.0f};
    std::size_t integration_spectra = 320;
    bf_tracker_config c() const {
        return bf_tracker_config{{direction_start[0], direction_start[1], direction_start[2]},
                                 {direction_rate_per_sample[0], direction_rate_per_sample[1]},
                                 integration_spectra};
    }
};

struct TrackerExecution {
    Kernel kernel = Kernel::Auto;
    bool enable_graph = false;
    bf_tracker_exec c() const {
        return bf_tracker_exec{static_cast<int>(kernel), enable_graph ? 1 : 0};
    }
};

// Device-resident single-beam tracker: [T][F][1] per-spectrum power; syncs
// only when stream is null.
inline void beam_tracker_device_resident(const std::uint8_t* d_packed, float* d_intensity,
                                         const Dimensions& d, const TrackerConfig& t,
                                         const TrackerExecution& e = {},
                                         void* stream = nullptr) {
    const bf_dims c = d.c();
    const bf_tracker_config tc = t.c();
    const bf_tracker_exec ec = e.c();
    check(bf_beam_tracker_device_resident(d_packed, d_intensity, &c, &tc, &ec, stream));
}

inline std::vector<float> beam_tracker_into(const std::vector<std::uint8_t>& packed,
                                            const Dimensions& d, const TrackerConfig& t,
                                            const TrackerExecution& e = {}) {
    const bf_dims c = d.c();
    const bf_tracker_config tc = t.c();
    const bf_tracker_exec ec = e.c();
    std::vector<float> out(d.n_time * d.n_freq);
    check(bf_beam_tracker_into(packed.data(), packed.size(), &c, &tc, out.data(), out.size(),
                               &ec));
    return out;
}

}  // namespace amd
}  // namespace beamformer
