// reference_api.hpp -- the reference's C++ host API for the beamformer and
// tracker hot path, by name and type, over the C-ABI in bf_capi.h
// (header-only; no HIP headers needed).  A caller of the reference's
// CUDA path replaces its includes of
//   beamformer/{config,formats,quantization,temporal_integration,
//   beam_tracker,cuda_stage_api,cuda_beamformer,cuda_frame,
//   cuda_offline_runner,cuda_tracker_v2,cuda_beam_tracker_fused_warp_shuffle,
//   cuda_beam_tracker_v3,cuda_beam_tracker_v4,cuda_beam_tracker_v5}.hpp
// by this one header and links libbeamform_hip.so.
//
// Mirrored surfaces (reference file:line):
//   Dimensions, ShardDescriptor              include/beamformer/config.hpp:29-79
//   ComplexFloat, PackedVoltage, Weights ... include/beamformer/complex.hpp:7-10,
//                                            include/beamformer/formats.hpp:13-24
//   TemporalIntegrationConfig                include/beamformer/temporal_integration.hpp:14-21
//   Int8QuantizationParameters,
//   QuantizedIntegratedOutput                include/beamformer/quantization.hpp:23-48
//   CudaBeamformerKernel, launch_*          include/beamformer/cuda_stage_api.hpp:14-42
//   CudaBeamformerOutput, CudaBeamformerTimings, CudaDeviceInfo,
//   CudaBeamformerWorkspace, cuda_beamform_packed_*
//                                            include/beamformer/cuda_beamformer.hpp:12-123
//   CudaDataType, CudaFrameView, make_*_frame_view
//                                            include/beamformer/cuda_frame.hpp:20-230
//   CudaOfflineFrameRunner                   include/beamformer/cuda_offline_runner.hpp:17-52
//   TrackerTrajectoryConfig, TrackerConfig   include/beamformer/beam_tracker.hpp:42-57
//   cuda_tracker_v2_*                        include/beamformer/cuda_tracker_v2.hpp:20-53
//   cuda_beam_tracker_fused_warp_shuffle_*   include/beamformer/cuda_beam_tracker_fused_warp_shuffle.hpp:47-85
//   cuda_beam_tracker_v3_*                   include/beamformer/cuda_beam_tracker_v3.hpp:52-113
//   cuda_beam_tracker_v4_*                   include/beamformer/cuda_beam_tracker_v4.hpp:40-103
//   cuda_beam_tracker_v5_*, BatchedTrackerStreamV5
//                                            include/beamformer/cuda_beam_tracker_v5.hpp:20-125
//
// Differences a caller sees:
//   * CudaBeamformerKernel gains SplitF16, F16, Fp32 and Auto; Direct and
//     Tiled keep their meaning (canonical or tiled f32 weights, FP32 math).
//   * The tracker generations V2, FWS, V3, V4 and V5 all run the one
//     tracker of this library; their execution-config fields (unroll, chunk
//     sizes, kernel modes, load strategies, stream counts) have no
//     counterpart and are accepted and ignored, except enable_cuda_graph.
//     Their results are therefore identical to one another.
//   * Errors are the reference's exception types carrying this library's
//     message.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../bf_capi.h"

namespace beamformer {

namespace detail {
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
}  // namespace detail

// ------------------------------------------------------------ data contract
inline constexpr std::size_t local_frequency_channels = 336;
inline constexpr std::size_t frequency_shard_count = 2;

struct Dimensions {
    std::size_t n_time = 15360;
    std::size_t n_freq = local_frequency_channels;
    std::size_t n_ant = 64;
    std::size_t n_beams = 64;
};

struct ShardDescriptor {
    std::size_t shard_id = 0;
    std::size_t shard_count = frequency_shard_count;
    std::size_t local_frequency_count = local_frequency_channels;
    std::size_t absolute_frequency_start = 0;
    std::uint64_t timestamp_start = 0;
    std::uint64_t timestamp_step = 1;
    std::uint64_t loss_mask_id = 0;
    bool loss_mask_independent = true;
};

struct ComplexFloat {
    float real;
    float imag;
};
static_assert(sizeof(ComplexFloat) == 2 * sizeof(float));

using PackedVoltage = std::vector<std::uint8_t>;
using Weights = std::vector<ComplexFloat>;
using TiledWeights = std::vector<ComplexFloat>;
using Intensities = std::vector<float>;
using IntegratedIntensities = std::vector<float>;
using QuantizedIntensities = std::vector<std::int8_t>;
using Vec2 = std::array<float, 2>;
using Vec3 = std::array<float, 3>;

struct TemporalIntegrationConfig {
    std::size_t integration_spectra;
};
inline constexpr TemporalIntegrationConfig integration_after_upchan{10};
inline constexpr TemporalIntegrationConfig integration_direct{320};

struct Int8QuantizationParameters {
    float offset = 0.0F;
    float scale = 0.0F;
};
static_assert(sizeof(Int8QuantizationParameters) == 2 * sizeof(float));

struct QuantizedIntegratedOutput {
    QuantizedIntensities codes;
    std::vector<Int8QuantizationParameters> parameters;
};

namespace detail {
inline bf_dims c_dims(const Dimensions& d) { return bf_dims{d.n_time, d.n_freq, d.n_ant, d.n_beams}; }
inline Dimensions from_c(const bf_dims& d) { return Dimensions{d.n_time, d.n_freq, d.n_ant, d.n_beams}; }
inline bf_shard c_shard(const ShardDescriptor& s) {
    return bf_shard{s.shard_id,        s.shard_count,    s.local_frequency_count,
                    s.absolute_frequency_start, s.timestamp_start, s.timestamp_step,
                    s.loss_mask_id,    s.loss_mask_independent ? 1 : 0};
}
inline const float* floats(const std::vector<ComplexFloat>& w) {
    return reinterpret_cast<const float*>(w.data());
}
}  // namespace detail

inline void validate_dimensions(const Dimensions& d) {
    const bf_dims c = detail::c_dims(d);
    detail::check(bf_validate_dimensions(&c));
}

inline void validate_shard_descriptor(const ShardDescriptor& s) {
    const bf_shard c = detail::c_shard(s);
    detail::check(bf_validate_shard_descriptor(&c));
}

inline std::size_t integrated_time_count(std::size_t n_time, std::size_t spectra) {
    std::size_t out = 0;
    detail::check(bf_integrated_time_count(n_time, spectra, &out));
    return out;
}

inline std::size_t quantization_parameter_count(const Dimensions& integrated) {
    const bf_dims c = detail::c_dims(integrated);
    std::size_t out = 0;
    detail::check(bf_quantization_parameter_count(&c, &out));
    return out;
}

// -------------------------------------------------------- stage launchers
enum class CudaBeamformerKernel : int {
    Direct = BF_KERNEL_DIRECT,
    Tiled = BF_KERNEL_TILED,
    Fp32 = BF_KERNEL_FP32,
    SplitF16 = BF_KERNEL_SPLIT_F16,
    F16 = BF_KERNEL_F16,
    Auto = BF_KERNEL_AUTO,
};

// Caller-owned device buffers; asynchronous on `stream`; no allocation, no
// synchronisation; validation errors are thrown before any launch.
inline void launch_packed_beamformer(CudaBeamformerKernel kernel, void* stream,
                                     const std::uint8_t* packed, const ComplexFloat* weights,
                                     float* intensity, const Dimensions& dims) {
    const bf_dims c = detail::c_dims(dims);
    detail::check(bf_launch_packed_beamformer(static_cast<int>(kernel), stream, packed,
                                              reinterpret_cast<const float*>(weights),
                                              intensity, &c));
}

inline void launch_packed_integrated_beamformer(CudaBeamformerKernel kernel, void* stream,
                                                const std::uint8_t* packed,
                                                const ComplexFloat* weights, float* integrated,
                                                const Dimensions& dims,
                                                const TemporalIntegrationConfig& cfg) {
    const bf_dims c = detail::c_dims(dims);
    detail::check(bf_launch_packed_integrated_beamformer(
        static_cast<int>(kernel), stream, packed, reinterpret_cast<const float*>(weights),
        integrated, &c, cfg.integration_spectra));
}

inline void launch_quantize_integrated_intensity(void* stream, const float* integrated,
                                                 std::int8_t* codes,
                                                 Int8QuantizationParameters* parameters,
                                                 const Dimensions& integrated_dims) {
    const bf_dims c = detail::c_dims(integrated_dims);
    detail::check(bf_launch_quantize_integrated_intensity(
        stream, integrated, codes, reinterpret_cast<float*>(parameters), &c));
}

// ------------------------------------------------------------ workspace
enum class CudaBeamformerOutput {
    Float32,
    QuantizedInt8,
};

struct CudaBeamformerTimings {
    double setup_ms = 0.0;
    double host_to_device_ms = 0.0;
    double kernel_ms = 0.0;
    double temporal_integration_ms = 0.0;
    double quantization_ms = 0.0;
    double device_to_device_ms = 0.0;
    double device_to_host_ms = 0.0;
};

struct CudaDeviceInfo {
    std::string name;
    int compute_major = 0;  // gfx950 reported as 9.5
    int compute_minor = 0;
    std::size_t global_memory_bytes = 0;
    int driver_version = 0;
    int runtime_version = 0;
};

inline CudaDeviceInfo cuda_device_info() {
    char name[256] = {0};
    std::size_t mem = 0;
    int cus = 0, clk = 0;
    detail::check(bf_device_info(name, sizeof(name), &mem, &cus, &clk));
    CudaDeviceInfo info;
    info.name = name;
    info.compute_major = 9;
    info.compute_minor = 5;
    info.global_memory_bytes = mem;
    return info;
}

namespace detail {
inline CudaBeamformerTimings timings(const bf_timings& t) {
    return CudaBeamformerTimings{t.setup_ms, t.host_to_device_ms, t.kernel_ms,
                                 t.temporal_integration_ms, t.quantization_ms,
                                 t.device_to_device_ms, t.device_to_host_ms};
}
inline int c_output(CudaBeamformerOutput o) {
    return o == CudaBeamformerOutput::QuantizedInt8 ? BF_OUTPUT_INT8 : BF_OUTPUT_FLOAT32;
}
}  // namespace detail

// --------------------------------------------------------- frame views
enum class CudaDataType {
    PackedInt4x2 = BF_DTYPE_PACKED_INT4X2,
    ComplexFloat32 = BF_DTYPE_COMPLEX_FLOAT32,
    Float32 = BF_DTYPE_FLOAT32,
    SignedInt8 = BF_DTYPE_SIGNED_INT8,
};

// A non-owning device buffer with its frame metadata (the reference's
// CudaBufferView + CudaFrameMetadata, flattened into the C-ABI's view).
struct CudaFrameView {
    bf_frame_view view{};
};

inline CudaFrameView make_contiguous_cuda_frame_view(void* device_data, std::size_t bytes,
                                                     CudaDataType type,
                                                     const std::array<std::size_t, 4>& shape,
                                                     std::size_t rank, std::uint64_t frame_id,
                                                     const ShardDescriptor& shard) {
    validate_shard_descriptor(shard);
    static constexpr std::size_t elem[] = {1, 8, 4, 1};
    CudaFrameView f;
    f.view.device_data = device_data;
    f.view.bytes = bytes;
    f.view.type = static_cast<int>(type);
    f.view.rank = rank;
    std::size_t stride = elem[static_cast<int>(type)];
    for (std::size_t i = rank; i-- > 0;) {
        f.view.shape[i] = shape[i];
        f.view.strides_bytes[i] = stride;
        stride *= shape[i];
    }
    f.view.frame_id = frame_id;
    f.view.shard = detail::c_shard(shard);
    if (stride > bytes)
        throw std::invalid_argument("frame bytes are smaller than its shape and strides");
    return f;
}

inline CudaFrameView make_packed_voltage_frame_view(void* d, const Dimensions& dims,
                                                    std::uint64_t id, const ShardDescriptor& s) {
    validate_dimensions(dims);
    return make_contiguous_cuda_frame_view(d, dims.n_time * dims.n_freq * dims.n_ant,
                                           CudaDataType::PackedInt4x2,
                                           {dims.n_time, dims.n_freq, dims.n_ant, 0}, 3, id, s);
}
inline CudaFrameView make_intensity_frame_view(void* d, const Dimensions& dims,
                                               std::uint64_t id, const ShardDescriptor& s) {
    validate_dimensions(dims);
    return make_contiguous_cuda_frame_view(d, dims.n_time * dims.n_freq * dims.n_beams * 4,
                                           CudaDataType::Float32,
                                           {dims.n_time, dims.n_freq, dims.n_beams, 0}, 3, id, s);
}
inline CudaFrameView make_quantized_intensity_frame_view(void* d, const Dimensions& dims,
                                                         std::uint64_t id,
                                                         const ShardDescriptor& s) {
    validate_dimensions(dims);
    return make_contiguous_cuda_frame_view(d, dims.n_time * dims.n_freq * dims.n_beams,
                                           CudaDataType::SignedInt8,
                                           {dims.n_time, dims.n_freq, dims.n_beams, 0}, 3, id, s);
}
inline CudaFrameView make_weights_frame_view(void* d, const Dimensions& dims, std::uint64_t id,
                                             const ShardDescriptor& s) {
    validate_dimensions(dims);
    return make_contiguous_cuda_frame_view(d, dims.n_beams * dims.n_freq * dims.n_ant * 8,
                                           CudaDataType::ComplexFloat32,
                                           {dims.n_beams, dims.n_freq, dims.n_ant, 0}, 3, id, s);
}
inline CudaFrameView make_tiled_weights_frame_view(void* d, const Dimensions& dims,
                                                   std::uint64_t id, const ShardDescriptor& s) {
    validate_dimensions(dims);
    const std::size_t tiles = (dims.n_beams + 31) / 32;
    return make_contiguous_cuda_frame_view(d, dims.n_freq * tiles * dims.n_ant * 32 * 8,
                                           CudaDataType::ComplexFloat32,
                                           {dims.n_freq, tiles, dims.n_ant, 32}, 4, id, s);
}
inline CudaFrameView make_quantization_parameters_frame_view(void* d,
                                                             const Dimensions& integrated,
                                                             std::uint64_t id,
                                                             const ShardDescriptor& s) {
    validate_dimensions(integrated);
    const std::size_t ft = (integrated.n_freq + 15) / 16, bt = (integrated.n_beams + 15) / 16;
    return make_contiguous_cuda_frame_view(d, integrated.n_time * ft * bt * 8,
                                           CudaDataType::Float32, {integrated.n_time, ft, bt, 2},
                                           4, id, s);
}

class CudaBeamformerWorkspace {
  public:
    explicit CudaBeamformerWorkspace(
        const Dimensions& capacity, CudaBeamformerKernel kernel = CudaBeamformerKernel::Direct,
        std::optional<TemporalIntegrationConfig> temporal_integration = std::nullopt,
        CudaBeamformerOutput output = CudaBeamformerOutput::Float32)
        : kernel_(kernel) {
        const bf_dims c = detail::c_dims(capacity);
        bf_workspace* ws = nullptr;
        detail::check(bf_workspace_create(
            &c, static_cast<int>(kernel),
            temporal_integration ? temporal_integration->integration_spectra : 0,
            detail::c_output(output), &ws));
        ws_.reset(ws);
    }
    CudaBeamformerWorkspace(const CudaBeamformerWorkspace&) = delete;
    CudaBeamformerWorkspace& operator=(const CudaBeamformerWorkspace&) = delete;

    double setup_ms() const {
        double v = 0.0;
        detail::check(bf_workspace_setup_ms(ws_.get(), &v));
        return v;
    }
    CudaBeamformerKernel kernel() const { return kernel_; }
    double upload_packed_voltage(const PackedVoltage& packed, const Dimensions& dims) {
        const bf_dims c = detail::c_dims(dims);
        double ms = 0.0;
        detail::check(bf_workspace_upload_packed_voltage(ws_.get(), packed.data(), packed.size(),
                                                         &c, &ms));
        return ms;
    }
    double upload_weights(const Weights& weights, const Dimensions& dims) {
        const bf_dims c = detail::c_dims(dims);
        double ms = 0.0;
        detail::check(bf_workspace_upload_weights(ws_.get(), detail::floats(weights),
                                                  weights.size(), &c, &ms));
        return ms;
    }
    double run_kernel(const Dimensions& dims) {
        const bf_dims c = detail::c_dims(dims);
        double ms = 0.0;
        detail::check(bf_workspace_run_kernel(ws_.get(), &c, &ms));
        return ms;
    }
    double download_intensity(Intensities& intensity, const Dimensions& dims) {
        const bf_dims c = detail::c_dims(dims);
        double ms = 0.0;
        detail::check(bf_workspace_download_intensity(ws_.get(), intensity.data(),
                                                      intensity.size(), &c, &ms));
        return ms;
    }
    CudaBeamformerTimings run_pipeline(const PackedVoltage& packed, Intensities& intensity,
                                       const Dimensions& dims) {
        const bf_dims c = detail::c_dims(dims);
        bf_timings t{};
        detail::check(bf_workspace_run_pipeline(ws_.get(), packed.data(), packed.size(),
                                                intensity.data(), intensity.size(), &c, &t));
        return detail::timings(t);
    }
    CudaBeamformerTimings run_integrated_pipeline(const PackedVoltage& packed,
                                                  IntegratedIntensities& intensity,
                                                  const Dimensions& dims) {
        const bf_dims c = detail::c_dims(dims);
        bf_timings t{};
        detail::check(bf_workspace_run_integrated_pipeline(
            ws_.get(), packed.data(), packed.size(), intensity.data(), intensity.size(), &c, &t));
        return detail::timings(t);
    }
    CudaBeamformerTimings run_quantized_integrated_pipeline(const PackedVoltage& packed,
                                                            QuantizedIntegratedOutput& out,
                                                            const Dimensions& dims) {
        const bf_dims c = detail::c_dims(dims);
        bf_timings t{};
        detail::check(bf_workspace_run_quantized_integrated_pipeline(
            ws_.get(), packed.data(), packed.size(), out.codes.data(), out.codes.size(),
            reinterpret_cast<float*>(out.parameters.data()), out.parameters.size(), &c, &t));
        return detail::timings(t);
    }
    CudaBeamformerTimings run_device_frame(const CudaFrameView& voltage,
                                           const CudaFrameView& weights, CudaFrameView& output,
                                           CudaFrameView* quantization_parameters = nullptr) {
        bf_timings t{};
        detail::check(bf_workspace_run_device_frame(
            ws_.get(), &voltage.view, &weights.view, &output.view,
            quantization_parameters ? &quantization_parameters->view : nullptr, &t));
        return detail::timings(t);
    }
    bool has_temporal_integration() const {
        int v = 0;
        detail::check(bf_workspace_has_temporal_integration(ws_.get(), &v));
        return v != 0;
    }
    bool has_int8_quantization() const {
        int v = 0;
        detail::check(bf_workspace_has_int8_quantization(ws_.get(), &v));
        return v != 0;
    }
    std::size_t packed_voltage_capacity_bytes() const {
        std::size_t v = 0;
        detail::check(bf_workspace_packed_voltage_capacity_bytes(ws_.get(), &v));
        return v;
    }

  private:
    struct Del {
        void operator()(bf_workspace* w) const noexcept { (void)bf_workspace_destroy(w); }
    };
    std::unique_ptr<bf_workspace, Del> ws_;
    CudaBeamformerKernel kernel_;
};

// --------------------------------------------------------- one-shot API
inline Intensities cuda_beamform_packed_intensity(
    const PackedVoltage& packed, const Weights& weights, const Dimensions& dims,
    CudaBeamformerTimings* timings = nullptr,
    CudaBeamformerKernel kernel = CudaBeamformerKernel::Direct) {
    const bf_dims c = detail::c_dims(dims);
    Intensities out(dims.n_time * dims.n_freq * dims.n_beams);
    bf_timings t{};
    detail::check(bf_beamform_packed_intensity(packed.data(), packed.size(),
                                               detail::floats(weights), weights.size(), &c,
                                               static_cast<int>(kernel), out.data(), out.size(),
                                               &t));
    if (timings) *timings = detail::timings(t);
    return out;
}

inline IntegratedIntensities cuda_beamform_packed_integrated_intensity(
    const PackedVoltage& packed, const Weights& weights, const Dimensions& dims,
    const TemporalIntegrationConfig& temporal_integration,
    CudaBeamformerTimings* timings = nullptr,
    CudaBeamformerKernel kernel = CudaBeamformerKernel::Tiled) {
    const bf_dims c = detail::c_dims(dims);
    const std::size_t S = temporal_integration.integration_spectra;
    IntegratedIntensities out(integrated_time_count(dims.n_time, S) * dims.n_freq *
                              dims.n_beams);
    bf_timings t{};
    detail::check(bf_beamform_packed_integrated_intensity(
        packed.data(), packed.size(), detail::floats(weights), weights.size(), &c, S,
        static_cast<int>(kernel), out.data(), out.size(), &t));
    if (timings) *timings = detail::timings(t);
    return out;
}

inline QuantizedIntegratedOutput cuda_beamform_packed_quantized_integrated_intensity(
    const PackedVoltage& packed, const Weights& weights, const Dimensions& dims,
    const TemporalIntegrationConfig& temporal_integration,
    CudaBeamformerTimings* timings = nullptr,
    CudaBeamformerKernel kernel = CudaBeamformerKernel::Tiled) {
    const bf_dims c = detail::c_dims(dims);
    const std::size_t S = temporal_integration.integration_spectra;
    const Dimensions od{integrated_time_count(dims.n_time, S), dims.n_freq, dims.n_ant,
                        dims.n_beams};
    QuantizedIntegratedOutput out;
    out.codes.resize(od.n_time * od.n_freq * od.n_beams);
    out.parameters.resize(quantization_parameter_count(od));
    bf_timings t{};
    detail::check(bf_beamform_packed_quantized_integrated_intensity(
        packed.data(), packed.size(), detail::floats(weights), weights.size(), &c, S,
        static_cast<int>(kernel), out.codes.data(), out.codes.size(),
        reinterpret_cast<float*>(out.parameters.data()), out.parameters.size(), &t));
    if (timings) *timings = detail::timings(t);
    return out;
}

// -------------------------------------------------- offline frame runner
struct CudaOfflineFrameResult {
    Dimensions output_dims;
    CudaBeamformerTimings timings;
    Intensities float32_output;
    QuantizedIntegratedOutput quantized_output;
};

class CudaOfflineFrameRunner {
  public:
    explicit CudaOfflineFrameRunner(
        const Dimensions& capacity, CudaBeamformerKernel kernel = CudaBeamformerKernel::Direct,
        std::optional<TemporalIntegrationConfig> temporal_integration = std::nullopt,
        CudaBeamformerOutput output = CudaBeamformerOutput::Float32)
        : kernel_(kernel), output_(output), integration_(temporal_integration) {
        const bf_dims c = detail::c_dims(capacity);
        bf_offline_frame_runner* r = nullptr;
        detail::check(bf_offline_frame_runner_create(
            &c, static_cast<int>(kernel),
            temporal_integration ? temporal_integration->integration_spectra : 0,
            detail::c_output(output), &r));
        r_.reset(r);
    }
    CudaOfflineFrameRunner(const CudaOfflineFrameRunner&) = delete;
    CudaOfflineFrameRunner& operator=(const CudaOfflineFrameRunner&) = delete;

    CudaOfflineFrameResult run(const PackedVoltage& packed, const Weights& weights,
                               const Dimensions& dims, std::uint64_t frame_id,
                               const ShardDescriptor& shard) {
        const bf_dims c = detail::c_dims(dims);
        const bf_shard sh = detail::c_shard(shard);
        validate_dimensions(dims);
        Dimensions od = dims;
        if (integration_)
            od.n_time = integrated_time_count(dims.n_time, integration_->integration_spectra);
        CudaOfflineFrameResult res;
        bf_dims out_dims{};
        bf_timings t{};
        if (output_ == CudaBeamformerOutput::QuantizedInt8) {
            res.quantized_output.codes.resize(od.n_time * od.n_freq * od.n_beams);
            res.quantized_output.parameters.resize(
                integration_ ? quantization_parameter_count(od) : 0);
            detail::check(bf_offline_frame_runner_run(
                r_.get(), packed.data(), packed.size(), detail::floats(weights),
                weights.size(), &c, frame_id, &sh, nullptr, 0,
                res.quantized_output.codes.data(), res.quantized_output.codes.size(),
                reinterpret_cast<float*>(res.quantized_output.parameters.data()),
                res.quantized_output.parameters.size(), &out_dims, &t));
        } else {
            res.float32_output.resize(od.n_time * od.n_freq * od.n_beams);
            detail::check(bf_offline_frame_runner_run(
                r_.get(), packed.data(), packed.size(), detail::floats(weights),
                weights.size(), &c, frame_id, &sh, res.float32_output.data(),
                res.float32_output.size(), nullptr, 0, nullptr, 0, &out_dims, &t));
        }
        res.output_dims = detail::from_c(out_dims);
        res.timings = detail::timings(t);
        return res;
    }
    CudaBeamformerKernel kernel() const { return kernel_; }
    CudaBeamformerOutput output() const { return output_; }
    bool has_temporal_integration() const { return integration_.has_value(); }

  private:
    struct Del {
        void operator()(bf_offline_frame_runner* r) const noexcept {
            (void)bf_offline_frame_runner_destroy(r);
        }
    };
    std::unique_ptr<bf_offline_frame_runner, Del> r_;
    CudaBeamformerKernel kernel_;
    CudaBeamformerOutput output_;
    std::optional<TemporalIntegrationConfig> integration_;
};

// ---------------------------------------------------------------- tracker
inline Vec3 direction_from_lm(float l, float m) {
    Vec3 d{};
    detail::check(bf_direction_from_lm(l, m, d.data()));
    return d;
}

struct TrackerTrajectoryConfig {
    Vec3 direction_start = Vec3{0.0F, 0.0F, 1.0F};
    Vec2 direction_rate_per_sample = {0.0F, 0.0F};
};

struct TrackerConfig {
    TrackerTrajectoryConfig trajectory;
    std::size_t integration_spectra = integration_direct.integration_spectra;
};

inline constexpr std::size_t tracker_beam_count = 1;

namespace detail {
inline bf_tracker_config c_tracker(const TrackerConfig& t) {
    const auto& d = t.trajectory.direction_start;
    const auto& r = t.trajectory.direction_rate_per_sample;
    return bf_tracker_config{{d[0], d[1], d[2]}, {r[0], r[1]}, t.integration_spectra};
}
inline void tracker_into(const PackedVoltage& packed, const Dimensions& dims,
                         const TrackerConfig& tracker, Intensities& intensity, bool graph) {
    const bf_dims c = c_dims(dims);
    const bf_tracker_config tc = c_tracker(tracker);
    const bf_tracker_exec ec{BF_KERNEL_AUTO, graph ? 1 : 0};
    if (intensity.size() < dims.n_time * dims.n_freq)
        intensity.resize(dims.n_time * dims.n_freq);
    check(bf_beam_tracker_into(packed.data(), packed.size(), &c, &tc, intensity.data(),
                               intensity.size(), &ec));
}
inline void tracker_device(const std::uint8_t* d_packed, float* d_intensity,
                           const Dimensions& dims, const TrackerConfig& tracker, bool graph,
                           void* stream) {
    const bf_dims c = c_dims(dims);
    const bf_tracker_config tc = c_tracker(tracker);
    const bf_tracker_exec ec{BF_KERNEL_AUTO, graph ? 1 : 0};
    check(bf_beam_tracker_device_resident(d_packed, d_intensity, &c, &tc, &ec, stream));
}
}  // namespace detail

// V5 (the reference's best tracker)
struct V5ExecutionConfig {
    std::size_t time_chunk_size = 80;
    std::size_t time_unroll = 8;
    bool enable_cuda_graph = false;
};

inline void cuda_beam_tracker_v5_into(const PackedVoltage& packed, const Dimensions& dims,
                                      const TrackerConfig& tracker, Intensities& intensity,
                                      const V5ExecutionConfig& config = V5ExecutionConfig{}) {
    detail::tracker_into(packed, dims, tracker, intensity, config.enable_cuda_graph);
}
inline Intensities cuda_beam_tracker_v5(const PackedVoltage& packed, const Dimensions& dims,
                                        const TrackerConfig& tracker,
                                        const V5ExecutionConfig& config = V5ExecutionConfig{}) {
    Intensities out;
    cuda_beam_tracker_v5_into(packed, dims, tracker, out, config);
    return out;
}
inline void cuda_beam_tracker_v5_device_resident(
    const std::uint8_t* d_packed, float* d_intensity, const Dimensions& dims,
    const TrackerConfig& tracker, const V5ExecutionConfig& config = V5ExecutionConfig{},
    void* stream = nullptr) {
    detail::tracker_device(d_packed, d_intensity, dims, tracker, config.enable_cuda_graph,
                           stream);
}

class BatchedTrackerStreamV5 {
  public:
    BatchedTrackerStreamV5(const Dimensions& single_window_dims, const TrackerConfig& tracker,
                           std::size_t batch_size,
                           const V5ExecutionConfig& config = V5ExecutionConfig{}) {
        const bf_dims c = detail::c_dims(single_window_dims);
        const bf_tracker_config tc = detail::c_tracker(tracker);
        const bf_tracker_exec ec{BF_KERNEL_AUTO, config.enable_cuda_graph ? 1 : 0};
        bf_tracker_stream* s = nullptr;
        detail::check(bf_tracker_stream_create(&c, &tc, batch_size, &ec, &s));
        s_.reset(s);
    }
    BatchedTrackerStreamV5(const BatchedTrackerStreamV5&) = delete;
    BatchedTrackerStreamV5& operator=(const BatchedTrackerStreamV5&) = delete;

    // Host to host; honours first_window_index (the windows are the absolute
    // windows first_window_index .. + batch_size - 1 of the trajectory).
    void process_batch(std::size_t first_window_index, const std::uint8_t* host_packed,
                       float* host_intensity) {
        detail::check(bf_tracker_stream_process_batch(s_.get(), first_window_index,
                                                      host_packed, host_intensity));
    }
    void process_batch_kernel_only(std::size_t first_window_index) {
        detail::check(bf_tracker_stream_process_batch_kernel_only(s_.get(), first_window_index));
    }
    float last_kernel_time_ms() const {
        float v = 0.0F;
        detail::check(bf_tracker_stream_last_kernel_time_ms(s_.get(), &v));
        return v;
    }
    std::uint8_t* device_packed_buffer() {
        std::uint8_t* p = nullptr;
        detail::check(bf_tracker_stream_device_packed_buffer(s_.get(), &p));
        return p;
    }
    float* device_intensity_buffer() {
        float* p = nullptr;
        detail::check(bf_tracker_stream_device_intensity_buffer(s_.get(), &p));
        return p;
    }
    void* device_stream() {
        void* p = nullptr;
        detail::check(bf_tracker_stream_device_stream(s_.get(), &p));
        return p;
    }
    std::size_t batch_size() const { return sizes()[0]; }
    std::size_t window_bytes() const { return sizes()[1]; }
    std::size_t batch_voltage_bytes() const { return sizes()[2]; }
    std::size_t batch_output_floats() const { return sizes()[3]; }

  private:
    std::array<std::size_t, 4> sizes() const {
        std::array<std::size_t, 4> v{};
        detail::check(bf_tracker_stream_sizes(s_.get(), &v[0], &v[1], &v[2], &v[3]));
        return v;
    }
    struct Del {
        void operator()(bf_tracker_stream* s) const noexcept { (void)bf_tracker_stream_destroy(s); }
    };
    std::unique_ptr<bf_tracker_stream, Del> s_;
};

// V2 (cuda_tracker_v2.hpp)
enum class CudaTrackerKernelV2 {
    TwoPass,
    Fused,
    WarpReduction,
};
inline void cuda_tracker_v2_packed_intensity_into(
    const PackedVoltage& packed, const Dimensions& dims, const TrackerConfig& tracker,
    Intensities& intensity, CudaTrackerKernelV2 = CudaTrackerKernelV2::TwoPass) {
    detail::tracker_into(packed, dims, tracker, intensity, false);
}
inline Intensities cuda_tracker_v2_packed_intensity(
    const PackedVoltage& packed, const Dimensions& dims, const TrackerConfig& tracker,
    CudaTrackerKernelV2 kernel = CudaTrackerKernelV2::TwoPass) {
    Intensities out;
    cuda_tracker_v2_packed_intensity_into(packed, dims, tracker, out, kernel);
    return out;
}

// FWS (cuda_beam_tracker_fused_warp_shuffle.hpp)
enum class FwsLoadStrategy {
    Direct,
    SharedMemory,
};
inline void cuda_beam_tracker_fused_warp_shuffle_into(
    const PackedVoltage& packed, const Dimensions& dims, const TrackerConfig& tracker,
    Intensities& intensity, FwsLoadStrategy = FwsLoadStrategy::Direct) {
    detail::tracker_into(packed, dims, tracker, intensity, false);
}
inline Intensities cuda_beam_tracker_fused_warp_shuffle(
    const PackedVoltage& packed, const Dimensions& dims, const TrackerConfig& tracker,
    FwsLoadStrategy strategy = FwsLoadStrategy::Direct) {
    Intensities out;
    cuda_beam_tracker_fused_warp_shuffle_into(packed, dims, tracker, out, strategy);
    return out;
}
inline void cuda_beam_tracker_fused_warp_shuffle_stream(
    const PackedVoltage& packed, const Dimensions& dims, const TrackerConfig& tracker,
    Intensities& intensity, std::size_t = 3, FwsLoadStrategy = FwsLoadStrategy::Direct) {
    detail::tracker_into(packed, dims, tracker, intensity, false);
}

// V3 (cuda_beam_tracker_v3.hpp)
struct V3ExecutionConfig {
    std::size_t time_chunk_size = 80;
    std::size_t time_unroll = 2;
    bool enable_cuda_graph = false;
};
inline void cuda_beam_tracker_v3_into(const PackedVoltage& packed, const Dimensions& dims,
                                      const TrackerConfig& tracker, Intensities& intensity,
                                      const V3ExecutionConfig& config = V3ExecutionConfig{}) {
    detail::tracker_into(packed, dims, tracker, intensity, config.enable_cuda_graph);
}
inline Intensities cuda_beam_tracker_v3(const PackedVoltage& packed, const Dimensions& dims,
                                        const TrackerConfig& tracker,
                                        const V3ExecutionConfig& config = V3ExecutionConfig{}) {
    Intensities out;
    cuda_beam_tracker_v3_into(packed, dims, tracker, out, config);
    return out;
}
inline void cuda_beam_tracker_v3_stream(const PackedVoltage& packed, const Dimensions& dims,
                                        const TrackerConfig& tracker, Intensities& intensity,
                                        std::size_t = 3,
                                        const V3ExecutionConfig& config = V3ExecutionConfig{}) {
    detail::tracker_into(packed, dims, tracker, intensity, config.enable_cuda_graph);
}
inline void cuda_beam_tracker_v3_device_resident(
    const std::uint8_t* d_packed, float* d_intensity, const Dimensions& dims,
    const TrackerConfig& tracker, const V3ExecutionConfig& config = V3ExecutionConfig{},
    void* stream = nullptr) {
    detail::tracker_device(d_packed, d_intensity, dims, tracker, config.enable_cuda_graph,
                           stream);
}

// V4 (cuda_beam_tracker_v4.hpp)
enum class V4KernelMode {
    Auto = 0,
    WmmaTensorCore = 1,
    Half2VectorSimd = 2,
    DeepIlpWarpShuffle = 3,
    BlockReduction = 4,
};
struct V4ExecutionConfig {
    V4KernelMode mode = V4KernelMode::Auto;
    std::size_t time_chunk_size = 80;
    std::size_t time_unroll = 4;
    bool enable_cuda_graph = false;
    bool enable_async_copy = true;
};
inline void cuda_beam_tracker_v4_into(const PackedVoltage& packed, const Dimensions& dims,
                                      const TrackerConfig& tracker, Intensities& intensity,
                                      const V4ExecutionConfig& config = V4ExecutionConfig{}) {
    detail::tracker_into(packed, dims, tracker, intensity, config.enable_cuda_graph);
}
inline Intensities cuda_beam_tracker_v4(const PackedVoltage& packed, const Dimensions& dims,
                                        const TrackerConfig& tracker,
                                        const V4ExecutionConfig& config = V4ExecutionConfig{}) {
    Intensities out;
    cuda_beam_tracker_v4_into(packed, dims, tracker, out, config);
    return out;
}
inline void cuda_beam_tracker_v4_stream(const PackedVoltage& packed, const Dimensions& dims,
                                        const TrackerConfig& tracker, Intensities& intensity,
                                        std::size_t = 3,
                                        const V4ExecutionConfig& config = V4ExecutionConfig{}) {
    detail::tracker_into(packed, dims, tracker, intensity, config.enable_cuda_graph);
}
inline void cuda_beam_tracker_v4_device_resident(
    const std::uint8_t* d_packed, float* d_intensity, const Dimensions& dims,
    const TrackerConfig& tracker, const V4ExecutionConfig& config = V4ExecutionConfig{},
    void* stream = nullptr) {
    detail::tracker_device(d_packed, d_intensity, dims, tracker, config.enable_cuda_graph,
                           stream);
}

}  // namespace beamformer
