// amd_io.hpp -- header-only C++ layer over the on-disk format functions of
// bf_capi.h, with the reference's names, types and exceptions
// (include/beamformer/io.hpp, formats.hpp:13-77, config.hpp:29-72,
// quantization.hpp:75-81).  IO failures throw std::runtime_error, size and
// contract violations std::invalid_argument, as src/io.cpp does.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "amd.hpp"

namespace beamformer {
namespace amd {

using PackedVoltage = std::vector<std::uint8_t>;
using Weights = std::vector<float>;       // interleaved complex f32 [b][f][a]
using TiledWeights = std::vector<float>;  // interleaved complex f32 [f][bt][a][32]
using Intensities = std::vector<float>;
using QuantizedIntensities = std::vector<std::int8_t>;
using QuantizationParameters = std::vector<float>;  // {offset, scale} pairs
using LossMask = std::vector<std::uint8_t>;

struct ShardDescriptor {
    std::size_t shard_id = 0;
    std::size_t shard_count = 2;
    std::size_t local_frequency_count = 336;
    std::size_t absolute_frequency_start = 0;
    std::uint64_t timestamp_start = 0;
    std::uint64_t timestamp_step = 1;
    std::uint64_t loss_mask_id = 0;
    bool loss_mask_independent = true;
    bf_shard c() const {
        return bf_shard{shard_id,        shard_count,    local_frequency_count,
                        absolute_frequency_start, timestamp_start, timestamp_step,
                        loss_mask_id,    loss_mask_independent ? 1 : 0};
    }
    static ShardDescriptor from(const bf_shard& s) {
        return ShardDescriptor{s.shard_id,        s.shard_count,    s.local_frequency_count,
                               s.absolute_frequency_start, s.timestamp_start,
                               s.timestamp_step,  s.loss_mask_id,   s.loss_mask_independent != 0};
    }
};

inline std::array<ShardDescriptor, 2> default_shard_descriptors() {
    std::array<ShardDescriptor, 2> out;
    for (std::size_t i = 0; i < 2; ++i) {
        bf_shard s{};
        check(bf_default_shard_descriptor(i, &s));
        out[i] = ShardDescriptor::from(s);
    }
    return out;
}

inline void validate_shard_descriptor(const ShardDescriptor& shard) {
    const bf_shard s = shard.c();
    check(bf_validate_shard_descriptor(&s));
}

struct PackedShard {
    ShardDescriptor descriptor;
    PackedVoltage payload;
    LossMask loss_mask;
};

struct QuantizedOutputMetadata {
    Dimensions output_dims;
    std::size_t input_n_time = 0;
    TemporalIntegrationConfig temporal_integration{0};
    ShardDescriptor shard;
    std::string parameters_file;
};

inline std::size_t tiled_weight_count(const Dimensions& d) {
    return d.n_freq * ((d.n_beams + 31) / 32) * d.n_ant * 32;
}

inline PackedVoltage read_packed_voltage(const std::filesystem::path& path,
                                         const Dimensions& dims) {
    const bf_dims c = dims.c();
    PackedVoltage out(dims.n_time * dims.n_freq * dims.n_ant);
    check(bf_read_packed_voltage(path.c_str(), &c, out.data(), out.size()));
    return out;
}

inline void write_packed_voltage(const std::filesystem::path& path, const PackedVoltage& v,
                                 const Dimensions& dims) {
    const bf_dims c = dims.c();
    check(bf_write_packed_voltage(path.c_str(), v.data(), v.size(), &c));
}

inline Weights read_weights(const std::filesystem::path& path, const Dimensions& dims) {
    const bf_dims c = dims.c();
    Weights out(2 * dims.n_beams * dims.n_freq * dims.n_ant);
    check(bf_read_weights(path.c_str(), &c, out.data(), out.size() / 2));
    return out;
}

inline void write_weights(const std::filesystem::path& path, const Weights& w,
                          const Dimensions& dims) {
    const bf_dims c = dims.c();
    check(bf_write_weights(path.c_str(), w.data(), w.size() / 2, &c));
}

inline TiledWeights read_tiled_weights(const std::filesystem::path& path,
                                       const Dimensions& dims) {
    const bf_dims c = dims.c();
    TiledWeights out(2 * tiled_weight_count(dims));
    check(bf_read_tiled_weights(path.c_str(), &c, out.data(), out.size() / 2));
    return out;
}

inline void write_tiled_weights(const std::filesystem::path& path, const TiledWeights& w,
                                const Dimensions& dims) {
    const bf_dims c = dims.c();
    check(bf_write_tiled_weights(path.c_str(), w.data(), w.size() / 2, &c));
}

inline void write_intensities(const std::filesystem::path& path, const Intensities& i,
                              const Dimensions& dims) {
    const bf_dims c = dims.c();
    check(bf_write_intensities(path.c_str(), i.data(), i.size(), &c));
}

inline QuantizedIntensities read_quantized_intensities(const std::filesystem::path& path,
                                                       const Dimensions& dims) {
    const bf_dims c = dims.c();
    QuantizedIntensities out(dims.n_time * dims.n_freq * dims.n_beams);
    check(bf_read_quantized_intensities(path.c_str(), &c, out.data(), out.size()));
    return out;
}

inline void write_quantized_intensities(const std::filesystem::path& path,
                                        const QuantizedIntensities& codes,
                                        const Dimensions& dims) {
    const bf_dims c = dims.c();
    check(bf_write_quantized_intensities(path.c_str(), codes.data(), codes.size(), &c));
}

inline QuantizationParameters read_quantization_parameters(const std::filesystem::path& path,
                                                           const Dimensions& integrated) {
    const bf_dims c = integrated.c();
    QuantizationParameters out(2 * quantization_parameter_count(integrated));
    check(bf_read_quantization_parameters(path.c_str(), &c, out.data(), out.size() / 2));
    return out;
}

inline void write_quantization_parameters(const std::filesystem::path& path,
                                          const QuantizationParameters& p,
                                          const Dimensions& integrated) {
    const bf_dims c = integrated.c();
    check(bf_write_quantization_parameters(path.c_str(), p.data(), p.size() / 2, &c));
}

inline void write_quantized_output_metadata(const std::filesystem::path& path,
                                            const QuantizedOutputMetadata& m) {
    bf_quantized_output_metadata c{};
    c.output_dims = m.output_dims.c();
    c.input_n_time = m.input_n_time;
    c.integration_spectra = m.temporal_integration.integration_spectra;
    c.shard = m.shard.c();
    if (m.parameters_file.size() >= sizeof(c.parameters_file))
        throw std::invalid_argument("parameters_file name is too long");
    std::memcpy(c.parameters_file, m.parameters_file.c_str(), m.parameters_file.size() + 1);
    check(bf_write_quantized_output_metadata(path.c_str(), &c));
}

inline QuantizedOutputMetadata read_quantized_output_metadata(
    const std::filesystem::path& path) {
    bf_quantized_output_metadata c{};
    check(bf_read_quantized_output_metadata(path.c_str(), &c));
    QuantizedOutputMetadata m;
    m.output_dims = Dimensions{c.output_dims.n_time, c.output_dims.n_freq, c.output_dims.n_ant,
                               c.output_dims.n_beams};
    m.input_n_time = c.input_n_time;
    m.temporal_integration.integration_spectra = c.integration_spectra;
    m.shard = ShardDescriptor::from(c.shard);
    m.parameters_file = c.parameters_file;
    return m;
}

inline void write_packed_shard(const std::filesystem::path& payload_path,
                               const std::filesystem::path& metadata_path,
                               const std::filesystem::path& loss_mask_path,
                               const PackedShard& shard, const Dimensions& dims) {
    const bf_dims c = dims.c();
    const bf_shard s = shard.descriptor.c();
    check(bf_write_packed_shard(payload_path.c_str(), metadata_path.c_str(),
                                loss_mask_path.c_str(), &s, shard.payload.data(),
                                shard.payload.size(), shard.loss_mask.data(),
                                shard.loss_mask.size(), &c));
}

inline PackedShard read_packed_shard(const std::filesystem::path& payload_path,
                                     const std::filesystem::path& metadata_path,
                                     const std::filesystem::path& loss_mask_path,
                                     const Dimensions& dims) {
    const bf_dims c = dims.c();
    PackedShard out;
    out.payload.resize(dims.n_time * dims.n_freq * dims.n_ant);
    out.loss_mask.resize(dims.n_time * dims.n_freq);
    bf_shard s{};
    check(bf_read_packed_shard(payload_path.c_str(), metadata_path.c_str(),
                               loss_mask_path.c_str(), &c, &s, out.payload.data(),
                               out.payload.size(), out.loss_mask.data(), out.loss_mask.size()));
    out.descriptor = ShardDescriptor::from(s);
    return out;
}

}  // namespace amd
}  // namespace beamformer
