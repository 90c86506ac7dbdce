// capi_io.cpp -- extern "C" on-disk formats of the beamformer: packed
// voltage, canonical and tiled weights, f32 intensities, int8 codes and
// their {offset, scale} parameters, the quantized-output metadata and the
// packed shard (payload + loss mask + metadata).  Host code only.
//
// Byte-compatible with the reference's files so that its tools and Python
// validation can exchange data with this build:
//   raw little-endian arrays with exact size checks   src/io.cpp:13-57
//   typed readers/writers                             src/io.cpp:60-146
//   beamformer-quantized-int8-v1 metadata             src/io.cpp:149-242
//   beamformer-packed-shard-v1 metadata + loss mask   src/io.cpp:244-342
//   shard descriptor and its validation               include/beamformer/config.hpp:29-72
//   quantized-metadata validation                     src/quantization.cpp:235-249
//   packed-shard validation                           src/synthetic_data.cpp:139-153
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <map>
#include <cmath>
#include <random>
#include <string>
#include <sys/stat.h>

#include "bf_common.hpp"

namespace bfhip {
namespace {

constexpr size_t kChunkTime = 1, kChunkFreq = 16, kChunkBeam = 16;  // quantization.hpp:16-18
constexpr int kInvalidCode = -128;                    // quantization.hpp:19

std::string str(const char* p, const char* what) {
    if (p == nullptr || *p == '\0') invalid(std::string(what) + " must not be empty");
    return std::string(p);
}

[[noreturn]] void io_fail(const std::string& msg) { fail(BF_ERR_RUNTIME, msg); }

// Owns a FILE*; closes it on every path.
struct File {
    FILE* f = nullptr;
    File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
    File(const File&) = delete;
    File& operator=(const File&) = delete;
};

// Reads exactly `bytes` bytes; the file must hold exactly that many.
void read_exact(const std::string& path, void* dst, size_t bytes) {
    struct stat st;
    if (::stat(path.c_str(), &st) != 0) io_fail("cannot inspect input file: " + path);
    if ((unsigned long long)st.st_size != (unsigned long long)bytes)
        io_fail(path + " has " + std::to_string((unsigned long long)st.st_size) +
                " bytes; expected " + std::to_string((unsigned long long)bytes));
    if (bytes > 0 && dst == nullptr) invalid("output buffer must not be null");
    File in(path, "rb");
    if (!in.f) io_fail("cannot open input file: " + path);
    if (bytes > 0 && std::fread(dst, 1, bytes, in.f) != bytes)
        io_fail("failed to read input file: " + path);
}

void write_all(const std::string& path, const void* src, size_t bytes) {
    if (bytes > 0 && src == nullptr) invalid("input buffer must not be null");
    File out(path, "wb");
    if (!out.f) io_fail("cannot open output file: " + path);
    if (bytes > 0 && std::fwrite(src, 1, bytes, out.f) != bytes)
        io_fail("failed to write output file: " + path);
    if (std::fflush(out.f) != 0) io_fail("failed to write output file: " + path);
}

void require_capacity(size_t have, size_t need, const char* what) {
    if (have < need)
        invalid(std::string(what) + " holds " + std::to_string(have) + " elements; " +
                std::to_string(need) + " are needed");
}

// ------------------------------------------------------- key=value text
using Fields = std::map<std::string, std::string>;

Fields read_fields(const std::string& path, const char* kind) {
    File in(path, "r");
    if (!in.f) io_fail(std::string("cannot open ") + kind + " metadata file: " + path);
    Fields fields;
    std::string line;
    int c;
    bool any = false;
    auto take = [&]() {
        const size_t sep = line.find('=');
        if (sep == std::string::npos || sep == 0)
            io_fail(std::string("invalid ") + kind + " metadata line in: " + path);
        fields[line.substr(0, sep)] = line.substr(sep + 1);
        line.clear();
    };
    while ((c = std::fgetc(in.f)) != EOF) {
        any = true;
        if (c == '\n')
            take();
        else
            line.push_back((char)c);
    }
    if (any && !line.empty()) take();  // a last line without its newline
    return fields;
}

struct FieldReader {
    const Fields& fields;
    const std::string& path;
    const char* kind;
    const std::string& text(const char* name) const {
        const auto it = fields.find(name);
        if (it == fields.end())
            io_fail(std::string("missing ") + kind + " metadata field " + name + " in: " + path);
        return it->second;
    }
    unsigned long long number(const char* name) const {
        const std::string& v = text(name);
        // std::stoull semantics: leading whitespace and a prefix number
        errno = 0;
        char* end = nullptr;
        const unsigned long long x = std::strtoull(v.c_str(), &end, 10);
        if (end == v.c_str()) invalid(std::string("stoull"));
        if (errno == ERANGE) fail(BF_ERR_OUT_OF_RANGE, "stoull");
        return x;
    }
    bool flag(const char* name, const std::string& msg) const {
        const std::string& v = text(name);
        if (v != "0" && v != "1") io_fail(msg);
        return v == "1";
    }
};

bf_shard descriptor_from(const FieldReader& r, const char* flag_msg) {
    bf_shard d{};
    d.shard_id = (size_t)r.number("shard_id");
    d.shard_count = (size_t)r.number("shard_count");
    d.local_frequency_count = (size_t)r.number("local_frequency_count");
    d.absolute_frequency_start = (size_t)r.number("absolute_frequency_start");
    d.timestamp_start = (uint64_t)r.number("timestamp_start");
    d.timestamp_step = (uint64_t)r.number("timestamp_step");
    d.loss_mask_id = (uint64_t)r.number("loss_mask_id");
    d.loss_mask_independent = r.flag("loss_mask_independent", flag_msg) ? 1 : 0;
    return d;
}

std::string descriptor_lines(const bf_shard& d) {
    std::string s;
    s += "shard_id=" + std::to_string(d.shard_id) + "\n";
    s += "shard_count=" + std::to_string(d.shard_count) + "\n";
    s += "local_frequency_count=" + std::to_string(d.local_frequency_count) + "\n";
    s += "absolute_frequency_start=" + std::to_string(d.absolute_frequency_start) + "\n";
    s += "timestamp_start=" + std::to_string((unsigned long long)d.timestamp_start) + "\n";
    s += "timestamp_step=" + std::to_string((unsigned long long)d.timestamp_step) + "\n";
    s += "loss_mask_id=" + std::to_string((unsigned long long)d.loss_mask_id) + "\n";
    s += std::string("loss_mask_independent=") + (d.loss_mask_independent ? "1" : "0") + "\n";
    return s;
}

std::string dims_lines(const bf_dims& d) {
    return "n_time=" + std::to_string(d.n_time) + "\nn_freq=" + std::to_string(d.n_freq) +
           "\nn_ant=" + std::to_string(d.n_ant) + "\nn_beams=" + std::to_string(d.n_beams) +
           "\n";
}

void write_text(const std::string& path, const std::string& body, const char* kind) {
    File out(path, "w");
    if (!out.f) io_fail(std::string("cannot open ") + kind + " metadata file: " + path);
    if (std::fwrite(body.data(), 1, body.size(), out.f) != body.size() || std::fflush(out.f) != 0)
        io_fail(std::string("failed to write ") + kind + " metadata file: " + path);
}

// Reference src/quantization.cpp:235-249.
void validate_metadata(const bf_quantized_output_metadata& m) {
    validate_dimensions(m.output_dims);
    if (m.integration_spectra == 0) invalid("integration_spectra must be positive");
    validate_shard(m.shard);
    if (m.input_n_time == 0) invalid("quantized metadata input_n_time must be positive");
    if (integrated_time_count(m.input_n_time, m.integration_spectra) != m.output_dims.n_time)
        invalid("quantized metadata time dimensions are inconsistent");
    if (m.parameters_file[0] == '\0')
        invalid("quantized metadata parameters_file must not be empty");
    if (std::memchr(m.parameters_file, '\0', sizeof(m.parameters_file)) == nullptr)
        invalid("quantized metadata parameters_file is not terminated");
}

// Reference src/synthetic_data.cpp:139-153.
void validate_shard_payload(const bf_shard& s, const uint8_t* mask, size_t payload,
                            size_t mask_bytes, const bf_dims& d) {
    validate_dimensions(d);
    validate_shard(s);
    if (payload != d.n_time * d.n_freq * d.n_ant)
        invalid("shard payload size does not match dimensions");
    if (mask_bytes != d.n_time * d.n_freq) invalid("loss mask size does not match dimensions");
    for (size_t i = 0; i < mask_bytes; ++i)
        if (mask[i] > 1u) invalid("loss mask entries must be 0 or 1");
}

const bf_shard& deref(const bf_shard* p) {
    if (p == nullptr) invalid("shard descriptor must not be null");
    return *p;
}

}  // namespace
}  // namespace bfhip

using namespace bfhip;

extern "C" {

bf_status bf_default_shard_descriptor(size_t shard_id, bf_shard* out) {
    return guarded([&] {
        if (out == nullptr) invalid("out must not be null");
        if (shard_id >= kFrequencyShardCount) invalid("shard_id is outside the two-shard band");
        // config.hpp:65-72: shard s starts at channel 336 s, loss mask id s
        *out = bf_shard{shard_id, kFrequencyShardCount, kLocalFrequencyChannels,
                                   shard_id * kLocalFrequencyChannels, 0, 1, shard_id, 1};
    });
}

bf_status bf_validate_shard_descriptor(const bf_shard* shard) {
    return guarded([&] { validate_shard(deref(shard)); });
}

bf_status bf_read_packed_voltage(const char* path, const bf_dims* dims, uint8_t* out,
                                 size_t out_bytes) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        const size_t n = d.n_time * d.n_freq * d.n_ant;
        require_capacity(out_bytes, n, "out");
        read_exact(str(path, "path"), out, n);
    });
}

bf_status bf_write_packed_voltage(const char* path, const uint8_t* voltage, size_t bytes,
                                  const bf_dims* dims) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        if (bytes != d.n_time * d.n_freq * d.n_ant)
            invalid("packed voltage size does not match dimensions");
        write_all(str(path, "path"), voltage, bytes);
    });
}

bf_status bf_read_weights(const char* path, const bf_dims* dims, float* out, size_t out_complex) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        const size_t n = d.n_beams * d.n_freq * d.n_ant;
        require_capacity(out_complex, n, "out");
        read_exact(str(path, "path"), out, n * 2 * sizeof(float));
    });
}

bf_status bf_write_weights(const char* path, const float* weights, size_t n_complex,
                           const bf_dims* dims) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        if (n_complex != d.n_beams * d.n_freq * d.n_ant)
            invalid("weight count does not match dimensions");
        write_all(str(path, "path"), weights, n_complex * 2 * sizeof(float));
    });
}

bf_status bf_read_tiled_weights(const char* path, const bf_dims* dims, float* out,
                                size_t out_complex) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        const size_t n = tiled_weight_count(d);
        require_capacity(out_complex, n, "out");
        read_exact(str(path, "path"), out, n * 2 * sizeof(float));
    });
}

bf_status bf_write_tiled_weights(const char* path, const float* weights, size_t n_complex,
                                 const bf_dims* dims) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        if (n_complex != tiled_weight_count(d))
            invalid("tiled weight count does not match dimensions");
        write_all(str(path, "path"), weights, n_complex * 2 * sizeof(float));
    });
}

bf_status bf_write_intensities(const char* path, const float* intensity, size_t count,
                               const bf_dims* dims) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        if (count != d.n_time * d.n_freq * d.n_beams)
            invalid("intensity count does not match dimensions");
        write_all(str(path, "path"), intensity, count * sizeof(float));
    });
}

bf_status bf_read_quantized_intensities(const char* path, const bf_dims* dims, int8_t* out,
                                        size_t out_count) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        const size_t n = d.n_time * d.n_freq * d.n_beams;
        require_capacity(out_count, n, "out");
        read_exact(str(path, "path"), out, n);
    });
}

bf_status bf_write_quantized_intensities(const char* path, const int8_t* codes, size_t count,
                                         const bf_dims* dims) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        if (count != d.n_time * d.n_freq * d.n_beams)
            invalid("quantized intensity count does not match dimensions");
        write_all(str(path, "path"), codes, count);
    });
}

bf_status bf_read_quantization_parameters(const char* path, const bf_dims* integrated_dims,
                                          float* out, size_t out_params) {
    return guarded([&] {
        const size_t n = quantization_parameter_count(deref_dims(integrated_dims));
        require_capacity(out_params, n, "out");
        read_exact(str(path, "path"), out, n * 2 * sizeof(float));
    });
}

bf_status bf_write_quantization_parameters(const char* path, const float* params,
                                           size_t n_params, const bf_dims* integrated_dims) {
    return guarded([&] {
        if (n_params != quantization_parameter_count(deref_dims(integrated_dims)))
            invalid("quantization parameter count does not match dimensions");
        write_all(str(path, "path"), params, n_params * 2 * sizeof(float));
    });
}

bf_status bf_write_quantized_output_metadata(const char* path,
                                             const bf_quantized_output_metadata* metadata) {
    return guarded([&] {
        if (metadata == nullptr) invalid("metadata must not be null");
        const bf_quantized_output_metadata& m = *metadata;
        validate_metadata(m);
        std::string s = "format=beamformer-quantized-int8-v1\n" + dims_lines(m.output_dims);
        s += "input_n_time=" + std::to_string(m.input_n_time) + "\n";
        s += "integration_spectra=" + std::to_string(m.integration_spectra) + "\n";
        s += "chunk_time=" + std::to_string(kChunkTime) + "\n";
        s += "chunk_frequency=" + std::to_string(kChunkFreq) + "\n";
        s += "chunk_beam=" + std::to_string(kChunkBeam) + "\n";
        s += "invalid_code=" + std::to_string(kInvalidCode) + "\n";
        s += "parameters_format=little-endian-float32[offset,scale]\n";
        s += std::string("parameters_file=") + m.parameters_file + "\n";
        s += "parameters_count=" +
             std::to_string(quantization_parameter_count(m.output_dims)) + "\n";
        s += descriptor_lines(m.shard);
        write_text(str(path, "path"), s, "quantized");
    });
}

bf_status bf_read_quantized_output_metadata(const char* path,
                                            bf_quantized_output_metadata* out) {
    return guarded([&] {
        if (out == nullptr) invalid("out must not be null");
        const std::string p = str(path, "path");
        const Fields fields = read_fields(p, "quantized");
        const FieldReader r{fields, p, "quantized"};
        if (r.text("format") != "beamformer-quantized-int8-v1" ||
            r.text("parameters_format") != "little-endian-float32[offset,scale]" ||
            r.number("chunk_time") != kChunkTime || r.number("chunk_frequency") != kChunkFreq ||
            r.number("chunk_beam") != kChunkBeam ||
            std::stoi(r.text("invalid_code")) != kInvalidCode)
            io_fail("unsupported quantized metadata format: " + p);
        bf_quantized_output_metadata m{};
        m.output_dims = bf_dims{(size_t)r.number("n_time"), (size_t)r.number("n_freq"),
                                (size_t)r.number("n_ant"), (size_t)r.number("n_beams")};
        m.input_n_time = (size_t)r.number("input_n_time");
        m.integration_spectra = (size_t)r.number("integration_spectra");
        m.shard = descriptor_from(r, "invalid loss_mask_independent metadata value");
        const std::string& file = r.text("parameters_file");
        if (file.size() >= sizeof(m.parameters_file))
            invalid("parameters_file is longer than the metadata struct holds");
        std::memcpy(m.parameters_file, file.c_str(), file.size() + 1);
        if (r.number("parameters_count") != quantization_parameter_count(m.output_dims))
            io_fail("quantization parameter count is inconsistent: " + p);
        validate_metadata(m);
        *out = m;
    });
}

bf_status bf_write_packed_shard(const char* payload_path, const char* metadata_path,
                                const char* loss_mask_path, const bf_shard* shard,
                                const uint8_t* payload, size_t payload_bytes,
                                const uint8_t* loss_mask, size_t loss_mask_bytes,
                                const bf_dims* dims) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        const bf_shard& s = deref(shard);
        if (loss_mask_bytes > 0 && loss_mask == nullptr) invalid("loss_mask must not be null");
        validate_shard_payload(s, loss_mask, payload_bytes, loss_mask_bytes, d);
        write_all(str(payload_path, "payload_path"), payload, payload_bytes);
        write_all(str(loss_mask_path, "loss_mask_path"), loss_mask, loss_mask_bytes);
        std::string m = "format=beamformer-packed-shard-v1\n" + dims_lines(d) +
                        descriptor_lines(s);
        m += "payload_bytes=" + std::to_string(payload_bytes) + "\n";
        m += "loss_mask_bytes=" + std::to_string(loss_mask_bytes) + "\n";
        write_text(str(metadata_path, "metadata_path"), m, "shard");
    });
}

bf_status bf_read_packed_shard(const char* payload_path, const char* metadata_path,
                               const char* loss_mask_path, const bf_dims* dims,
                               bf_shard* shard_out, uint8_t* payload,
                               size_t payload_capacity, uint8_t* loss_mask,
                               size_t loss_mask_capacity) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        if (shard_out == nullptr) invalid("shard_out must not be null");
        const std::string mp = str(metadata_path, "metadata_path");
        const Fields fields = read_fields(mp, "shard");
        const FieldReader r{fields, mp, "shard"};
        const size_t payload_bytes = d.n_time * d.n_freq * d.n_ant;
        const size_t mask_bytes = d.n_time * d.n_freq;
        if (r.text("format") != "beamformer-packed-shard-v1" || r.number("n_time") != d.n_time ||
            r.number("n_freq") != d.n_freq || r.number("n_ant") != d.n_ant ||
            r.number("n_beams") != d.n_beams || r.number("payload_bytes") != payload_bytes ||
            r.number("loss_mask_bytes") != mask_bytes)
            io_fail("shard metadata does not match dimensions: " + mp);
        const bf_shard s = descriptor_from(
            r, "invalid boolean shard metadata field: loss_mask_independent");
        require_capacity(payload_capacity, payload_bytes, "payload");
        require_capacity(loss_mask_capacity, mask_bytes, "loss_mask");
        read_exact(str(payload_path, "payload_path"), payload, payload_bytes);
        read_exact(str(loss_mask_path, "loss_mask_path"), loss_mask, mask_bytes);
        validate_shard_payload(s, loss_mask, payload_bytes, mask_bytes, d);
        *shard_out = s;
    });
}

bf_status bf_validate_packed_shard(const bf_shard* shard, const uint8_t* payload,
                                   size_t payload_bytes, const uint8_t* loss_mask,
                                   size_t loss_mask_bytes, const bf_dims* dims) {
    return guarded([&] {
        if (payload == nullptr && payload_bytes != 0) invalid("payload must not be null");
        if (loss_mask == nullptr && loss_mask_bytes != 0) invalid("loss_mask must not be null");
        // synthetic_data.cpp:139-153
        validate_shard_payload(deref(shard), loss_mask, payload_bytes, loss_mask_bytes,
                               deref_dims(dims));
    });
}

bf_status bf_validate_packed_shards(const bf_shard* shards, const uint8_t* const* payloads,
                                    size_t payload_bytes, const uint8_t* const* loss_masks,
                                    size_t loss_mask_bytes, const bf_dims* dims) {
    return guarded([&] {
        if (shards == nullptr || payloads == nullptr || loss_masks == nullptr)
            invalid("shard arrays must not be null");
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        // synthetic_data.cpp:155-171
        for (size_t i = 0; i < kFrequencyShardCount; ++i) {
            if (payloads[i] == nullptr || loss_masks[i] == nullptr)
                invalid("shard buffers must not be null");
            validate_shard_payload(shards[i], loss_masks[i], payload_bytes, loss_mask_bytes, d);
            if (shards[i].shard_id != i) invalid("shard IDs must identify their array slot");
        }
        if (shards[0].absolute_frequency_start == shards[1].absolute_frequency_start)
            invalid("the two shards must have distinct frequency origins");
        if (shards[0].loss_mask_id == shards[1].loss_mask_id)
            invalid("the two shards must have independent loss masks");
    });
}

bf_status bf_make_loss_mask(const bf_dims* dims, uint32_t seed, float loss_probability,
                            uint8_t* out, size_t n) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        validate_dimensions(d);
        // make_loss_mask (synthetic_data.cpp:121-137): the same engine and
        // distribution of the C++ standard library, one draw per frame
        if (!std::isfinite(loss_probability) || loss_probability < 0.0f ||
            loss_probability > 1.0f)
            invalid("loss probability must be in [0, 1]");
        if (out == nullptr) invalid("out must not be null");
        if (n < d.n_time * d.n_freq) invalid("loss mask output is smaller than dimensions");
        std::mt19937 random(seed);
        std::bernoulli_distribution lost(loss_probability);
        for (size_t i = 0; i < d.n_time * d.n_freq; ++i) out[i] = lost(random) ? 0u : 1u;
    });
}

bf_status bf_shard_frequencies(const bf_shard* shard, float start_hz, float channel_width_hz,
                               float* out) {
    return guarded([&] {
        const bf_shard& s = deref(shard);
        validate_shard(s);
        if (out == nullptr) invalid("out must not be null");
        if (!std::isfinite(start_hz) || start_hz <= 0.0f || !std::isfinite(channel_width_hz) ||
            channel_width_hz <= 0.0f)
            invalid("frequency start and channel width must be positive");
        // channel absolute_frequency_start + k of the full-band plan
        // (channelized_frequencies over both shards, geometry.cpp:79-95)
        for (size_t k = 0; k < s.local_frequency_count; ++k)
            out[k] = start_hz + (float)(s.absolute_frequency_start + k) * channel_width_hz;
    });
}

}  // extern "C"
