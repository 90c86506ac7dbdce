// beamformer_amd -- fixed-grid beamforming of a packed voltage file on the
// GPU.  It answers to the flags of the reference's GPU tool
// (tools/beamformer_cuda.cpp) and writes the same files: f32 intensities,
// or int8 codes with their {offset, scale} parameters and the key=value
// metadata file; an optional metrics CSV gains one row per run with the
// reference's columns (backend "hip").  Kernels: direct and tiled (FP32 on
// canonical or tiled weights, as in the reference), split-f16 and f16
// (MFMA on canonical weights), auto.
//
// A run is one Job: the option table fills it, validate() checks it before
// any file is read, and execute() produces an Outcome that the report and
// the CSV row are printed from.
#include <cstddef>
#include <exception>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "beamformer/amd.hpp"
#include "beamformer/amd_io.hpp"
#include "cli_common.hpp"

namespace {

namespace bfa = beamformer::amd;
namespace fs = std::filesystem;

enum class Format { Float32, Int8 };

struct Job {
    fs::path input, weights, output;
    std::optional<fs::path> metrics, params_file, metadata_file;
    std::size_t n_time = 15360, n_ant = 64, n_beams = 5, shard = 0;
    std::optional<std::size_t> integrate;
    Format format = Format::Float32;
    bfa::Kernel kernel = bfa::Kernel::Direct;

    bfa::Dimensions dims() const { return {n_time, 336, n_ant, n_beams}; }
    bfa::Dimensions output_dims() const {
        bfa::Dimensions d = dims();
        if (integrate) d.n_time = bfa::integrated_time_count(d.n_time, *integrate);
        return d;
    }
};

const std::vector<std::pair<std::string, bfa::Kernel>> kKernels = {
    {"direct", bfa::Kernel::Direct}, {"tiled", bfa::Kernel::Tiled},
    {"split-f16", bfa::Kernel::SplitF16}, {"f16", bfa::Kernel::F16}, {"auto", bfa::Kernel::Auto}};

std::string kernel_label(bfa::Kernel k) {
    for (const auto& [name, value] : kKernels)
        if (value == k) return name;
    return "fp32";
}

bfcli::OptionTable options_for(Job& j) {
    using namespace bfcli;
    OptionTable t;
    t.add("--input", "FILE", "packed int4 voltages [T][336][A]", path_into(j.input), true)
        .add("--weights", "FILE", "complex f32 weights (tiled layout for --kernel tiled)",
             path_into(j.weights), true)
        .add("--output", "FILE", "f32 intensities, or int8 codes", path_into(j.output), true)
        .add("--n-time", "N", "spectra; default 15360", size_into(j.n_time))
        .add("--n-ant", "N", "32, 64, 128 or 256; default 64", size_into(j.n_ant))
        .add("--n-beams", "N", "1 to 128; default 5", size_into(j.n_beams))
        .add("--kernel", "NAME", "direct, tiled, split-f16, f16 or auto; default direct",
             choice_into(j.kernel, kKernels))
        .add("--integration-spectra", "N", "integrate N spectra per output row",
             size_into(j.integrate))
        .add("--output-format", "NAME", "float32 or int8; default float32",
             choice_into(j.format, std::vector<std::pair<std::string, Format>>{
                                       {"float32", Format::Float32}, {"int8", Format::Int8}}))
        .add("--quantization-params", "FILE", "int8 {offset, scale} output (int8 only)",
             path_into(j.params_file))
        .add("--quantization-metadata", "FILE", "int8 key=value metadata output (int8 only)",
             path_into(j.metadata_file))
        .add("--shard-id", "N", "0 or 1, recorded in the int8 metadata; default 0",
             size_into(j.shard))
        .add("--metrics", "FILE", "append one timing row to this CSV", path_into(j.metrics))
        .help_flags({"--help"});
    return t;
}

void validate(const Job& j, const bfcli::OptionTable& t) {
    const std::string missing = t.missing_required();
    if (!missing.empty()) throw std::invalid_argument("required: " + missing);
    if (j.format == Format::Int8) {
        if (!j.integrate) throw std::invalid_argument("int8 output needs --integration-spectra");
        if (!j.params_file || !j.metadata_file)
            throw std::invalid_argument(
                "int8 output needs --quantization-params and --quantization-metadata");
    }
    if (j.shard > 1) throw std::invalid_argument("--shard-id must be 0 or 1");
    bfa::validate_dimensions(j.dims());
}

struct Outcome {
    std::size_t elements = 0, bytes = 0;
    std::optional<std::size_t> strongest_beam;  // f32 output only
    bf_timings gpu{};
    double load_ms = 0, compute_ms = 0, write_ms = 0, total_ms = 0;
};

// Index of the beam whose intensity summed over every (row, channel) is
// largest; rows of the [..][B] output are B floats each.
std::size_t strongest_beam(const std::vector<float>& rows, std::size_t B) {
    std::vector<double> power(B, 0.0);
    for (std::size_t r = 0; r + B <= rows.size(); r += B)
        for (std::size_t b = 0; b < B; ++b) power[b] += rows[r + b];
    std::size_t best = 0;
    for (std::size_t b = 1; b < B; ++b)
        if (power[b] > power[best]) best = b;
    return best;
}

Outcome execute(const Job& j) {
    Outcome o;
    bfcli::Laps laps;
    const bfa::Dimensions d = j.dims();
    const bfa::PackedVoltage voltages = bfa::read_packed_voltage(j.input, d);
    const std::vector<float> w = j.kernel == bfa::Kernel::Tiled ? bfa::read_tiled_weights(j.weights, d)
                                                                : bfa::read_weights(j.weights, d);
    o.load_ms = laps.lap();
    const bfa::Dimensions od = j.output_dims();
    if (j.format == Format::Int8) {
        const bfa::TemporalIntegrationConfig integ{*j.integrate};
        const auto q = bfa::beamform_packed_quantized_integrated_intensity(voltages, w, d, integ,
                                                                           j.kernel, &o.gpu);
        o.compute_ms = laps.lap();
        bfa::write_quantized_intensities(j.output, q.codes, od);
        bfa::write_quantization_parameters(*j.params_file, q.parameters, od);
        bfa::QuantizedOutputMetadata meta;
        meta.output_dims = od;
        meta.input_n_time = d.n_time;
        meta.temporal_integration = integ;
        meta.shard = bfa::default_shard_descriptors()[j.shard];
        meta.parameters_file = j.params_file->filename().string();
        bfa::write_quantized_output_metadata(*j.metadata_file, meta);
        o.elements = q.codes.size();
        o.bytes = q.codes.size() + q.parameters.size() * sizeof(float);
    } else {
        std::vector<float> out;
        if (j.integrate)
            out = bfa::beamform_packed_integrated_intensity(
                voltages, w, d, bfa::TemporalIntegrationConfig{*j.integrate}, j.kernel, &o.gpu);
        else
            out = bfa::beamform_packed_intensity(voltages, w, d, j.kernel, &o.gpu);
        o.compute_ms = laps.lap();
        bfa::write_intensities(j.output, out, od);
        o.elements = out.size();
        o.bytes = out.size() * sizeof(float);
        o.strongest_beam = strongest_beam(out, od.n_beams);
    }
    o.write_ms = laps.lap();
    o.total_ms = laps.total();
    return o;
}

// Device time of the beamforming stages (kernel, integration, quantiser).
double device_ms(const Outcome& o) {
    return o.gpu.kernel_ms + o.gpu.temporal_integration_ms + o.gpu.quantization_ms;
}

void report(const Job& j, const Outcome& o, std::ostream& out) {
    const bfa::Dimensions d = j.dims(), od = j.output_dims();
    const double s = device_ms(o) / 1e3;
    const double elems_per_s = s > 0 ? double(o.elements) / s : 0.0;
    const double gmac_per_s =
        s > 0 ? double(d.n_time) * double(d.n_freq) * double(d.n_beams) * double(d.n_ant) / s / 1e9
              : 0.0;
    if (j.metrics) {
        bfcli::CsvAppender csv(*j.metrics);
        csv.column("backend", "hip").column("n_time", d.n_time).column("n_freq", d.n_freq)
            .column("n_ant", d.n_ant).column("n_beams", d.n_beams).column("load_ms", o.load_ms)
            .column("unpack_ms", 0.0).column("setup_ms", o.gpu.setup_ms)
            .column("host_to_device_ms", o.gpu.host_to_device_ms).column("compute_ms", device_ms(o))
            .column("device_to_host_ms", o.gpu.device_to_host_ms).column("write_ms", o.write_ms)
            .column("total_ms", o.total_ms).column("output_elements_per_second", elems_per_s)
            .column("complex_gmac_per_second", gmac_per_s)
            .append();
    }
    auto layout = [](const bfa::Dimensions& x) {
        std::ostringstream s;
        s << "[T=" << x.n_time << "][F=" << x.n_freq << "][B=" << x.n_beams << "]";
        return s.str();
    };
    out << std::fixed << std::setprecision(3) << "HIP beamforming complete: kernel="
        << kernel_label(j.kernel) << " input_layout=" << layout(d)
        << " output_layout=" << layout(od) << "\n";
    const std::vector<std::pair<const char*, double>> times = {
        {"load_ms", o.load_ms},
        {"setup_ms", o.gpu.setup_ms},
        {"host_to_device_ms", o.gpu.host_to_device_ms},
        {"kernel_ms", device_ms(o)},
        {"beamformer_ms", o.gpu.kernel_ms},
        {"temporal_integration_ms", o.gpu.temporal_integration_ms},
        {"quantization_ms", o.gpu.quantization_ms},
        {"device_to_host_ms", o.gpu.device_to_host_ms},
        {"write_ms", o.write_ms},
        {"total_ms", o.total_ms}};
    for (std::size_t i = 0; i < times.size(); ++i)
        out << (i ? " " : "") << times[i].first << '=' << times[i].second;
    out << "\nkernel_output_elements_per_second=" << elems_per_s
        << " complex_GMAC_per_second=" << gmac_per_s;
    if (o.strongest_beam) out << " peak_integrated_beam=" << *o.strongest_beam;
    out << "\nWrote " << o.bytes << " bytes of output data to " << j.output << "\n";
}

}  // namespace

int main(int argc, char** argv) {
    try {
        Job job;
        bfcli::OptionTable table = options_for(job);
        if (!table.parse(argc, argv, bfcli::Unknown::Reject, std::cout)) return 0;
        validate(job, table);
        report(job, execute(job), std::cout);
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "beamformer_amd: " << e.what() << "\n";
        return 1;
    }
}
