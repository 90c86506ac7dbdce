// beamformer_amd -- the reference's fixed-grid GPU CLI
// (tools/beamformer_cuda.cpp) on the MI355X build: packed voltage and
// weight files in, f32 intensities or int8 codes + parameters + metadata
// out, with the same flags, output files and metrics CSV columns (backend
// "hip").  Kernels: direct and tiled (FP32, canonical or tiled weights, as
// in the reference), split-f16 and f16 (MFMA, canonical weights), auto.
#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cstdlib>
#include <exception>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <optional>
#include <stdexcept>
#include <string>
#include <system_error>
#include <vector>

#include "beamformer/amd.hpp"
#include "beamformer/amd_io.hpp"

namespace {

namespace bfa = beamformer::amd;
using Clock = std::chrono::steady_clock;

struct Options {
    std::filesystem::path input, weights, output;
    std::optional<std::filesystem::path> metrics;
    std::size_t n_time = 15360, n_ant = 64, n_beams = 5;
    std::optional<std::size_t> integration_spectra;
    bool output_int8 = false;
    std::optional<std::filesystem::path> quantization_parameters, quantization_metadata;
    std::size_t shard_id = 0;
    bfa::Kernel kernel = bfa::Kernel::Direct;
};

void print_usage(const char* program) {
    std::cout << "Usage: " << program << " --input FILE --weights FILE --output FILE [options]\n\n"
              << "  --n-time N              default: 15360\n"
              << "  --n-ant N               32, 64, 128 or 256; default: 64\n"
              << "  --n-beams N             1 to 128; default: 5\n"
              << "  --kernel NAME           direct, tiled, split-f16, f16 or auto; default: direct\n"
              << "  --integration-spectra N GPU integration\n"
              << "  --output-format NAME    float32 or int8; default: float32\n"
              << "  --quantization-params F int8 parameter output; required for int8\n"
              << "  --quantization-metadata F int8 metadata output; required for int8\n"
              << "  --shard-id N            0 or 1; default: 0\n"
              << "  --metrics FILE          append timing row to CSV\n";
}

const char* require_value(int argc, char** argv, int& i) {
    if (i + 1 >= argc) throw std::invalid_argument(std::string("missing value after ") + argv[i]);
    return argv[++i];
}

bfa::Kernel parse_kernel(const std::string& name) {
    if (name == "direct") return bfa::Kernel::Direct;
    if (name == "tiled") return bfa::Kernel::Tiled;
    if (name == "split-f16") return bfa::Kernel::SplitF16;
    if (name == "f16") return bfa::Kernel::F16;
    if (name == "auto") return bfa::Kernel::Auto;
    throw std::invalid_argument("kernel must be direct, tiled, split-f16, f16 or auto");
}

const char* kernel_name(bfa::Kernel k) {
    switch (k) {
    case bfa::Kernel::Direct: return "direct";
    case bfa::Kernel::Tiled: return "tiled";
    case bfa::Kernel::SplitF16: return "split-f16";
    case bfa::Kernel::F16: return "f16";
    case bfa::Kernel::Auto: return "auto";
    default: return "fp32";
    }
}

std::size_t parse_size(const char* value, const char* option) {
    const std::string text(value);
    std::size_t used = 0;
    const auto parsed = std::stoull(text, &used);
    if (used != text.size()) throw std::invalid_argument(std::string("invalid integer for ") + option);
    return static_cast<std::size_t>(parsed);
}

Options parse_options(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a(argv[i]);
        if (a == "--help") {
            print_usage(argv[0]);
            std::exit(0);
        } else if (a == "--input") {
            o.input = require_value(argc, argv, i);
        } else if (a == "--weights") {
            o.weights = require_value(argc, argv, i);
        } else if (a == "--output") {
            o.output = require_value(argc, argv, i);
        } else if (a == "--metrics") {
            o.metrics = require_value(argc, argv, i);
        } else if (a == "--n-time") {
            o.n_time = parse_size(require_value(argc, argv, i), "--n-time");
        } else if (a == "--n-ant") {
            o.n_ant = parse_size(require_value(argc, argv, i), "--n-ant");
        } else if (a == "--n-beams") {
            o.n_beams = parse_size(require_value(argc, argv, i), "--n-beams");
        } else if (a == "--kernel") {
            o.kernel = parse_kernel(require_value(argc, argv, i));
        } else if (a == "--integration-spectra") {
            o.integration_spectra = parse_size(require_value(argc, argv, i), "--integration-spectra");
        } else if (a == "--output-format") {
            const std::string f = require_value(argc, argv, i);
            if (f != "float32" && f != "int8")
                throw std::invalid_argument("output format must be float32 or int8");
            o.output_int8 = f == "int8";
        } else if (a == "--quantization-params") {
            o.quantization_parameters = require_value(argc, argv, i);
        } else if (a == "--quantization-metadata") {
            o.quantization_metadata = require_value(argc, argv, i);
        } else if (a == "--shard-id") {
            o.shard_id = parse_size(require_value(argc, argv, i), "--shard-id");
        } else {
            throw std::invalid_argument("unknown option: " + a);
        }
    }
    if (o.input.empty() || o.weights.empty() || o.output.empty())
        throw std::invalid_argument("--input, --weights, and --output are required");
    if (o.output_int8 &&
        (!o.integration_spectra || !o.quantization_parameters || !o.quantization_metadata))
        throw std::invalid_argument(
            "int8 output requires --integration-spectra, --quantization-params, and "
            "--quantization-metadata");
    if (o.shard_id >= 2) throw std::invalid_argument("--shard-id must be 0 or 1");
    return o;
}

double elapsed_ms(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Beam with the largest summed intensity (tools/beamformer_cuda.cpp:163-173).
std::size_t peak_beam(const std::vector<float>& intensity, std::size_t n_beams) {
    std::vector<double> sum(n_beams, 0.0);
    for (std::size_t i = 0; i < intensity.size(); ++i) sum[i % n_beams] += intensity[i];
    return (std::size_t)std::distance(sum.begin(), std::max_element(sum.begin(), sum.end()));
}

struct Timings {
    double load_ms = 0, unpack_ms = 0, setup_ms = 0, host_to_device_ms = 0, compute_ms = 0,
           device_to_host_ms = 0, write_ms = 0, total_ms = 0;
};

void append_metrics(const std::filesystem::path& path, const bfa::Dimensions& d,
                    const Timings& t, double rate, double gmac) {
    std::error_code ec;
    const bool exists = std::filesystem::exists(path, ec);
    if (ec) throw std::runtime_error("cannot inspect metrics file: " + path.string());
    const auto size = exists ? std::filesystem::file_size(path, ec) : 0;
    if (ec) throw std::runtime_error("cannot inspect metrics file: " + path.string());
    std::ofstream out(path, std::ios::app);
    if (!out) throw std::runtime_error("cannot open metrics file: " + path.string());
    if (!exists || size == 0)
        out << "backend,n_time,n_freq,n_ant,n_beams,load_ms,unpack_ms,setup_ms,"
               "host_to_device_ms,compute_ms,device_to_host_ms,write_ms,total_ms,"
               "output_elements_per_second,complex_gmac_per_second\n";
    out << std::fixed << std::setprecision(6) << "hip," << d.n_time << ',' << d.n_freq << ','
        << d.n_ant << ',' << d.n_beams << ',' << t.load_ms << ',' << t.unpack_ms << ','
        << t.setup_ms << ',' << t.host_to_device_ms << ',' << t.compute_ms << ','
        << t.device_to_host_ms << ',' << t.write_ms << ',' << t.total_ms << ',' << rate << ','
        << gmac << '\n';
    if (!out) throw std::runtime_error("failed to write metrics file: " + path.string());
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Options o = parse_options(argc, argv);
        const bfa::Dimensions dims{o.n_time, 336, o.n_ant, o.n_beams};
        bfa::validate_dimensions(dims);
        const auto t0 = Clock::now();
        const bfa::PackedVoltage packed = bfa::read_packed_voltage(o.input, dims);
        const std::vector<float> weights = o.kernel == bfa::Kernel::Tiled
                                               ? bfa::read_tiled_weights(o.weights, dims)
                                               : bfa::read_weights(o.weights, dims);
        const auto t_load = Clock::now();
        bf_timings gt{};
        bfa::Dimensions out_dims = dims;
        if (o.integration_spectra)
            out_dims.n_time = bfa::integrated_time_count(dims.n_time, *o.integration_spectra);
        std::size_t out_bytes = 0, out_elements = 0;
        std::optional<std::size_t> peak;
        Clock::time_point t_compute;
        if (o.output_int8) {
            const bfa::TemporalIntegrationConfig cfg{*o.integration_spectra};
            const auto q = bfa::beamform_packed_quantized_integrated_intensity(
                packed, weights, dims, cfg, o.kernel, &gt);
            t_compute = Clock::now();
            bfa::write_quantized_intensities(o.output, q.codes, out_dims);
            bfa::write_quantization_parameters(*o.quantization_parameters, q.parameters,
                                               out_dims);
            bfa::QuantizedOutputMetadata meta;
            meta.output_dims = out_dims;
            meta.input_n_time = dims.n_time;
            meta.temporal_integration = cfg;
            meta.shard = bfa::default_shard_descriptors()[o.shard_id];
            meta.parameters_file = o.quantization_parameters->filename().string();
            bfa::write_quantized_output_metadata(*o.quantization_metadata, meta);
            out_elements = q.codes.size();
            out_bytes = q.codes.size() + q.parameters.size() * sizeof(float);
        } else {
            const std::vector<float> intensity =
                o.integration_spectra
                    ? bfa::beamform_packed_integrated_intensity(
                          packed, weights, dims,
                          bfa::TemporalIntegrationConfig{*o.integration_spectra}, o.kernel, &gt)
                    : bfa::beamform_packed_intensity(packed, weights, dims, o.kernel, &gt);
            t_compute = Clock::now();
            bfa::write_intensities(o.output, intensity, out_dims);
            out_elements = intensity.size();
            out_bytes = intensity.size() * sizeof(float);
            peak = peak_beam(intensity, out_dims.n_beams);
        }
        const auto t_write = Clock::now();
        Timings t;
        t.load_ms = elapsed_ms(t0, t_load);
        t.setup_ms = gt.setup_ms;
        t.host_to_device_ms = gt.host_to_device_ms;
        t.compute_ms = gt.kernel_ms + gt.temporal_integration_ms + gt.quantization_ms;
        t.device_to_host_ms = gt.device_to_host_ms;
        t.write_ms = elapsed_ms(t_compute, t_write);
        t.total_ms = elapsed_ms(t0, t_write);
        const double secs = t.compute_ms / 1000.0;
        const double raw = double(dims.n_time * dims.n_freq * dims.n_beams);
        const double rate = secs > 0 ? double(out_elements) / secs : 0.0;
        const double gmac = secs > 0 ? raw * double(dims.n_ant) / secs / 1e9 : 0.0;
        if (o.metrics) append_metrics(*o.metrics, dims, t, rate, gmac);
        std::cout << std::fixed << std::setprecision(3)
                  << "HIP beamforming complete: kernel=" << kernel_name(o.kernel)
                  << " input_layout=[T=" << dims.n_time << "][F=" << dims.n_freq
                  << "][B=" << dims.n_beams << "] output_layout=[T=" << out_dims.n_time
                  << "][F=" << out_dims.n_freq << "][B=" << out_dims.n_beams << "]\n"
                  << "load_ms=" << t.load_ms << " setup_ms=" << t.setup_ms
                  << " host_to_device_ms=" << t.host_to_device_ms
                  << " kernel_ms=" << t.compute_ms << " beamformer_ms=" << gt.kernel_ms
                  << " temporal_integration_ms=" << gt.temporal_integration_ms
                  << " quantization_ms=" << gt.quantization_ms
                  << " device_to_host_ms=" << t.device_to_host_ms
                  << " write_ms=" << t.write_ms << " total_ms=" << t.total_ms << "\n"
                  << "kernel_output_elements_per_second=" << rate
                  << " complex_GMAC_per_second=" << gmac;
        if (peak) std::cout << " peak_integrated_beam=" << *peak;
        std::cout << "\nWrote " << out_bytes << " bytes of output data to " << o.output << "\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "beamformer_amd: " << e.what() << "\n";
        return 1;
    }
}
