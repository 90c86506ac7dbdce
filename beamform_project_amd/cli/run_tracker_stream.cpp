// run_tracker_stream -- the reference's tracker CLI bridge
// (tools/run_tracker_stream.cpp:56-186) on the MI355X build: reads a packed
// voltage file, runs the single-beam dynamic tracker and writes the
// per-spectrum f32 intensities [T][F][1].  Same flags and engine names, so
// the reference's astronomical validation runner
// (tools/astronomical_validation/runner.py:43-90) can call it unchanged.
//
// Engines: amd_split (split-FP16 MFMA, the default), amd_f16 (single FP16
// MFMA), amd_fp32 (FP32 VALU).  The reference's GPU engine names
// (cuda_twopass, cuda_fused, cuda_warp, cuda_fws, cuda_v3*, cuda_v4*,
// cuda_v5*) select amd_split: they compute the same tracker.  Its CPU
// engines (cpu_naive, cpu_v1, cpu_v2) are not part of this build.
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <filesystem>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "beamformer/amd.hpp"
#include "beamformer/amd_io.hpp"

namespace {

namespace bfa = beamformer::amd;

void print_usage(const char* prog) {
    std::cout << "Usage: " << prog << " [options]\n"
              << "  --engine ENGINE       amd_split (default), amd_f16, amd_fp32; the\n"
              << "                        reference's cuda_* names run amd_split\n"
              << "  --n-time N            number of time samples\n"
              << "  --n-freq N            number of frequency channels (default 336)\n"
              << "  --n-ant N             number of antenna elements (32, 64, 128, 256)\n"
              << "  --spectra N           integration spectra per window (default 320)\n"
              << "  --source-l0 L0        steering start l0\n"
              << "  --source-m0 M0        steering start m0\n"
              << "  --source-dl DL        steering rate dl/sample\n"
              << "  --source-dm DM        steering rate dm/sample\n"
              << "  --input FILE          input packed voltage binary file\n"
              << "  --output FILE         output float32 intensity binary file\n";
}

bool starts_with(const std::string& s, const char* p) { return s.rfind(p, 0) == 0; }

// Engine name -> kernel; throws for names this build does not provide.
bfa::Kernel engine_kernel(const std::string& engine) {
    if (engine == "amd_split" || engine == "auto") return bfa::Kernel::SplitF16;
    if (engine == "amd_f16") return bfa::Kernel::F16;
    if (engine == "amd_fp32") return bfa::Kernel::Fp32;
    if (starts_with(engine, "cuda_")) return bfa::Kernel::SplitF16;
    if (starts_with(engine, "cpu_"))
        throw std::invalid_argument("engine " + engine +
                                    " is the reference's CPU tracker, not part of this build");
    throw std::invalid_argument("Unknown engine: " + engine);
}

// direction_from_lm (src/geometry.cpp:97-106), float arithmetic.
void direction_from_lm(float l, float m, float out[3]) {
    if (!std::isfinite(l) || !std::isfinite(m))
        throw std::invalid_argument("direction cosines l and m must be finite");
    const float t2 = l * l + m * m;
    if (t2 > 1.0F) throw std::invalid_argument("direction cosines must satisfy l*l + m*m <= 1");
    out[0] = l;
    out[1] = m;
    out[2] = std::sqrt(1.0F - t2);
}

}  // namespace

int main(int argc, char** argv) {
    std::string engine = "amd_split";
    std::size_t n_time = 15360, n_freq = 336, n_ant = 64, spectra = 320;
    float l0 = 0.0F, m0 = 0.0F, dl = 0.0F, dm = 0.0F;
    std::string input_file, output_file;
    try {
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            const bool more = i + 1 < argc;
            if (a == "--engine" && more) engine = argv[++i];
            else if (a == "--n-time" && more) n_time = std::stoull(argv[++i]);
            else if (a == "--n-freq" && more) n_freq = std::stoull(argv[++i]);
            else if (a == "--n-ant" && more) n_ant = std::stoull(argv[++i]);
            else if (a == "--spectra" && more) spectra = std::stoull(argv[++i]);
            else if (a == "--source-l0" && more) l0 = std::stof(argv[++i]);
            else if (a == "--source-m0" && more) m0 = std::stof(argv[++i]);
            else if (a == "--source-dl" && more) dl = std::stof(argv[++i]);
            else if (a == "--source-dm" && more) dm = std::stof(argv[++i]);
            else if (a == "--input" && more) input_file = argv[++i];
            else if (a == "--output" && more) output_file = argv[++i];
            else if (a == "--help" || a == "-h") {
                print_usage(argv[0]);
                return 0;
            }
        }
        if (input_file.empty() || output_file.empty()) {
            std::cerr << "Error: --input and --output are required.\n";
            print_usage(argv[0]);
            return 1;
        }
        const bfa::Kernel kernel = engine_kernel(engine);
        const bfa::Dimensions dims{n_time, n_freq, n_ant, 1};
        bfa::validate_dimensions(dims);
        bfa::TrackerConfig cfg;
        direction_from_lm(l0, m0, cfg.direction_start);
        cfg.direction_rate_per_sample[0] = dl;
        cfg.direction_rate_per_sample[1] = dm;
        cfg.integration_spectra = spectra;
        // Inputs and arguments are validated before the GPU is touched.
        const bfa::PackedVoltage packed = bfa::read_packed_voltage(input_file, dims);
        bfa::TrackerExecution exec;
        exec.kernel = kernel;
        const std::vector<float> intensity = bfa::beam_tracker_into(packed, dims, cfg, exec);
        bfa::write_intensities(output_file, intensity, dims);
        std::cout << "Successfully ran " << engine << " on " << input_file << " -> wrote "
                  << output_file << " (" << intensity.size() * sizeof(float) << " bytes)\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << "\n";
        return 1;
    }
}
