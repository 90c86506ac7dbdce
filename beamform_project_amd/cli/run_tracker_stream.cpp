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
#include "cli_common.hpp"

namespace {

namespace bfa = beamformer::amd;

struct Run {
    std::string engine = "amd_split";
    std::size_t n_time = 15360, n_freq = 336, n_ant = 64, spectra = 320;
    float l0 = 0.0F, m0 = 0.0F, dl = 0.0F, dm = 0.0F;
    std::string input, output;
};

bfcli::OptionTable options_for(Run& r) {
    using namespace bfcli;
    OptionTable t;
    // lenient numbers (a leading number is read, as std::stoull / std::stof do)
    t.add("--engine", "ENGINE", "amd_split (default), amd_f16, amd_fp32; the reference's cuda_* "
                                "names run amd_split", text_into(r.engine))
        .add("--n-time", "N", "number of time samples", size_into(r.n_time, false))
        .add("--n-freq", "N", "number of frequency channels (default 336)", size_into(r.n_freq, false))
        .add("--n-ant", "N", "number of antenna elements (32, 64, 128, 256)", size_into(r.n_ant, false))
        .add("--spectra", "N", "integration spectra per window (default 320)",
             size_into(r.spectra, false))
        .add("--source-l0", "L0", "steering start l0", float_into(r.l0))
        .add("--source-m0", "M0", "steering start m0", float_into(r.m0))
        .add("--source-dl", "DL", "steering rate dl/sample", float_into(r.dl))
        .add("--source-dm", "DM", "steering rate dm/sample", float_into(r.dm))
        .add("--input", "FILE", "input packed voltage binary file", text_into(r.input))
        .add("--output", "FILE", "output float32 intensity binary file", text_into(r.output))
        .help_flags({"--help", "-h"});
    return t;
}

// Engine name -> kernel; throws for names this build does not provide.
bfa::Kernel engine_kernel(const std::string& engine) {
    auto prefixed = [&](const char* p) { return engine.rfind(p, 0) == 0; };
    if (engine == "amd_split" || engine == "auto" || prefixed("cuda_")) return bfa::Kernel::SplitF16;
    if (engine == "amd_f16") return bfa::Kernel::F16;
    if (engine == "amd_fp32") return bfa::Kernel::Fp32;
    if (prefixed("cpu_"))
        throw std::invalid_argument("engine " + engine +
                                    " is the reference's CPU tracker, not part of this build");
    throw std::invalid_argument("Unknown engine: " + engine);
}

// Unit vector of direction cosines (l, m), as src/geometry.cpp:97-106
// defines it, in float arithmetic.
void unit_direction(float l, float m, float out[3]) {
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
    Run run;
    bfcli::OptionTable table = options_for(run);
    try {
        if (!table.parse(argc, argv, bfcli::Unknown::Skip, std::cout)) return 0;
        if (run.input.empty() || run.output.empty()) {
            std::cerr << "Error: --input and --output are required.\n" << table.usage(argv[0]);
            return 1;
        }
        const bfa::Kernel kernel = engine_kernel(run.engine);
        const bfa::Dimensions dims{run.n_time, run.n_freq, run.n_ant, 1};
        bfa::validate_dimensions(dims);
        bfa::TrackerConfig cfg;
        unit_direction(run.l0, run.m0, cfg.direction_start);
        cfg.direction_rate_per_sample[0] = run.dl;
        cfg.direction_rate_per_sample[1] = run.dm;
        cfg.integration_spectra = run.spectra;
        // Inputs and arguments are validated before the GPU is touched.
        const bfa::PackedVoltage packed = bfa::read_packed_voltage(run.input, dims);
        bfa::TrackerExecution exec;
        exec.kernel = kernel;
        const std::vector<float> intensity = bfa::beam_tracker_into(packed, dims, cfg, exec);
        bfa::write_intensities(run.output, intensity, dims);
        std::cout << "Successfully ran " << run.engine << " on " << run.input << " -> wrote "
                  << run.output << " (" << intensity.size() * sizeof(float) << " bytes)\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << "\n";
        return 1;
    }
}
