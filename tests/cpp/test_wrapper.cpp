// test_wrapper.cpp -- the C++ layer (include/beamformer/amd.hpp) maps the
// C-ABI's status codes onto the reference's exception types, and its file
// layer (amd_io.hpp) round-trips the reference's formats.  Runs without a
// GPU: every device call below fails validation before touching the device.
#include <cstdio>
#include <cstdlib>
#include <unistd.h>
#include <cstring>
#include <stdexcept>
#include <string>

#include "beamformer/amd.hpp"
#include "beamformer/amd_io.hpp"

namespace bfa = beamformer::amd;

static int failures = 0;

#define EXPECT(cond)                                                     \
    do {                                                                 \
        if (!(cond)) {                                                   \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                  \
        }                                                                \
    } while (0)

template <typename Exc, typename Fn>
static bool throws(Fn fn, const char* needle) {
    try {
        fn();
    } catch (const Exc& e) {
        return std::strstr(e.what(), needle) != nullptr;
    } catch (...) {
        return false;
    }
    return false;
}

int main() {
    // config.hpp:81-94 contract
    bfa::validate_dimensions({1, 336, 64, 1});
    EXPECT((throws<std::invalid_argument>([] { bfa::validate_dimensions({0, 336, 64, 1}); },
                                          "n_time")));
    EXPECT((throws<std::invalid_argument>([] { bfa::validate_dimensions({1, 672, 64, 1}); },
                                          "336")));
    EXPECT((throws<std::invalid_argument>([] { bfa::validate_dimensions({1, 336, 48, 1}); },
                                          "n_ant")));
    EXPECT((throws<std::invalid_argument>([] { bfa::validate_dimensions({1, 336, 64, 129}); },
                                          "n_beams")));
    EXPECT(bfa::integrated_time_count(32768, 10) == 3277);
    EXPECT(bfa::integrated_time_count(65536, 320) == 205);
    EXPECT(bfa::quantization_parameter_count({2, 336, 32, 17}) == 84);
    EXPECT((throws<std::invalid_argument>([] { bfa::integrated_time_count(10, 0); },
                                          "integration")));
    // stage launchers validate before launching (cuda_stage_api.hpp:19-20)
    EXPECT((throws<std::invalid_argument>(
        [] {
            bfa::launch_packed_beamformer(bfa::Kernel::Fp32, nullptr, nullptr, nullptr, nullptr,
                                          {256, 336, 64, 4});
        },
        "null")));
    EXPECT((throws<std::invalid_argument>(
        [] {
            bfa::launch_packed_integrated_beamformer(
                bfa::Kernel::Direct, nullptr, reinterpret_cast<const std::uint8_t*>(256),
                reinterpret_cast<const float*>(256), reinterpret_cast<float*>(256),
                {640, 336, 64, 4}, {10});
        },
        "Direct")));
    // one-shot API checks buffer sizes before any device work
    EXPECT((throws<std::invalid_argument>(
        [] {
            std::vector<std::uint8_t> x(10);
            std::vector<float> w(2 * 4 * 336 * 64);
            bfa::beamform_packed_intensity(x, w, {4, 336, 64, 4});
        },
        "")));
    // tracker contract: one beam (src/beam_tracker.cpp:18-21)
    EXPECT((throws<std::invalid_argument>(
        [] {
            std::vector<std::uint8_t> x(10 * 336 * 64);
            bfa::beam_tracker_into(x, {10, 336, 64, 2}, bfa::TrackerConfig{});
        },
        "n_beams")));
    // on-disk formats (io.hpp): round trips and the reference's error types
    {
        const char* tmp = std::getenv("TMPDIR");
        const std::string base = std::string(tmp ? tmp : "/tmp") + "/bf_wrapper_io_" +
                                 std::to_string(::getpid());
        const bfa::Dimensions d{3, 336, 32, 2};
        bfa::PackedVoltage v(3 * 336 * 32);
        for (std::size_t i = 0; i < v.size(); ++i) v[i] = static_cast<std::uint8_t>(i * 7);
        bfa::write_packed_voltage(base + "_v.bin", v, d);
        EXPECT(bfa::read_packed_voltage(base + "_v.bin", d) == v);
        EXPECT((throws<std::runtime_error>(
            [&] { bfa::read_packed_voltage(base + "_v.bin", {4, 336, 32, 2}); }, "expected")));
        EXPECT((throws<std::invalid_argument>(
            [&] { bfa::write_packed_voltage(base + "_w.bin", bfa::PackedVoltage(5), d); },
            "does not match")));
        bfa::PackedShard sh;
        sh.descriptor = bfa::default_shard_descriptors()[1];
        EXPECT(sh.descriptor.absolute_frequency_start == 336 && sh.descriptor.loss_mask_id == 1);
        sh.payload = v;
        sh.loss_mask.assign(3 * 336, 1);
        sh.loss_mask[5] = 0;
        bfa::write_packed_shard(base + "_p.bin", base + "_m.txt", base + "_k.bin", sh, d);
        const bfa::PackedShard back =
            bfa::read_packed_shard(base + "_p.bin", base + "_m.txt", base + "_k.bin", d);
        EXPECT(back.payload == sh.payload && back.loss_mask == sh.loss_mask);
        EXPECT(back.descriptor.shard_id == 1 && back.descriptor.loss_mask_independent);
        bfa::QuantizedOutputMetadata m;
        m.output_dims = {32, 336, 64, 16};
        m.input_n_time = 320;
        m.temporal_integration.integration_spectra = 10;
        m.shard = sh.descriptor;
        m.parameters_file = "params.bin";
        bfa::write_quantized_output_metadata(base + "_q.txt", m);
        const bfa::QuantizedOutputMetadata mq =
            bfa::read_quantized_output_metadata(base + "_q.txt");
        EXPECT(mq.output_dims.n_time == 32 && mq.parameters_file == "params.bin" &&
               mq.shard.absolute_frequency_start == 336);
        m.output_dims.n_time = 33;
        EXPECT((throws<std::invalid_argument>(
            [&] { bfa::write_quantized_output_metadata(base + "_q2.txt", m); }, "inconsistent")));
        for (const char* suf : {"_v.bin", "_p.bin", "_m.txt", "_k.bin", "_q.txt"})
            std::remove((base + suf).c_str());
    }
    std::printf("%s (%d failures)\n", failures ? "FAIL" : "OK", failures);
    return failures ? 1 : 0;
}
