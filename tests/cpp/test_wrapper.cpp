// test_wrapper.cpp -- the C++ layer (include/beamformer/amd.hpp) maps the
// C-ABI's status codes onto the reference's exception types, and its file
// layer (amd_io.hpp) round-trips the reference's formats.  Runs without a
// GPU: every device call below fails validation before touching the device.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <cstdlib>
#include <unistd.h>
#include <cstring>
#include <stdexcept>
#include <string>

#include "beamformer/amd.hpp"
#include "beamformer/amd_io.hpp"
#include "beamformer/amd_opt_tracker.hpp"

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
    {
        // tests/cpu/test_geometry.cpp, through the C++ layer
        auto close = [](float a, float b) { return std::fabs(a - b) < 1.0e-6F; };
        const auto p32 = bfa::default_positions(32);
        EXPECT(p32.size() == 32 && (p32[0] == bfa::Vec3{0.0F, 0.0F, 0.0F}));
        EXPECT(close(p32[7][0], 4.2F) && close(p32[8][1], 0.6F) && close(p32[31][1], 1.8F));
        EXPECT((bfa::default_positions(64, 0.5F)[63] == bfa::Vec3{3.5F, 3.5F, 0.0F}));
        const auto fr = bfa::channelized_frequencies();
        EXPECT(fr.size() == 336 && fr.front() == 300'000'000.0F && fr.back() == 400'500'000.0F);
        const auto b5 = bfa::default_beam_grid(5);
        EXPECT(b5[0][0] == -0.04F && b5[2][0] == 0.0F && b5[4][0] == 0.04F);
        EXPECT(bfa::default_beam_grid(10)[7][0] == 0.04F);
        const float lambda = static_cast<float>(299792458.0 / 400'000'000.0);
        const auto r32 = bfa::rectangular_beam_grid(32);
        EXPECT(r32.size() == 32 && close(r32[1][0] - r32[0][0], lambda / (8.0F * 0.6F)));
        EXPECT((bfa::centered_integer_range(4) == std::vector<long long>{-2, -1, 0, 1}));
        const float du = lambda / (2.0F * 8.0F * 0.6F);
        const auto s32 = bfa::select_fft_beam_centers_rectangular(du, 2.0F * du, 128, 8, 4);
        EXPECT(s32.centers.size() == 128 && s32.n_u == 16 && s32.n_v == 8);
        const auto s64 = bfa::select_fft_beam_centers_rectangular(du, du, 128, 8, 8);
        EXPECT(s64.n_u == 12 && s64.n_v == 11 && s64.n_bank_u == 16 && s64.n_bank_v == 16);
        EXPECT(bfa::fft_beam_grid(32, 128).size() == 128);
        EXPECT(bfa::estimate_nbeams_hex_formula(0.6F, 8, 8, 108.0F, 74.0F, lambda) > 0);
        const auto hx = bfa::generate_hex_targets_cropped_fov(128, 0.8F, 0.6F);
        bool inside = hx.size() == 128;
        for (const auto& t : hx)
            inside = inside && t[0] * t[0] / 0.64F + t[1] * t[1] / 0.36F <= 1.0F + 1.0e-6F;
        EXPECT(inside);
        const std::string base = std::string("/tmp/bf_geom_") + std::to_string(getpid());
        {
            std::ofstream o(base + "_p.txt");
            o << "# x, y, z in metres\n1.0, 2.0, 3.0\n4.0 5.0 6.0\n";
        }
        const auto lp = bfa::load_positions(base + "_p.txt", 2);
        EXPECT((lp[1] == bfa::Vec3{4.0F, 5.0F, 6.0F}));
        EXPECT((throws<std::runtime_error>([&] { bfa::load_positions(base + "_p.txt", 3); },
                                           "expected 3")));
        std::remove((base + "_p.txt").c_str());
        EXPECT((throws<std::invalid_argument>([] { bfa::default_positions(48); }, "256")));
        const auto w = bfa::generate_weights({1, 336, 32, 2}, p32, fr, bfa::default_beam_grid(2));
        EXPECT(w.size() == 2 * 2 * 336 * 32 && w[0] == 1.0F && w[1] == 0.0F);
        const auto t = bfa::generate_tiled_weights({1, 336, 32, 2}, p32, fr,
                                                   bfa::default_beam_grid(2));
        EXPECT(t.size() == 2 * 336 * 1 * 32 * 32);
        // adaptive tracker: constructor checks of cpu_opt_beam_tracker.cpp:1084-1113
        bfa::OptTrackerConfig bad;
        bad.forgetting_factor = 1.5F;
        EXPECT((throws<std::invalid_argument>(
            [&] { bfa::OptBeamTracker tr({}, {}, {4, 336, 32, 1}, bad); }, "forgetting")));
        bad = {};
        bad.spatial_smoothing_subarray_size = 100;
        EXPECT((throws<std::invalid_argument>(
            [&] { bfa::OptBeamTracker tr({}, {}, {4, 336, 32, 1}, bad); }, "subarray")));
        EXPECT((throws<std::invalid_argument>(
            [] { bfa::OptBeamTracker tr({}, {}, {4, 336, 32, 2}, {}); }, "n_beams == 1")));
    }
    std::printf("%s (%d failures)\n", failures ? "FAIL" : "OK", failures);
    return failures ? 1 : 0;
}
