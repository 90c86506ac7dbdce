// test_wrapper.cpp -- the C++ layer (include/beamformer/amd.hpp) maps the
// C-ABI's status codes onto the reference's exception types.  Runs without
// a GPU: every call below fails validation before touching the device.
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "beamformer/amd.hpp"

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
    std::printf("%s (%d failures)\n", failures ? "FAIL" : "OK", failures);
    return failures ? 1 : 0;
}
