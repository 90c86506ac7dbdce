// amd_opt_tracker.hpp -- header-only C++ layer over bf_opt_tracker_* with
// the reference's CpuOptBeamTracker surface
// (include/beamformer/cpu_opt_beam_tracker.hpp:39-196): TrackerEstimator,
// the configuration, the stateful tracker (seed_trajectory, run_into,
// window_direction) and the stateless free function.  The search runs on
// the GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "amd_geometry.hpp"

namespace beamformer {
namespace amd {

enum class TrackerEstimator : int {
    Bartlett = BF_ESTIMATOR_BARTLETT,
    Capon = BF_ESTIMATOR_CAPON,
};

// CpuOptTrackerConfig; `kernel` computes the emitted intensities.
struct OptTrackerConfig {
    TrackerEstimator estimator = TrackerEstimator::Bartlett;
    std::size_t coarse_grid_resolution = 1;
    std::size_t refinement_levels = 0;
    float search_fov_l = 0.2F;
    float search_fov_m = 0.2F;
    std::size_t spatial_smoothing_subarray_size = 0;
    float forgetting_factor = 1.0F;
    float diagonal_load = 1.0e-3F;
    bool enable_quadratic_peak_interp = true;
    std::size_t integration_spectra = 320;
    Kernel kernel = Kernel::Auto;
    bf_opt_tracker_config c() const {
        bf_opt_tracker_config o{};
        o.estimator = static_cast<int>(estimator);
        o.coarse_grid_resolution = coarse_grid_resolution;
        o.refinement_levels = refinement_levels;
        o.search_fov_l = search_fov_l;
        o.search_fov_m = search_fov_m;
        o.spatial_smoothing_subarray_size = spatial_smoothing_subarray_size;
        o.forgetting_factor = forgetting_factor;
        o.diagonal_load = diagonal_load;
        o.enable_quadratic_peak_interp = enable_quadratic_peak_interp ? 1 : 0;
        o.integration_spectra = integration_spectra;
        o.kernel = static_cast<int>(kernel);
        return o;
    }
};

class OptBeamTracker {
public:
    // Empty positions / frequencies select the default geometry.
    OptBeamTracker(const std::vector<Vec3>& positions_m, const std::vector<float>& frequencies_hz,
                   const Dimensions& dims, const OptTrackerConfig& config)
        : dims_(dims), config_(config) {
        if (!positions_m.empty() && positions_m.size() != dims.n_ant)
            throw std::invalid_argument("position count must match n_ant");
        if (!frequencies_hz.empty() && frequencies_hz.size() != dims.n_freq)
            throw std::invalid_argument("frequency count must match n_freq");
        const bf_dims d = dims.c();
        const bf_opt_tracker_config c = config.c();
        check(bf_opt_tracker_create(positions_m.empty() ? nullptr : positions_m.front().data(),
                                    frequencies_hz.empty() ? nullptr : frequencies_hz.data(), &d,
                                    &c, &t_));
    }
    ~OptBeamTracker() {
        if (t_) (void)bf_opt_tracker_destroy(t_);
    }
    OptBeamTracker(const OptBeamTracker&) = delete;
    OptBeamTracker& operator=(const OptBeamTracker&) = delete;
    OptBeamTracker(OptBeamTracker&& o) noexcept
        : dims_(o.dims_), config_(o.config_), t_(std::exchange(o.t_, nullptr)) {}

    void seed_trajectory(const TrackerConfig& trajectory) {
        check(bf_opt_tracker_seed_trajectory(t_, trajectory.direction_start,
                                             trajectory.direction_rate_per_sample));
    }
    void reset() { check(bf_opt_tracker_reset(t_)); }

    // Host buffers: packed [T][F][A] -> intensity [T][F][1].
    void run_into(const std::vector<std::uint8_t>& packed, std::vector<float>& intensity) {
        check(bf_opt_tracker_run(t_, packed.data(), packed.size(), intensity.data(),
                                 intensity.size()));
    }
    // Device buffers, asynchronous on `stream`.
    void run_device(const std::uint8_t* d_packed, float* d_intensity, void* stream = nullptr) {
        check(bf_opt_tracker_run_device(t_, stream, d_packed, d_intensity));
    }
    std::vector<Vec3> window_directions() const {
        std::size_t n = 0;
        check(bf_opt_tracker_window_directions(t_, nullptr, 0, &n));
        std::vector<Vec3> out(n);
        if (n) check(bf_opt_tracker_window_directions(t_, out.front().data(), n, &n));
        return out;
    }
    Vec3 window_direction(std::size_t window) const {
        const auto d = window_directions();
        return window < d.size() ? d[window] : Vec3{0.0F, 0.0F, 1.0F};
    }
    const Dimensions& dimensions() const noexcept { return dims_; }
    const OptTrackerConfig& config() const noexcept { return config_; }

private:
    Dimensions dims_;
    OptTrackerConfig config_;
    bf_opt_tracker* t_ = nullptr;
};

// cpu_opt_beam_tracker_packed_intensity (stateless: default geometry, a
// fresh covariance recursion, the trajectory's integration_spectra).
inline std::vector<float> opt_beam_tracker_packed_intensity(
    const std::vector<std::uint8_t>& packed, const Dimensions& dims,
    const TrackerConfig& trajectory, const OptTrackerConfig& opt = {}) {
    const bf_dims d = dims.c();
    const bf_tracker_config t = trajectory.c();
    const bf_opt_tracker_config c = opt.c();
    std::vector<float> out(dims.n_time * dims.n_freq * dims.n_beams);
    check(bf_opt_tracker_packed_intensity(packed.data(), packed.size(), &d, &t, &c, out.data(),
                                          out.size()));
    return out;
}

}  // namespace amd
}  // namespace beamformer
