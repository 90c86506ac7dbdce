// bf_tracker.hpp -- tracker launch description shared by the tracker
// kernels and the C-ABI (internal).
#pragma once

#include "bf_common.hpp"

namespace bfhip {

inline constexpr int kDirTrajectory = 0;
inline constexpr int kDirTable = 1;

// Array geometry: device tables, or (when a pointer is null) the
// reference's hard-coded geometry: a regular grid of `cols` columns at
// `spacing` metres and channels f_start + k * f_step.
struct Geom {
    const float* pos;    // [A][3] or null
    const float* freqs;  // [F] or null
    int cols;
    float spacing;
    float f_start;
    float f_step;
};

struct TrackerLaunch {
    const uint8_t* x = nullptr;  // packed [T][F][A]
    float* out = nullptr;        // [T][F][B] or [W][F][B]
    size_t T = 0, F = 0, A = 0, B = 0, S = 0;
    size_t n_windows = 0;
    Geom geo{};
    int dir_source = kDirTrajectory;
    const float* dir_table = nullptr;     // [n_windows][B][3]
    const float* trajectories = nullptr;  // [B][4]
    size_t window_offset = 0;             // absolute index of window 0
    bool integrate = false;
    int kernel = BF_KERNEL_SPLIT_F16;
};

void launch_tracker(hipStream_t stream, const TrackerLaunch& launch);

}  // namespace bfhip
