// bf_serve.hpp -- device-resident state of the streaming tracker server
// (internal; shared by the serving kernels and the C-ABI).
//
// The server keeps one voltage ring [R][F][A] (R a power of two, spectrum t
// at row t mod R), a request table of kServeMaxSlots slots in device memory,
// and a control block.  A step processes every window of every active slot
// whose last spectrum lies in (processed, head]: a one-workgroup plan kernel
// turns the slot table into per-slot item ranges, and one persistent kernel
// works through the items (slot, channel, run of windows).
#pragma once

#include "bf_tracker.hpp"

namespace bfhip {

inline constexpr int kServeMaxSlots = 256;
inline constexpr int kServeMaxBeams = 64;

// One request as the kernels see it.  Windows w = 0, 1, ... cover spectra
// [t0 + w*S, min(t0 + (w+1)*S, t_end)); window w's integrated [F][B] output
// goes to out + (w mod out_windows) * F * B.  Beam b points along the
// trajectory traj[b] = (l0, m0, rate_l, rate_m) evaluated at the window's
// first spectrum, counted from t0.
struct ServeSlot {
    int active;
    int B;
    int S;
    int wpi;  // windows per work item
    long long t0;
    long long t_end;
    long long out_windows;
    float* out;
    long long reserved[2];
    float traj[kServeMaxBeams][4];
};
static_assert(sizeof(ServeSlot) == 64 + 16 * kServeMaxBeams, "ServeSlot layout");

struct ServeCtrl {
    long long head;       // spectra landed and released to this step
    long long processed;  // head of the previous step
    long long n_items;    // items of the current step
    long long reserved;
};

// Per-step work table written by the plan kernel: slot s owns items
// [off[s], off[s+1]), ordered (run of windows, channel), and its windows
// [w_lo[s], w_lo[s] + n_win[s]).
struct ServeWork {
    int off[kServeMaxSlots + 1];
    int w_lo[kServeMaxSlots];
    int n_win[kServeMaxSlots];
};

struct ServeLaunch {
    const uint8_t* ring = nullptr;
    long long ring_mask = 0;
    size_t F = 0, A = 0;
    Geom geo{};
    const ServeSlot* slots = nullptr;
    ServeCtrl* ctrl = nullptr;
    ServeWork* work = nullptr;
    int grid = 0;  // persistent workgroups of the item kernel
    bool split = true;
};

// Windows per work item for an S-spectrum window (about 16 tiles per item).
inline int serve_windows_per_item(int S) {
    const int tpw = (S + 15) / 16;
    const int w = 16 / tpw;
    return w < 1 ? 1 : w;
}

// ctrl->head = head (one tiny kernel; keeps the captured step invariant).
void launch_serve_set_head(hipStream_t stream, ServeCtrl* ctrl, long long head);
// The step itself: plan kernel + item kernel (capturable).
void launch_serve_step(hipStream_t stream, const ServeLaunch& launch);

}  // namespace bfhip
