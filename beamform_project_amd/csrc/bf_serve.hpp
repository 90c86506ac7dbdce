// bf_serve.hpp -- device-resident state of the streaming tracker server
// (internal; shared by the serving kernels and the C-ABI).
//
// The server keeps one voltage ring [R][F][A] (R a power of two, spectrum t
// at row t mod R), a request table of kServeMaxSlots slots in device memory,
// and a control block.  A step processes every window of every active slot
// whose last spectrum lies in (previous head, head]: a one-workgroup plan
// kernel reads the step's head and turns the slot table into work offsets,
// and one persistent kernel works through workgroup items (multi-tile
// windows) and then wave tasks (one-tile windows).
#pragma once

#include "bf_tracker.hpp"

namespace bfhip {

inline constexpr int kServeMaxSlots = 256;
inline constexpr int kServeMaxBeams = 64;

// One request as the kernels see it.  Windows w = 0, 1, ... cover spectra
// [t0 + w*S, min(t0 + (w+1)*S, t_end)); window w's integrated [F][B] output
// goes to out + (w mod out_windows) * F * B.  Beam b points along the
// trajectory traj[b] = (l0, m0, rate_l, rate_m) evaluated at the window's
// first spectrum, counted from t0.  w_next is the first window no step has
// issued yet (0 at admission, advanced by the plan kernel): a request whose
// start lies before the previous step's head has its early windows issued by
// the first step after its admission.
struct ServeSlot {
    int active;
    int B;
    int S;
    int wpi;  // windows per work item
    long long t0;
    long long t_end;
    long long out_windows;
    float* out;
    long long w_next;
    long long reserved;
    float traj[kServeMaxBeams][4];
};
static_assert(sizeof(ServeSlot) == 64 + 16 * kServeMaxBeams, "ServeSlot layout");

// Plan state: steps planned so far and the head of the previous step.
struct ServeCtrl {
    long long step;
    long long processed;
    long long reserved[2];
};

// Heads of the steps, written by the host into pinned host memory before it
// launches each step: step k releases spectra up to heads[k mod
// kServeHeadRing].  The host keeps fewer than kServeHeadRing / 2 steps in
// flight, so an entry is never rewritten before its step has read it.
inline constexpr int kServeHeadRing = 4096;

// Windows of one wave task of a one-tile-window slot.
inline constexpr int kServeShortRun = 16;

// Per-step work table written by the plan kernel.  Slots whose windows span
// several 16-spectrum tiles make workgroup items (run of windows, channel),
// [off_long[s], off_long[s+1]); slots whose windows fit one tile make wave
// tasks (run of up to kServeShortRun windows, 16-beam block, channel),
// [off_short[s], off_short[s+1]).  Channels are fastest in both.
struct ServeWork {
    int off_long[kServeMaxSlots + 1];
    int off_short[kServeMaxSlots + 1];
    int w_lo[kServeMaxSlots];
    int n_win[kServeMaxSlots];
};

struct ServeLaunch {
    const uint8_t* ring = nullptr;
    const uint8_t* valid = nullptr;  // [R][F] frame-validity ring, or null (all valid)
    long long ring_mask = 0;
    size_t F = 0, A = 0;
    Geom geo{};
    ServeSlot* slots = nullptr;  // the plan kernel advances each slot's window cursor
    ServeCtrl* ctrl = nullptr;
    ServeWork* work = nullptr;
    const long long* heads = nullptr;  // device view of the pinned head ring
    int grid = 0;  // persistent workgroups of the item kernel
    bool split = true;
};

// Windows per workgroup item for multi-tile windows (about 16 tiles).
inline int serve_windows_per_item(int S) {
    const int tpw = (S + 15) / 16;
    const int w = 16 / tpw;
    return w < 1 ? 1 : w;
}

// The step: plan kernel + item kernel (capturable; the head comes from the
// pinned ring, so the captured launch never changes).
void launch_serve_step(hipStream_t stream, const ServeLaunch& launch);

}  // namespace bfhip
