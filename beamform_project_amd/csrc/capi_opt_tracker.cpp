// capi_opt_tracker.cpp -- extern "C" adaptive planar DOA tracker: the GPU
// counterpart of the reference's CpuOptBeamTracker
// (include/beamformer/cpu_opt_beam_tracker.hpp:94-196,
// src/cpu_opt_beam_tracker.cpp:1084-1360, 1619-2494).
//
// run_into semantics followed:
//   - scan disabled (coarse_grid_resolution <= 1): the open-loop trajectory
//     tracker, one direction per window from the seeded prior (:1663-1688);
//   - scan enabled: per window the covariance recursion, the coarse grid
//     around the previous estimate (the prior's window-0 direction for the
//     first window of every call, :1693), the refinement levels, and then
//     the window's emission along the estimate (:2405-2432).  The covariance
//     state persists across calls; the estimate does not.
// The per-window search runs as a chain of small launches on the caller's
// stream (k_opt_tracker.hip); the emission of every window is one tracker
// launch over the whole shard with the estimated directions as its table.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "bf_doa.hpp"
#include "bf_tracker.hpp"

namespace bfhip {
namespace {

void require_ptr(const void* p, const char* what) {
    if (p == nullptr) invalid(std::string(what) + " must not be null");
}

template <typename T>
T* dmalloc(size_t n) {
    T* p = nullptr;
    hip_check(hipMalloc(reinterpret_cast<void**>(&p), (n ? n : 1) * sizeof(T)), "hipMalloc");
    return p;
}

constexpr size_t kMaxGrid = 64;    // coarse cells per axis
constexpr size_t kMaxLevels = 24;  // refinement halvings

// direction_from_lm (src/geometry.cpp:97-106), float arithmetic.
void direction_from_lm(float l, float m, float out[3]) {
    if (!std::isfinite(l) || !std::isfinite(m))
        invalid("direction cosines l and m must be finite");
    const float t2 = l * l + m * m;
    if (t2 > 1.0f) invalid("direction cosines must satisfy l*l + m*m <= 1");
    out[0] = l;
    out[1] = m;
    out[2] = std::sqrt(1.0f - t2);
}

// tracker_window_direction (src/beam_tracker.cpp:52-76).
void window_direction(const float start[3], const float rate[2], size_t w, size_t S,
                      float out[3]) {
    const float tf = (float)(w * S);
    direction_from_lm(start[0] + tf * rate[0], start[1] + tf * rate[1], out);
}

}  // namespace
}  // namespace bfhip

using namespace bfhip;

struct bf_opt_tracker {
    bf_dims dims{};
    bf_opt_tracker_config cfg{};
    int kernel = BF_KERNEL_SPLIT_F16;
    size_t M = 0, L = 0, G = 0;
    bool smoothing = false, coplanar = true, capon = false;
    bool initialised = false;
    float start[3] = {0.0f, 0.0f, 1.0f};
    float rate[2] = {0.0f, 0.0f};
    float h_traj[4] = {};     // staging of the trajectory row
    OptSearch h_search{};     // staging of the search state
    std::vector<float> h_dirs;  // [W][3] of the last run
    bool dirs_pending = false;  // h_dirs still on the device
    size_t n_windows = 0;
    hipStream_t last_stream = nullptr;
    float* d_pos = nullptr;
    float* d_freqs = nullptr;
    double* d_state = nullptr;
    // Per-window operands, two slots: the covariance, fold and factor of
    // window w + 1 run on the prep stream while window w is searched.
    struct Slot {
        int* rs = nullptr;
        float2* scan = nullptr;
        float* trace = nullptr;
        double2* scan64 = nullptr;  // Capon: R^-1 per channel
        double* trace64 = nullptr;
        int* mode = nullptr;
    } slot[2];
    hipStream_t prep = nullptr;
    // ev_done: the end of the last run on its stream; the next run's stream
    // (which may be another one) waits on it before it writes any state.
    hipEvent_t ev_done = nullptr;
    bool ran = false;
    hipEvent_t ev_start = nullptr, ev_ready[2] = {nullptr, nullptr},
               ev_used[2] = {nullptr, nullptr};
    float* d_part = nullptr;
    OptSearch* d_search = nullptr;
    float* d_dirs = nullptr;
    size_t dirs_cap = 0;
    float* d_traj = nullptr;

    ~bf_opt_tracker() {
        if (last_stream) (void)hipStreamSynchronize(last_stream);
        if (prep) (void)hipStreamSynchronize(prep);
        for (const Slot& q : slot)
            for (void* p : {(void*)q.rs, (void*)q.scan, (void*)q.trace, (void*)q.scan64,
                            (void*)q.trace64, (void*)q.mode})
                if (p) (void)hipFree(p);
        for (void* p : {(void*)d_pos, (void*)d_freqs, (void*)d_state, (void*)d_part,
                        (void*)d_search, (void*)d_dirs, (void*)d_traj})
            if (p) (void)hipFree(p);
        for (hipEvent_t e : {ev_start, ev_ready[0], ev_ready[1], ev_used[0], ev_used[1], ev_done})
            if (e) (void)hipEventDestroy(e);
        if (prep) (void)hipStreamDestroy(prep);
    }

    OptGridArgs grid_args() const {
        OptGridArgs g{};
        g.G = (int)G;
        g.M = (int)M;
        g.fov_l = cfg.search_fov_l;
        g.fov_m = cfg.search_fov_m;
        g.coplanar = coplanar ? 1 : 0;
        g.interp = cfg.enable_quadratic_peak_interp ? 1 : 0;
        return g;
    }

    void ensure_dirs(size_t W) {
        if (W <= dirs_cap) return;
        // a previous run may still write the old table
        if (last_stream != nullptr || dirs_cap != 0)
            hip_check(hipStreamSynchronize(last_stream), "hipStreamSynchronize");
        if (d_dirs) hip_check(hipFree(d_dirs), "hipFree");
        d_dirs = nullptr;
        d_dirs = dmalloc<float>(3 * W);
        dirs_cap = W;
    }

    void run(hipStream_t stream, const uint8_t* x, float* out) {
        const size_t T = dims.n_time, F = dims.n_freq, A = dims.n_ant;
        const size_t S = cfg.integration_spectra;
        const size_t W = ceil_div(T, S);
        // Every window's prior direction must lie on the unit disk, as the
        // reference's direction_from_lm requires; checked before any launch.
        std::vector<float> prior(3 * W);
        for (size_t w = 0; w < (G > 1 ? 1 : W); ++w) window_direction(start, rate, w, S, &prior[3 * w]);
        ensure_dirs(W);
        // Everything below (search state, trajectory, slot operands, the
        // direction table, the prep stream's first windows) is ordered after
        // the previous run, whichever stream it ran on: prep starts from
        // ev_start, recorded on `stream` after this wait.
        if (ran) hip_check(hipStreamWaitEvent(stream, ev_done, 0), "hipStreamWaitEvent");
        TrackerLaunch L_emit;
        L_emit.x = x;
        L_emit.out = out;
        L_emit.T = T;
        L_emit.F = F;
        L_emit.A = A;
        L_emit.B = 1;
        L_emit.S = S;
        L_emit.n_windows = W;
        L_emit.geo = Geom{d_pos, d_freqs, 8, 0.6f, 0.0f, 0.0f};
        L_emit.integrate = false;
        L_emit.kernel = kernel;
        last_stream = stream;
        if (G <= 1) {
            const float traj[4] = {start[0], start[1], rate[0], rate[1]};
            std::memcpy(h_traj, traj, sizeof(traj));
            hip_check(hipMemcpyAsync(d_traj, h_traj, sizeof(h_traj), hipMemcpyHostToDevice,
                                     stream),
                      "hipMemcpyAsync trajectory");
            L_emit.dir_source = kDirTrajectory;
            L_emit.trajectories = d_traj;
            launch_tracker(stream, L_emit);
            hip_check(hipEventRecord(ev_done, stream), "hipEventRecord");
            ran = true;
            h_dirs = prior;
            dirs_pending = false;
            n_windows = W;
            return;
        }
        std::memcpy(h_search.prev, &prior[0], sizeof(h_search.prev));
        std::memcpy(h_search.centre, &prior[0], sizeof(h_search.centre));
        hip_check(hipMemcpyAsync(d_search, &h_search, sizeof(h_search), hipMemcpyHostToDevice,
                                 stream),
                  "hipMemcpyAsync search");
        const OptGridArgs g = grid_args();
        const int last = (int)cfg.refinement_levels;
        // The prep stream starts after the caller's prior work (the input).
        hip_check(hipEventRecord(ev_start, stream), "hipEventRecord");
        hip_check(hipStreamWaitEvent(prep, ev_start, 0), "hipStreamWaitEvent");
        for (size_t w = 0; w < W; ++w) {
            const int k = (int)(w & 1);
            Slot& q = slot[k];
            // prep: window w's operands into slot k, once window w - 2's
            // search has released it
            if (w >= 2) hip_check(hipStreamWaitEvent(prep, ev_used[k], 0), "hipStreamWaitEvent");
            const size_t t0 = w * S;
            const size_t K = std::min(S, T - t0);
            launch_doa_covariance(prep, x, T, F, A, t0, K, M, q.rs, nullptr);
            launch_opt_fold(prep, q.rs, d_state, !initialised, F, M, K, L, smoothing, capon,
                            cfg.diagonal_load, cfg.forgetting_factor, q.scan, q.trace, q.scan64,
                            q.trace64, q.mode);
            initialised = true;
            hip_check(hipEventRecord(ev_ready[k], prep), "hipEventRecord");
            // search of window w on the caller's stream
            hip_check(hipStreamWaitEvent(stream, ev_ready[k], 0), "hipStreamWaitEvent");
            for (int level = 0; level <= last; ++level) {
                const size_t cells = level == 0 ? G * G : 9;
                launch_opt_scan(stream, q.scan, q.trace, q.scan64, q.trace64, q.mode, d_pos,
                                d_freqs, d_search, g, level, F, cells, capon, d_part);
                launch_opt_decide(stream, d_part, F, cells, g, level, last, w, d_search, d_dirs);
            }
            hip_check(hipEventRecord(ev_used[k], stream), "hipEventRecord");
        }
        L_emit.dir_source = kDirTable;
        L_emit.dir_table = d_dirs;
        launch_tracker(stream, L_emit);
        hip_check(hipEventRecord(ev_done, stream), "hipEventRecord");
        ran = true;
        n_windows = W;
        dirs_pending = true;
    }

    void fetch_dirs() {
        if (!dirs_pending) return;
        h_dirs.resize(3 * n_windows);
        hip_check(hipStreamSynchronize(last_stream), "hipStreamSynchronize");
        hip_check(hipMemcpy(h_dirs.data(), d_dirs, h_dirs.size() * sizeof(float),
                            hipMemcpyDeviceToHost),
                  "hipMemcpy directions");
        dirs_pending = false;
    }
};

namespace {

// Reference constructor checks (src/cpu_opt_beam_tracker.cpp:1084-1113) and
// this build's limits.
void create_tracker(const float* positions, const float* frequencies, const bf_dims* dp,
                    const bf_opt_tracker_config* c, std::unique_ptr<bf_opt_tracker>* out) {
    const bf_dims& d = deref_dims(dp);
    require_ptr(c, "config");
    validate_dimensions(d);
    if (d.n_beams != 1) invalid("tracker requires exactly n_beams == 1 (use tracker_beam_count)");
    if (d.n_ant != 32 && d.n_ant != 64)
        invalid("the adaptive tracker supports 32 or 64 antennas");
    if (c->integration_spectra == 0) invalid("tracker integration_spectra must be positive");
    if (!(c->forgetting_factor > 0.0f) || c->forgetting_factor > 1.0f)
        invalid("forgetting_factor must be in (0, 1]");
    if (c->spatial_smoothing_subarray_size > d.n_ant)
        invalid("spatial_smoothing_subarray_size must be <= n_ant");
    if (c->estimator != BF_ESTIMATOR_BARTLETT && c->estimator != BF_ESTIMATOR_CAPON)
        invalid("estimator must be BF_ESTIMATOR_BARTLETT or BF_ESTIMATOR_CAPON");
    if (c->coarse_grid_resolution > kMaxGrid)
        invalid("coarse_grid_resolution must be at most " + std::to_string(kMaxGrid));
    if (c->refinement_levels > kMaxLevels)
        invalid("refinement_levels must be at most " + std::to_string(kMaxLevels));
    if (!std::isfinite(c->search_fov_l) || !std::isfinite(c->search_fov_m) ||
        !std::isfinite(c->diagonal_load))
        invalid("search_fov_l, search_fov_m and diagonal_load must be finite");
    int kernel = c->kernel;
    if (kernel != BF_KERNEL_FP32 && kernel != BF_KERNEL_SPLIT_F16 && kernel != BF_KERNEL_F16 &&
        kernel != BF_KERNEL_AUTO)
        invalid("tracker kernel must be FP32, SPLIT_F16, F16 or AUTO");
    // AUTO stays AUTO: the one-beam emission then takes the same kernel as
    // the open-loop tracker's AUTO (the FP32 stream kernel).
    auto t = std::make_unique<bf_opt_tracker>();
    t->dims = d;
    t->cfg = *c;
    t->kernel = kernel;
    const size_t A = d.n_ant, F = d.n_freq;
    const size_t P = c->spatial_smoothing_subarray_size;
    t->smoothing = P != 0 && P < A;
    t->M = t->smoothing ? P : A;
    t->L = A - t->M + 1;
    t->G = c->coarse_grid_resolution;
    t->capon = c->estimator == BF_ESTIMATOR_CAPON;
    // Default geometry: default_positions(A) at 0.6 m and
    // channelized_frequencies(F) (src/geometry.cpp:53-95).
    std::vector<float> pos(3 * A), fr(F);
    for (size_t a = 0; a < A; ++a) {
        pos[3 * a] = positions ? positions[3 * a] : (float)(a % 8) * 0.6f;
        pos[3 * a + 1] = positions ? positions[3 * a + 1] : (float)(a / 8) * 0.6f;
        pos[3 * a + 2] = positions ? positions[3 * a + 2] : 0.0f;
    }
    for (size_t f = 0; f < F; ++f) {
        const float step = (float)f * 300000.0f;
        fr[f] = frequencies ? frequencies[f] : 300000000.0f + step;
    }
    for (float v : pos)
        if (!std::isfinite(v)) invalid("positions must be finite");
    for (float v : fr)
        if (!std::isfinite(v)) invalid("frequencies must be finite");
    // Coplanar aperture: the steered antennas all have |z| <= 1e-6 (:1197-1204).
    for (size_t a = 0; a < t->M; ++a)
        if (std::fabs(pos[3 * a + 2]) > 1.0e-6f) t->coplanar = false;
    const size_t M = t->M;
    t->d_pos = dmalloc<float>(pos.size());
    t->d_freqs = dmalloc<float>(F);
    hip_check(hipMemcpy(t->d_pos, pos.data(), pos.size() * sizeof(float), hipMemcpyHostToDevice),
              "hipMemcpy positions");
    hip_check(hipMemcpy(t->d_freqs, fr.data(), F * sizeof(float), hipMemcpyHostToDevice),
              "hipMemcpy frequencies");
    t->d_state = dmalloc<double>(F * M * M * 2);
    for (auto& q : t->slot) {
        q.rs = dmalloc<int>(F * M * M * 2);
        q.scan = dmalloc<float2>(F * M * M);
        q.trace = dmalloc<float>(F);
        if (t->capon) {
            q.scan64 = dmalloc<double2>(F * M * M);
            q.trace64 = dmalloc<double>(F);
        }
        q.mode = dmalloc<int>(F);
    }
    // default priority (a stream of another priority is never created here)
    hip_check(hipStreamCreateWithFlags(&t->prep, hipStreamNonBlocking), "hipStreamCreate");
    for (hipEvent_t* e : {&t->ev_done, &t->ev_start, &t->ev_ready[0], &t->ev_ready[1], &t->ev_used[0],
                          &t->ev_used[1]})
        hip_check(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreate");
    t->d_part = dmalloc<float>(F * std::max<size_t>(t->G * t->G, 9));
    t->d_search = dmalloc<OptSearch>(1);
    t->d_traj = dmalloc<float>(4);
    *out = std::move(t);
}

void seed(bf_opt_tracker* t, const float* start, const float* rate) {
    require_ptr(start, "direction_start");
    require_ptr(rate, "direction_rate_per_sample");
    // seed_trajectory (src/cpu_opt_beam_tracker.cpp:1335-1357)
    const double n2 = (double)start[0] * start[0] + (double)start[1] * start[1] +
                      (double)start[2] * start[2];
    if (!std::isfinite(n2) || std::fabs(n2 - 1.0) > 1.0e-3)
        invalid("seed_trajectory direction_start must be a finite unit vector");
    for (int i = 0; i < 2; ++i)
        if (!std::isfinite(rate[i]))
            invalid("seed_trajectory direction_rate_per_sample must be finite");
    std::memcpy(t->start, start, sizeof(t->start));
    std::memcpy(t->rate, rate, sizeof(t->rate));
}

void check_sizes(const bf_opt_tracker* t, size_t packed_count, size_t intensity_count) {
    const bf_dims& d = t->dims;
    if (packed_count < d.n_time * d.n_freq * d.n_ant)
        invalid("packed voltage is smaller than dimensions");
    if (intensity_count < d.n_time * d.n_freq)
        invalid("intensity output is smaller than dimensions");
}

}  // namespace

extern "C" {

bf_status bf_opt_tracker_default_config(bf_opt_tracker_config* c) {
    return guarded([&] {
        require_ptr(c, "config");
        // CpuOptTrackerConfig defaults (cpu_opt_beam_tracker.hpp:50-88)
        std::memset(c, 0, sizeof(*c));
        c->estimator = BF_ESTIMATOR_BARTLETT;
        c->coarse_grid_resolution = 1;
        c->refinement_levels = 0;
        c->search_fov_l = 0.2f;
        c->search_fov_m = 0.2f;
        c->spatial_smoothing_subarray_size = 0;
        c->forgetting_factor = 1.0f;
        c->diagonal_load = 1.0e-3f;
        c->enable_quadratic_peak_interp = 1;
        c->integration_spectra = 320;
        c->kernel = BF_KERNEL_AUTO;
    });
}

bf_status bf_opt_tracker_create(const float* positions, const float* frequencies,
                                const bf_dims* dims, const bf_opt_tracker_config* config,
                                bf_opt_tracker** out) {
    return guarded([&] {
        require_ptr(out, "out");
        *out = nullptr;
        std::unique_ptr<bf_opt_tracker> t;
        create_tracker(positions, frequencies, dims, config, &t);
        *out = t.release();
    });
}

bf_status bf_opt_tracker_destroy(bf_opt_tracker* t) {
    return guarded([&] { delete t; });
}

bf_status bf_opt_tracker_seed_trajectory(bf_opt_tracker* t, const float* direction_start,
                                         const float* direction_rate_per_sample) {
    return guarded([&] {
        require_ptr(t, "tracker");
        seed(t, direction_start, direction_rate_per_sample);
    });
}

bf_status bf_opt_tracker_reset(bf_opt_tracker* t) {
    return guarded([&] {
        require_ptr(t, "tracker");
        t->initialised = false;
    });
}

bf_status bf_opt_tracker_run_device(bf_opt_tracker* t, void* stream, const uint8_t* d_packed,
                                    float* d_intensity) {
    return guarded([&] {
        require_ptr(t, "tracker");
        require_ptr(d_packed, "packed");
        require_ptr(d_intensity, "intensity");
        t->run(static_cast<hipStream_t>(stream), d_packed, d_intensity);
    });
}

bf_status bf_opt_tracker_run(bf_opt_tracker* t, const uint8_t* packed, size_t packed_count,
                             float* intensity, size_t intensity_count) {
    return guarded([&] {
        require_ptr(t, "tracker");
        require_ptr(packed, "packed");
        require_ptr(intensity, "intensity");
        check_sizes(t, packed_count, intensity_count);
        const bf_dims& d = t->dims;
        const size_t nx = d.n_time * d.n_freq * d.n_ant, no = d.n_time * d.n_freq;
        uint8_t* dx = dmalloc<uint8_t>(nx);
        std::unique_ptr<uint8_t, void (*)(uint8_t*)> gx(dx, [](uint8_t* p) { (void)hipFree(p); });
        float* dout = dmalloc<float>(no);
        std::unique_ptr<float, void (*)(float*)> go(dout, [](float* p) { (void)hipFree(p); });
        hip_check(hipMemcpy(dx, packed, nx, hipMemcpyHostToDevice), "hipMemcpy packed");
        t->run(nullptr, dx, dout);
        hip_check(hipMemcpy(intensity, dout, no * sizeof(float), hipMemcpyDeviceToHost),
                  "hipMemcpy intensity");
        t->fetch_dirs();
    });
}

bf_status bf_opt_tracker_window_directions(bf_opt_tracker* t, float* dirs, size_t capacity,
                                           size_t* n_windows) {
    return guarded([&] {
        require_ptr(t, "tracker");
        require_ptr(n_windows, "n_windows");
        t->fetch_dirs();
        *n_windows = t->n_windows;
        if (dirs == nullptr) return;
        if (capacity < t->n_windows) invalid("direction buffer is smaller than the window count");
        std::memcpy(dirs, t->h_dirs.data(), t->h_dirs.size() * sizeof(float));
    });
}

bf_status bf_opt_tracker_packed_intensity(const uint8_t* packed, size_t packed_count,
                                          const bf_dims* dims,
                                          const bf_tracker_config* trajectory,
                                          const bf_opt_tracker_config* config, float* intensity,
                                          size_t intensity_count) {
    return guarded([&] {
        require_ptr(packed, "packed");
        require_ptr(intensity, "intensity");
        require_ptr(trajectory, "trajectory");
        require_ptr(config, "config");
        std::unique_ptr<bf_opt_tracker> t;
        bf_opt_tracker_config c = *config;
        c.integration_spectra = trajectory->integration_spectra;
        create_tracker(nullptr, nullptr, dims, &c, &t);
        seed(t.get(), trajectory->direction_start, trajectory->direction_rate_per_sample);
        check_sizes(t.get(), packed_count, intensity_count);
        const bf_dims& d = t->dims;
        const size_t nx = d.n_time * d.n_freq * d.n_ant, no = d.n_time * d.n_freq;
        uint8_t* dx = dmalloc<uint8_t>(nx);
        std::unique_ptr<uint8_t, void (*)(uint8_t*)> gx(dx, [](uint8_t* p) { (void)hipFree(p); });
        float* dout = dmalloc<float>(no);
        std::unique_ptr<float, void (*)(float*)> go(dout, [](float* p) { (void)hipFree(p); });
        hip_check(hipMemcpy(dx, packed, nx, hipMemcpyHostToDevice), "hipMemcpy packed");
        t->run(nullptr, dx, dout);
        hip_check(hipMemcpy(intensity, dout, no * sizeof(float), hipMemcpyDeviceToHost),
                  "hipMemcpy intensity");
    });
}

}  // extern "C"
