// capi_tracker.cpp -- extern "C" entry points of the dynamic tracker.
//
// Reference behaviour followed:
//   validation / pre-scan   src/cuda_beam_tracker_v5.cu:278-316,
//                           src/beam_tracker.cpp:18-21, 44-76
//   device-resident / into  src/cuda_beam_tracker_v5.cu:383-473
//   batched stream          src/cuda_beam_tracker_v5.cu:487-657
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "bf_common.hpp"
#include "bf_tracker.hpp"

namespace bfhip {
namespace {

void require_ptr(const void* p, const char* what) {
    if (p == nullptr) invalid(std::string(what) + " must not be null");
}

int tracker_kernel(const bf_tracker_exec* exec) {
    const int k = exec ? exec->kernel : BF_KERNEL_AUTO;
    if (k != BF_KERNEL_FP32 && k != BF_KERNEL_SPLIT_F16 && k != BF_KERNEL_F16 &&
        k != BF_KERNEL_AUTO)
        invalid("tracker kernel must be FP32, SPLIT_F16, F16 or AUTO");
    // AUTO stays AUTO: one beam with per-spectrum output takes the FP32
    // stream kernel, everything else the split-FP16 MFMA kernel.
    return k;
}

// Reference src/geometry.cpp:97-106 (direction_from_lm), float arithmetic.
bool on_unit_disk(float l, float m) {
    if (!std::isfinite(l) || !std::isfinite(m)) return false;
    const float ll = l * l;
    const float mm = m * m;
    return ll + mm <= 1.0f;
}

// Reference src/beam_tracker.cpp:18-50 (validate_tracker_config) and the
// per-window pre-scan of src/beam_tracker_opt_v2.cpp:112-122 for the windows
// [first_window, first_window + n_windows).
void validate_tracker(const bf_dims& d, const bf_tracker_config* cfg, size_t first_window,
                      size_t n_windows) {
    validate_dimensions(d);
    if (d.n_beams != 1) invalid("tracker requires exactly n_beams == 1 (use tracker_beam_count)");
    require_ptr(cfg, "tracker config");
    if (cfg->integration_spectra == 0)
        invalid("tracker integration_spectra must be positive");
    const float* s = cfg->direction_start;
    const double n2 = (double)s[0] * s[0] + (double)s[1] * s[1] + (double)s[2] * s[2];
    if (!std::isfinite(n2) || std::fabs(n2 - 1.0) > 1.0e-3)
        invalid("tracker direction_start must be a finite unit vector");
    for (float r : cfg->direction_rate_per_sample)
        if (!std::isfinite(r)) invalid("tracker direction_rate_per_sample must be finite");
    for (size_t w = first_window; w < first_window + n_windows; ++w) {
        const float tf = (float)(w * cfg->integration_spectra);
        const float dl = tf * cfg->direction_rate_per_sample[0];
        const float dm = tf * cfg->direction_rate_per_sample[1];
        if (!on_unit_disk(s[0] + dl, s[1] + dm))
            invalid("direction cosines must satisfy l*l + m*m <= 1");
    }
}

// Reference default geometry: default_positions(A) (src/geometry.cpp:53-67)
// at 0.6 m and channelized_frequencies (src/geometry.cpp:79-95).
Geom default_geometry(size_t A) {
    Geom g{};
    g.pos = nullptr;
    g.freqs = nullptr;
    g.cols = (A == 32 || A == 64) ? 8 : 16;
    g.spacing = 0.6f;
    g.f_start = 300000000.0f;
    g.f_step = 300000.0f;
    return g;
}

// Per-beam trajectory row (l0, m0, rate_l, rate_m) passed by value through a
// tiny device table owned by the caller's object or stream-ordered memory.
struct TrajTable {
    float v[4];
};

TrajTable traj_of(const bf_tracker_config& c) {
    return TrajTable{{c.direction_start[0], c.direction_start[1],
                      c.direction_rate_per_sample[0], c.direction_rate_per_sample[1]}};
}

TrackerLaunch single_beam_launch(const uint8_t* x, float* out, const bf_dims& d,
                                 const bf_tracker_config& cfg, const float* d_traj,
                                 size_t window_offset, int kernel) {
    TrackerLaunch L;
    L.x = x;
    L.out = out;
    L.T = d.n_time;
    L.F = d.n_freq;
    L.A = d.n_ant;
    L.B = 1;
    L.S = cfg.integration_spectra;
    L.n_windows = ceil_div(d.n_time, cfg.integration_spectra);
    L.geo = default_geometry(d.n_ant);
    L.dir_source = kDirTrajectory;
    L.trajectories = d_traj;
    L.window_offset = window_offset;
    L.integrate = false;
    L.kernel = kernel;
    return L;
}

}  // namespace
}  // namespace bfhip

using namespace bfhip;

struct bf_tracker_stream {
    bf_dims win_dims{};
    bf_dims batch_dims{};
    bf_tracker_config cfg{};
    size_t batch_size = 0;
    int kernel = BF_KERNEL_SPLIT_F16;
    bool use_graph = false;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    float last_ms = 0.0f;
    uint8_t* d_packed = nullptr;
    float* d_out = nullptr;
    float* d_traj = nullptr;      // [1][4]
    size_t* d_offset = nullptr;   // unused placeholder kept zero
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    size_t graph_window = (size_t)-1;

    ~bf_tracker_stream() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (d_packed) (void)hipFree(d_packed);
        if (d_out) (void)hipFree(d_out);
        if (d_traj) (void)hipFree(d_traj);
        if (start) (void)hipEventDestroy(start);
        if (stop) (void)hipEventDestroy(stop);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void enqueue(size_t first_window) {
        const TrackerLaunch L = single_beam_launch(d_packed, d_out, batch_dims, cfg, d_traj,
                                                   first_window, kernel);
        launch_tracker(stream, L);
    }

    void run(size_t first_window) {
        validate_tracker(batch_dims, &cfg, first_window, batch_size);
        if (use_graph) {
            // The window offset is a kernel argument: re-capture when the
            // batch position changes, replay otherwise.
            if (graph_exec == nullptr || graph_window != first_window) {
                if (graph_exec) hip_check(hipGraphExecDestroy(graph_exec), "hipGraphExecDestroy");
                if (graph) hip_check(hipGraphDestroy(graph), "hipGraphDestroy");
                graph_exec = nullptr;
                graph = nullptr;
                hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal),
                          "hipStreamBeginCapture");
                enqueue(first_window);
                hip_check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
                hip_check(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0),
                          "hipGraphInstantiate");
                graph_window = first_window;
            }
            hip_check(hipEventRecord(start, stream), "hipEventRecord");
            hip_check(hipGraphLaunch(graph_exec, stream), "hipGraphLaunch");
        } else {
            hip_check(hipEventRecord(start, stream), "hipEventRecord");
            enqueue(first_window);
        }
        hip_check(hipEventRecord(stop, stream), "hipEventRecord");
        hip_check(hipEventSynchronize(stop), "hipEventSynchronize");
        hip_check(hipEventElapsedTime(&last_ms, start, stop), "hipEventElapsedTime");
    }
};

extern "C" {

bf_status bf_beam_tracker_device_resident(const uint8_t* d_packed, float* d_intensity,
                                          const bf_dims* dims, const bf_tracker_config* tracker,
                                          const bf_tracker_exec* exec, void* stream) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        require_ptr(tracker, "tracker config");
        validate_tracker(d, tracker, 0, ceil_div(d.n_time, tracker->integration_spectra));
        require_ptr(d_packed, "d_packed");
        require_ptr(d_intensity, "d_intensity");
        if (reinterpret_cast<uintptr_t>(d_packed) % 16)
            invalid("d_packed must be 16-byte aligned");
        const int kernel = tracker_kernel(exec);
        hipStream_t s = static_cast<hipStream_t>(stream);
        // The trajectory row travels as a 16-byte kernel-visible table in
        // stream-ordered memory (released on the same stream).
        float* d_traj = nullptr;
        hip_check(hipMallocAsync(reinterpret_cast<void**>(&d_traj), 16, s), "hipMallocAsync");
        const TrajTable t = traj_of(*tracker);
        hip_check(hipMemcpyAsync(d_traj, t.v, sizeof(t.v), hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync trajectory");
        launch_tracker(s, single_beam_launch(d_packed, d_intensity, d, *tracker, d_traj, 0,
                                             kernel));
        hip_check(hipFreeAsync(d_traj, s), "hipFreeAsync");
        if (stream == nullptr) hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    });
}

bf_status bf_beam_tracker_into(const uint8_t* packed, size_t packed_bytes, const bf_dims* dims,
                               const bf_tracker_config* tracker, float* intensity,
                               size_t intensity_count, const bf_tracker_exec* exec) {
    return guarded([&] {
        const bf_dims& d = deref_dims(dims);
        require_ptr(tracker, "tracker config");
        validate_tracker(d, tracker, 0, ceil_div(d.n_time, tracker->integration_spectra));
        const size_t n_in = d.n_time * d.n_freq * d.n_ant;
        const size_t n_out = d.n_time * d.n_freq;
        if (packed == nullptr || packed_bytes != n_in)
            invalid("voltage buffer size does not match dimensions");
        if (intensity == nullptr || intensity_count < n_out)
            invalid("intensity output is smaller than dimensions");
        const int kernel = tracker_kernel(exec);
        hipStream_t s = nullptr;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        uint8_t* dx = nullptr;
        float* dout = nullptr;
        float* dtraj = nullptr;
        auto cleanup = [&] {
            if (dx) (void)hipFree(dx);
            if (dout) (void)hipFree(dout);
            if (dtraj) (void)hipFree(dtraj);
            (void)hipStreamDestroy(s);
        };
        try {
            hip_check(hipMalloc(reinterpret_cast<void**>(&dx), n_in), "hipMalloc d_packed");
            hip_check(hipMalloc(reinterpret_cast<void**>(&dout), n_out * sizeof(float)),
                      "hipMalloc d_intensity");
            hip_check(hipMalloc(reinterpret_cast<void**>(&dtraj), 16), "hipMalloc trajectory");
            const TrajTable t = traj_of(*tracker);
            hip_check(hipMemcpyAsync(dtraj, t.v, sizeof(t.v), hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync trajectory");
            hip_check(hipMemcpyAsync(dx, packed, n_in, hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync H2D d_packed");
            launch_tracker(s, single_beam_launch(dx, dout, d, *tracker, dtraj, 0, kernel));
            hip_check(hipMemcpyAsync(intensity, dout, n_out * sizeof(float),
                                     hipMemcpyDeviceToHost, s),
                      "hipMemcpyAsync D2H d_intensity");
            hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

bf_status bf_tracker_stream_create(const bf_dims* single_window_dims,
                                   const bf_tracker_config* tracker, size_t batch_size,
                                   const bf_tracker_exec* exec, bf_tracker_stream** out) {
    return guarded([&] {
        require_ptr(out, "out");
        *out = nullptr;
        const bf_dims& wd = deref_dims(single_window_dims);
        require_ptr(tracker, "tracker config");
        if (batch_size == 0) invalid("batch_size must be positive");
        bf_dims bd = wd;
        bd.n_time = wd.n_time * batch_size;
        validate_tracker(bd, tracker, 0, ceil_div(bd.n_time, tracker->integration_spectra));
        auto s = std::make_unique<bf_tracker_stream>();
        s->win_dims = wd;
        s->batch_dims = bd;
        s->cfg = *tracker;
        s->batch_size = batch_size;
        s->kernel = tracker_kernel(exec);
        s->use_graph = exec && exec->enable_graph;
        hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreate(&s->start), "hipEventCreate");
        hip_check(hipEventCreate(&s->stop), "hipEventCreate");
        hip_check(hipMalloc(reinterpret_cast<void**>(&s->d_packed),
                            bd.n_time * bd.n_freq * bd.n_ant),
                  "hipMalloc d_packed");
        hip_check(hipMalloc(reinterpret_cast<void**>(&s->d_out),
                            bd.n_time * bd.n_freq * sizeof(float)),
                  "hipMalloc d_intensity");
        hip_check(hipMalloc(reinterpret_cast<void**>(&s->d_traj), 16), "hipMalloc trajectory");
        const TrajTable t = traj_of(*tracker);
        hip_check(hipMemcpy(s->d_traj, t.v, sizeof(t.v), hipMemcpyHostToDevice),
                  "hipMemcpy trajectory");
        *out = s.release();
    });
}

bf_status bf_tracker_stream_destroy(bf_tracker_stream* s) {
    return guarded([&] { delete s; });
}

bf_status bf_tracker_stream_process_batch(bf_tracker_stream* s, size_t first_window,
                                          const uint8_t* host_packed, float* host_intensity) {
    return guarded([&] {
        require_ptr(s, "stream");
        require_ptr(host_packed, "host_packed");
        require_ptr(host_intensity, "host_intensity");
        const bf_dims& bd = s->batch_dims;
        const size_t n_in = bd.n_time * bd.n_freq * bd.n_ant;
        const size_t n_out = bd.n_time * bd.n_freq;
        hip_check(hipMemcpyAsync(s->d_packed, host_packed, n_in, hipMemcpyHostToDevice,
                                 s->stream),
                  "H2D d_packed");
        s->run(first_window);
        hip_check(hipMemcpyAsync(host_intensity, s->d_out, n_out * sizeof(float),
                                 hipMemcpyDeviceToHost, s->stream),
                  "D2H d_intensity");
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    });
}

bf_status bf_tracker_stream_process_batch_kernel_only(bf_tracker_stream* s,
                                                      size_t first_window) {
    return guarded([&] {
        require_ptr(s, "stream");
        s->run(first_window);
    });
}

bf_status bf_tracker_stream_last_kernel_time_ms(const bf_tracker_stream* s, float* out) {
    return guarded([&] {
        require_ptr(s, "stream");
        require_ptr(out, "out");
        *out = s->last_ms;
    });
}

bf_status bf_tracker_stream_device_packed_buffer(bf_tracker_stream* s, uint8_t** out) {
    return guarded([&] {
        require_ptr(s, "stream");
        require_ptr(out, "out");
        *out = s->d_packed;
    });
}

bf_status bf_tracker_stream_device_intensity_buffer(bf_tracker_stream* s, float** out) {
    return guarded([&] {
        require_ptr(s, "stream");
        require_ptr(out, "out");
        *out = s->d_out;
    });
}

bf_status bf_tracker_stream_device_stream(bf_tracker_stream* s, void** out) {
    return guarded([&] {
        require_ptr(s, "stream");
        require_ptr(out, "out");
        *out = s->stream;
    });
}

bf_status bf_tracker_stream_sizes(const bf_tracker_stream* s, size_t* batch_size,
                                  size_t* window_bytes, size_t* batch_voltage_bytes,
                                  size_t* batch_output_floats) {
    return guarded([&] {
        require_ptr(s, "stream");
        if (batch_size) *batch_size = s->batch_size;
        if (window_bytes) *window_bytes = s->win_dims.n_time * s->win_dims.n_freq * s->win_dims.n_ant;
        if (batch_voltage_bytes)
            *batch_voltage_bytes = s->batch_dims.n_time * s->batch_dims.n_freq * s->batch_dims.n_ant;
        if (batch_output_floats)
            *batch_output_floats = s->batch_dims.n_time * s->batch_dims.n_freq;
    });
}

bf_status bf_multibeam_tracker_launch(void* stream, const uint8_t* d_packed, float* d_out,
                                      const bf_multibeam_args* a) {
    return guarded([&] {
        require_ptr(a, "args");
        require_ptr(d_packed, "d_packed");
        require_ptr(d_out, "d_out");
        if (reinterpret_cast<uintptr_t>(d_packed) % 16)
            invalid("d_packed must be 16-byte aligned");
        if (a->n_time == 0) invalid("n_time must be positive");
        if (a->n_freq == 0) invalid("n_freq must be positive");
        if (a->n_ant != 32 && a->n_ant != 64 && a->n_ant != 128 && a->n_ant != 256)
            invalid("n_ant must be 32, 64, 128, or 256");
        if (a->n_beams == 0 || a->n_beams > kMaximumBeams)
            invalid("n_beams must be between 1 and 128");
        if (a->integration_spectra == 0) invalid("integration_spectra must be positive");
        if ((a->directions == nullptr) == (a->trajectories == nullptr))
            invalid("exactly one of directions and trajectories must be given");
        const int k = a->kernel == BF_KERNEL_AUTO ? BF_KERNEL_SPLIT_F16 : a->kernel;
        if (k != BF_KERNEL_FP32 && k != BF_KERNEL_SPLIT_F16 && k != BF_KERNEL_F16)
            invalid("tracker kernel must be FP32, SPLIT_F16, F16 or AUTO");
        TrackerLaunch L;
        L.x = d_packed;
        L.out = d_out;
        L.T = a->n_time;
        L.F = a->n_freq;
        L.A = a->n_ant;
        L.B = a->n_beams;
        L.S = a->integration_spectra;
        L.n_windows = ceil_div(a->n_time, a->integration_spectra);
        L.geo = default_geometry(a->n_ant);
        L.geo.pos = a->positions;
        L.geo.freqs = a->frequencies;
        L.dir_source = a->directions ? kDirTable : kDirTrajectory;
        L.dir_table = a->directions;
        L.trajectories = a->trajectories;
        L.window_offset = a->window_offset;
        L.integrate = a->integrate != 0;
        L.kernel = k;
        launch_tracker(static_cast<hipStream_t>(stream), L);
    });
}

}  // extern "C"
