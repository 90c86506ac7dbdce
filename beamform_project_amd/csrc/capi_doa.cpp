// capi_doa.cpp -- extern "C" entry points of the adaptive DOA path.
//
// Reference state semantics: the per-channel covariance recursion carried
// across run_into calls (include/beamformer/cpu_opt_beam_tracker.hpp:
// 190-196, src/cpu_opt_beam_tracker.cpp:1880-1892); here it is an explicit
// device state with get/set for checkpoint and resume.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "bf_doa.hpp"

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

}  // namespace
}  // namespace bfhip

using namespace bfhip;

struct bf_doa {
    bf_doa_config cfg{};
    size_t M = 0, L = 0;
    bool initialised = false;
    hipStream_t last_stream = nullptr;
    float* d_freqs = nullptr;
    double* d_theta = nullptr;
    double* d_sin = nullptr;
    int* d_rs = nullptr;  // [kCovParts][F][M][M][2] partial covariances
    size_t parts = 1;     // partial covariances the last update wrote
    double* d_state = nullptr;
    double* d_capon = nullptr;
    double* d_bart = nullptr;
    int* d_ok = nullptr;
    double* d_broadband = nullptr;
    double* d_peaks = nullptr;
    int* d_npeaks = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // around the 3 kernels

    ~bf_doa() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        for (void* p : {(void*)d_freqs, (void*)d_theta, (void*)d_sin, (void*)d_rs, (void*)d_state,
                        (void*)d_capon, (void*)d_bart, (void*)d_ok, (void*)d_broadband,
                        (void*)d_peaks, (void*)d_npeaks})
            if (p) (void)hipFree(p);
    }
    void sync() {
        hip_check(hipStreamSynchronize(last_stream), "hipStreamSynchronize");
    }
};

// Snapshot-range parts of the covariance (workgroups per channel).  Each
// part smooths its own partial sum, and at K = 1024, A = 64, M = 48 that
// repeated smoothing cost more than the added parallelism saved (4 parts:
// 49 us against 31 us for one on MI355X), so one part is used.
constexpr size_t kCovParts = 1;

extern "C" {

bf_status bf_doa_create(const bf_doa_config* c, bf_doa** out) {
    return guarded([&] {
        require_ptr(out, "out");
        *out = nullptr;
        require_ptr(c, "config");
        if (c->n_freq == 0) invalid("n_freq must be positive");
        if (c->n_ant != 32 && c->n_ant != 64) invalid("the DOA path supports 32 or 64 antennas");
        if (c->subarray == 0 || c->subarray > c->n_ant)
            invalid("subarray must satisfy 1 <= subarray <= n_ant");
        if (c->snapshots == 0 || c->snapshots > 65536)
            invalid("snapshots must be between 1 and 65536");
        if (c->n_theta < 3) invalid("n_theta must be at least 3");
        if (c->max_peaks == 0 || c->max_peaks > 64) invalid("max_peaks must be in 1..64");
        if (!(c->forgetting > 0.0 && c->forgetting <= 1.0))
            invalid("forgetting factor must be in (0, 1]");
        if (!(c->diagonal_load >= 0.0)) invalid("diagonal_load must be non-negative");
        if (!(c->theta_max_rad > c->theta_min_rad)) invalid("theta range must be increasing");
        if (!(c->element_spacing_m > 0.0)) invalid("element spacing must be positive");
        require_ptr(c->frequencies, "frequencies");
        auto d = std::make_unique<bf_doa>();
        d->cfg = *c;
        d->cfg.frequencies = nullptr;
        d->M = c->subarray;
        d->L = c->n_ant - c->subarray + 1;
        const size_t F = c->n_freq, M = d->M, n = c->n_theta;
        d->d_freqs = dmalloc<float>(F);
        hip_check(hipMemcpy(d->d_freqs, c->frequencies, F * sizeof(float), hipMemcpyHostToDevice),
                  "hipMemcpy frequencies");
        std::vector<double> th(n), sn(n);
        for (size_t q = 0; q < n; ++q) {
            th[q] = c->theta_min_rad + (c->theta_max_rad - c->theta_min_rad) * (double)q /
                                           (double)(n - 1);
            sn[q] = std::sin(th[q]);
        }
        d->d_theta = dmalloc<double>(n);
        d->d_sin = dmalloc<double>(n);
        hip_check(hipMemcpy(d->d_theta, th.data(), n * sizeof(double), hipMemcpyHostToDevice),
                  "hipMemcpy theta");
        hip_check(hipMemcpy(d->d_sin, sn.data(), n * sizeof(double), hipMemcpyHostToDevice),
                  "hipMemcpy sin theta");
        d->d_rs = dmalloc<int>(kCovParts * F * M * M * 2);
        d->d_state = dmalloc<double>(F * M * M * 2);
        hip_check(hipMemset(d->d_state, 0, F * M * M * 2 * sizeof(double)), "hipMemset state");
        d->d_capon = dmalloc<double>(F * n);
        d->d_bart = dmalloc<double>(F * n);
        d->d_ok = dmalloc<int>(F);
        d->d_broadband = dmalloc<double>(2 * n);
        d->d_peaks = dmalloc<double>(2 * c->max_peaks * 2);
        d->d_npeaks = dmalloc<int>(2);
        for (hipEvent_t& e : d->ev) hip_check(hipEventCreate(&e), "hipEventCreate");
        *out = d.release();
    });
}

bf_status bf_doa_destroy(bf_doa* d) {
    return guarded([&] { delete d; });
}

bf_status bf_doa_reset(bf_doa* d) {
    return guarded([&] {
        require_ptr(d, "doa");
        d->initialised = false;
    });
}

bf_status bf_doa_update(bf_doa* d, void* stream, const uint8_t* x, size_t n_time, size_t t0) {
    return guarded([&] {
        require_ptr(d, "doa");
        require_ptr(x, "d_packed");
        if (reinterpret_cast<uintptr_t>(x) % 4) invalid("d_packed must be 4-byte aligned");
        const bf_doa_config& c = d->cfg;
        if (t0 + c.snapshots > n_time) invalid("the update window exceeds n_time");
        hipStream_t s = static_cast<hipStream_t>(stream);
        hip_check(hipEventRecord(d->ev[0], s), "hipEventRecord");
        d->parts = launch_doa_covariance(s, x, n_time, c.n_freq, c.n_ant, t0, c.snapshots,
                                         d->M, d->d_rs, nullptr, kCovParts);
        hip_check(hipEventRecord(d->ev[1], s), "hipEventRecord");
        launch_doa_fold_scan(s, d->d_rs, d->parts, d->d_state, !d->initialised, c.n_freq, d->M,
                             c.snapshots, d->L, c.forgetting, c.diagonal_load, d->d_freqs,
                             c.element_spacing_m, c.n_theta, d->d_sin, d->d_capon, d->d_bart,
                             d->d_ok);
        hip_check(hipEventRecord(d->ev[2], s), "hipEventRecord");
        launch_doa_estimate(s, d->d_capon, d->d_bart, c.n_freq, c.n_theta, d->d_theta,
                            c.max_peaks, d->d_broadband, d->d_peaks, d->d_npeaks);
        hip_check(hipEventRecord(d->ev[3], s), "hipEventRecord");
        d->initialised = true;
        d->last_stream = s;
    });
}

bf_status bf_doa_last_timings(bf_doa* d, bf_doa_timings* t) {
    return guarded([&] {
        require_ptr(d, "doa");
        require_ptr(t, "timings");
        if (!d->initialised) fail(BF_ERR_LOGIC, "no update has run");
        hip_check(hipEventSynchronize(d->ev[3]), "hipEventSynchronize");
        float ms[3] = {0.0f, 0.0f, 0.0f};
        for (int i = 0; i < 3; ++i)
            hip_check(hipEventElapsedTime(&ms[i], d->ev[i], d->ev[i + 1]),
                      "hipEventElapsedTime");
        t->covariance_ms = ms[0];
        t->fold_scan_ms = ms[1];
        t->estimate_ms = ms[2];
    });
}

bf_status bf_doa_read_results(bf_doa* d, double* broadband, double* peaks, int* n_peaks) {
    return guarded([&] {
        require_ptr(d, "doa");
        d->sync();
        const size_t n = d->cfg.n_theta;
        if (broadband)
            hip_check(hipMemcpy(broadband, d->d_broadband, 2 * n * sizeof(double),
                                hipMemcpyDeviceToHost),
                      "hipMemcpy broadband");
        if (peaks)
            hip_check(hipMemcpy(peaks, d->d_peaks, 4 * d->cfg.max_peaks * sizeof(double),
                                hipMemcpyDeviceToHost),
                      "hipMemcpy peaks");
        if (n_peaks)
            hip_check(hipMemcpy(n_peaks, d->d_npeaks, 2 * sizeof(int), hipMemcpyDeviceToHost),
                      "hipMemcpy n_peaks");
    });
}

bf_status bf_doa_read_channel_spectra(bf_doa* d, double* capon, double* bartlett, int* ok) {
    return guarded([&] {
        require_ptr(d, "doa");
        d->sync();
        const size_t n = d->cfg.n_theta * d->cfg.n_freq;
        if (capon)
            hip_check(hipMemcpy(capon, d->d_capon, n * sizeof(double), hipMemcpyDeviceToHost),
                      "hipMemcpy capon");
        if (bartlett)
            hip_check(hipMemcpy(bartlett, d->d_bart, n * sizeof(double), hipMemcpyDeviceToHost),
                      "hipMemcpy bartlett");
        if (ok)
            hip_check(hipMemcpy(ok, d->d_ok, d->cfg.n_freq * sizeof(int), hipMemcpyDeviceToHost),
                      "hipMemcpy chol_ok");
    });
}

bf_status bf_doa_read_covariance(bf_doa* d, int32_t* rs) {
    return guarded([&] {
        require_ptr(d, "doa");
        require_ptr(rs, "rs");
        d->sync();
        const size_t n = d->cfg.n_freq * d->M * d->M * 2;
        std::vector<int32_t> all(d->parts * n);
        hip_check(hipMemcpy(all.data(), d->d_rs, all.size() * sizeof(int32_t),
                            hipMemcpyDeviceToHost),
                  "hipMemcpy covariance");
        for (size_t i = 0; i < n; ++i) {
            int32_t v = 0;
            for (size_t p = 0; p < d->parts; ++p) v += all[p * n + i];
            rs[i] = v;
        }
    });
}

bf_status bf_doa_get_state(bf_doa* d, double* state, int* initialised) {
    return guarded([&] {
        require_ptr(d, "doa");
        d->sync();
        if (state)
            hip_check(hipMemcpy(state, d->d_state,
                                d->cfg.n_freq * d->M * d->M * 2 * sizeof(double),
                                hipMemcpyDeviceToHost),
                      "hipMemcpy state");
        if (initialised) *initialised = d->initialised ? 1 : 0;
    });
}

bf_status bf_doa_set_state(bf_doa* d, const double* state, int initialised) {
    return guarded([&] {
        require_ptr(d, "doa");
        require_ptr(state, "state");
        d->sync();
        hip_check(hipMemcpy(d->d_state, state, d->cfg.n_freq * d->M * d->M * 2 * sizeof(double),
                            hipMemcpyHostToDevice),
                  "hipMemcpy state");
        d->initialised = initialised != 0;
    });
}

}  // extern "C"
