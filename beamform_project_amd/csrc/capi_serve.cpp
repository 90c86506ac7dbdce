// capi_serve.cpp -- extern "C" streaming tracker server.
//
// A superset of the reference's BatchedTrackerStreamV5
// (include/beamformer/cuda_beam_tracker_v5.hpp:90-125,
// src/cuda_beam_tracker_v5.cu:487-657), which runs fixed batches of one
// request synchronously.  Here voltages stream into a device ring, many
// tracking requests of different sizes share it, requests are admitted as
// request slots and output memory (an HBM arena of fixed budget) free up,
// and every step batches the windows that completed since the previous step
// -- across all active requests -- into one captured launch.
//
// Spectra reach the ring by copy (ingest, on an ingest stream) or are
// written in place by the caller's producer (acquire / commit on the
// producer's own stream); with loss masks enabled a validity ring [R][F]
// rides along and lost frames contribute nothing to any window.
//
// Threading: one host thread drives a server (ingest, step, read, release);
// the ring copy runs on an ingest stream and the steps on a step stream.
#include <hip/hip_runtime.h>

#include <climits>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

#include "bf_kernels.hpp"
#include "bf_serve.hpp"

namespace bfhip {
namespace {

void require_ptr(const void* p, const char* what) {
    if (p == nullptr) invalid(std::string(what) + " must not be null");
}

constexpr size_t kArenaAlign = 256;

// First-fit allocator over a byte range [0, size) with coalescing frees.
class Arena {
public:
    void reset(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
        size_ = size;
        used_ = peak_ = 0;
    }
    bool alloc(size_t bytes, size_t* offset) {
        bytes = (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *offset = it->first;
            const size_t rest = it->second - bytes;
            const size_t at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t offset, size_t bytes) {
        bytes = (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
        used_ -= bytes;
        auto it = free_.emplace(offset, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    void reset_peak() { peak_ = used_; }

private:
    std::map<size_t, size_t> free_;
    size_t size_ = 0, used_ = 0, peak_ = 0;
};

enum ReqState { kPending = 0, kActive = 1, kComplete = 2, kRejected = 3 };

struct Request {
    uint64_t id = 0;
    int state = kPending;
    size_t B = 0, S = 0, n_spectra = 0, out_windows = 0, windows_total = 0;
    long long start = -1;  // requested start (-1: head at admission)
    long long t0 = 0, t_end = 0;
    long long submit_head = 0;
    long long admit_head = 0;
    size_t windows_issued = 0, windows_done = 0;
    size_t arena_off = 0, arena_bytes = 0;
    int slot = -1;
    std::vector<float> traj;  // [B][4]
    std::string reject_reason;
};

struct Commit {
    long long h0, h1;
    hipEvent_t ev;
};

struct StepReq {
    uint64_t id;
    size_t w_lo, w_hi;
};

struct Step {
    long long P, H;
    hipEvent_t ev_start, ev_done;
    std::vector<StepReq> reqs;
    long long min_t;  // first spectrum any of its windows reads (LLONG_MAX: none)
};

// Nearest-rank percentile of (value, count) samples.
double percentile(std::vector<std::pair<float, uint32_t>>& v, double q, size_t total) {
    if (total == 0) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t rank = (size_t)std::ceil(q * (double)total);
    size_t seen = 0;
    for (const auto& e : v) {
        seen += e.second;
        if (seen >= std::max<size_t>(rank, 1)) return e.first;
    }
    return v.back().first;
}

// Direction validity of a linear trajectory over its window starts: l^2+m^2
// is convex along a line, so the first and last window starts bound it.
void validate_trajectory(const float* q, size_t n_windows, size_t S, size_t b) {
    for (int i = 0; i < 4; ++i)
        if (!std::isfinite(q[i])) invalid("trajectory " + std::to_string(b) + " is not finite");
    const size_t ends[2] = {0, n_windows - 1};
    for (size_t e : ends) {
        const float tf = (float)((long long)e * (long long)S);
        const float l = q[0] + tf * q[2];
        const float m = q[1] + tf * q[3];
        if (!(l * l + m * m <= 1.0f))
            invalid("trajectory " + std::to_string(b) + " leaves the sky (l*l + m*m > 1)");
    }
}

}  // namespace
}  // namespace bfhip

using namespace bfhip;

struct bf_server {
    bf_server_config cfg{};
    size_t F = 0, A = 0, R = 0;
    bool split = true;
    bool use_graph = false;
    hipStream_t step_stream = nullptr, ingest_stream = nullptr;
    uint8_t* d_ring = nullptr;
    uint8_t* d_valid = nullptr;  // [R][F] validity bytes when loss masks are on
    unsigned char* d_arena = nullptr;
    ServeSlot* d_slots = nullptr;
    ServeCtrl* d_ctrl = nullptr;
    ServeWork* d_work = nullptr;
    long long* h_heads = nullptr;  // pinned ring of step heads
    long long launched = 0;        // steps launched so far
    float* d_pos = nullptr;
    float* d_freqs = nullptr;
    ServeSlot* h_staging = nullptr;  // pinned, one descriptor per slot
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    ServeLaunch launch{};

    Arena arena;
    uint64_t next_id = 1;
    std::unordered_map<uint64_t, Request> requests;
    std::deque<uint64_t> pending;
    std::vector<uint64_t> slot_owner;  // 0: free
    std::vector<int> dirty_slots;
    size_t active_beams = 0;  // beams of the admitted, unfinished requests
    long long head = 0;       // spectra ingested
    long long processed = 0;  // head released to the last issued step
    long long done_head = 0;  // head of the last harvested step
    size_t acquired = 0;      // spectra handed to a producer by acquire, not committed
    std::deque<Commit> commits;
    std::deque<Step> steps;
    std::vector<hipEvent_t> event_pool;

    // statistics
    std::vector<std::pair<float, uint32_t>> lat;
    size_t lat_total = 0;
    double lat_sum = 0.0;
    float lat_max = 0.0f;
    size_t n_steps = 0, windows_processed = 0;
    double beam_samples = 0.0, busy_ms = 0.0;
    size_t n_submitted = 0, n_admitted = 0, n_completed = 0, n_rejected = 0;
    double admit_wait_sum = 0.0;

    ~bf_server() {
        if (step_stream) (void)hipStreamSynchronize(step_stream);
        if (ingest_stream) (void)hipStreamSynchronize(ingest_stream);
        for (auto& c : commits) event_pool.push_back(c.ev);
        for (auto& s : steps) {
            event_pool.push_back(s.ev_start);
            event_pool.push_back(s.ev_done);
        }
        for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : {(void*)d_ring, (void*)d_valid, (void*)d_arena, (void*)d_slots, (void*)d_ctrl,
                        (void*)d_work, (void*)d_pos, (void*)d_freqs})
            if (p) (void)hipFree(p);
        if (h_staging) (void)hipHostFree(h_staging);
        if (h_heads) (void)hipHostFree(h_heads);
        if (step_stream) (void)hipStreamDestroy(step_stream);
        if (ingest_stream) (void)hipStreamDestroy(ingest_stream);
    }

    hipEvent_t get_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }

    Request& req(uint64_t id) {
        auto it = requests.find(id);
        if (it == requests.end()) invalid("unknown request id " + std::to_string(id));
        return it->second;
    }

    size_t window_spectra(const Request& r, size_t w) const {
        const long long a = r.t0 + (long long)(w * r.S);
        const long long e = std::min(a + (long long)r.S, r.t_end);
        return (size_t)(e - a);
    }

    // Windows a step with head H issues: from the request's cursor (the
    // first window not yet issued) to the last window ending at or before H;
    // same rule as the plan kernel.
    static void window_range(const Request& r, long long H, size_t* lo, size_t* hi) {
        const long long S = (long long)r.S;
        const long long W = (long long)r.windows_total;
        const long long l = (long long)r.windows_issued;
        const long long h = H >= r.t_end ? W : (H < r.t0 ? 0 : (H - r.t0) / S);
        *lo = (size_t)l;
        *hi = (size_t)std::max(h, l);
    }

    // Commit holding spectrum t, or null when it has been retired.
    const Commit* commit_of(long long t) const {
        auto it = std::upper_bound(commits.begin(), commits.end(), t,
                                   [](long long v, const Commit& c) { return v < c.h1; });
        if (it == commits.end() || t < it->h0) return nullptr;
        return &*it;
    }

    void harvest_step(Step& s) {
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, s.ev_start, s.ev_done), "hipEventElapsedTime");
        busy_ms += ms;
        ++n_steps;
        for (const StepReq& sr : s.reqs) {
            auto it = requests.find(sr.id);
            if (it == requests.end()) continue;
            Request& r = it->second;
            const Commit* last = nullptr;
            float last_ms = 0.0f;
            for (size_t w = sr.w_lo; w < sr.w_hi; ++w) {
                const size_t n = window_spectra(r, w);
                beam_samples += (double)n * (double)F * (double)r.B;
                ++windows_processed;
                const long long end = r.t0 + (long long)(w * r.S) + (long long)n;
                if (end <= r.admit_head) continue;  // landed before admission
                const Commit* c = commit_of(end - 1);
                if (c == nullptr) continue;
                if (c != last) {
                    hip_check(hipEventElapsedTime(&last_ms, c->ev, s.ev_done),
                              "hipEventElapsedTime");
                    last = c;
                }
                if (!lat.empty() && lat.back().first == last_ms)
                    ++lat.back().second;
                else
                    lat.emplace_back(last_ms, 1u);
                ++lat_total;
                lat_sum += last_ms;
                lat_max = std::max(lat_max, last_ms);
            }
            r.windows_done = std::max(r.windows_done, sr.w_hi);
            if (r.state == kActive && r.windows_done == r.windows_total) {
                r.state = kComplete;
                active_beams -= r.B;
                ++n_completed;
            }
        }
        done_head = std::max(done_head, s.H);
    }

    // Retires finished steps in order; blocks on each when `wait`.
    void harvest(bool wait) {
        while (!steps.empty()) {
            Step& s = steps.front();
            if (wait) {
                hip_check(hipEventSynchronize(s.ev_done), "hipEventSynchronize");
            } else {
                const hipError_t q = hipEventQuery(s.ev_done);
                if (q == hipErrorNotReady) break;
                hip_check(q, "hipEventQuery");
            }
            harvest_step(s);
            event_pool.push_back(s.ev_start);
            event_pool.push_back(s.ev_done);
            steps.pop_front();
        }
        // A commit is no longer needed once every window ending in it has
        // been harvested.
        while (!commits.empty() && commits.front().h1 <= done_head) {
            if (hipEventQuery(commits.front().ev) != hipSuccess) break;
            event_pool.push_back(commits.front().ev);
            commits.pop_front();
        }
    }

    void upload_slot(int s) {
        hip_check(hipMemcpyAsync(d_slots + s, h_staging + s, sizeof(ServeSlot),
                                 hipMemcpyHostToDevice, step_stream),
                  "hipMemcpyAsync slot");
    }

    void admit() {
        while (!pending.empty()) {
            Request& r = req(pending.front());
            // Spectra below head - R have been overwritten.  Every window of a
            // request admitted here that ends at or before the step's head is
            // issued by this same step, before any later ingest can reach it
            // (the overrun guard waits for steps that read old spectra).
            if (r.start >= 0 && r.start < head - (long long)R) {
                r.state = kRejected;
                r.reject_reason = "start spectrum is no longer held by the ring";
                ++n_rejected;
                pending.pop_front();
                continue;
            }
            int s = -1;
            for (size_t i = 0; i < cfg.max_requests; ++i)
                if (slot_owner[i] == 0) {
                    s = (int)i;
                    break;
                }
            if (s < 0) return;
            if (cfg.max_active_beams != 0 && active_beams + r.B > cfg.max_active_beams)
                return;  // first in, first admitted
            size_t off = 0;
            if (!arena.alloc(r.arena_bytes, &off)) return;  // first in, first admitted
            r.arena_off = off;
            r.slot = s;
            r.t0 = r.start >= 0 ? r.start : head;
            r.t_end = r.t0 + (long long)r.n_spectra;
            r.admit_head = head;
            r.state = kActive;
            active_beams += r.B;
            slot_owner[s] = r.id;
            ServeSlot& d = h_staging[s];
            std::memset(&d, 0, sizeof(d));
            d.active = 1;
            d.B = (int)r.B;
            d.S = (int)r.S;
            d.wpi = serve_windows_per_item((int)r.S);
            d.t0 = r.t0;
            d.t_end = r.t_end;
            d.out_windows = (long long)r.out_windows;
            d.out = reinterpret_cast<float*>(d_arena + off);
            std::memcpy(&d.traj[0][0], r.traj.data(), r.traj.size() * sizeof(float));
            upload_slot(s);
            ++n_admitted;
            admit_wait_sum += (double)std::max(0LL, head - r.submit_head);
            pending.pop_front();
        }
    }

    size_t step() {
        harvest(false);
        for (int s : dirty_slots) upload_slot(s);
        dirty_slots.clear();
        admit();
        const long long H =
            std::max(processed, std::min(head, processed + (long long)cfg.max_step_spectra));
        // A step runs when it releases new spectra or when a request
        // admitted with a start behind the processed head has windows to
        // back-fill.
        Step st{processed, H, nullptr, nullptr, {}, LLONG_MAX};
        size_t n_windows = 0;
        for (size_t i = 0; i < cfg.max_requests; ++i) {
            if (slot_owner[i] == 0) continue;
            Request& r = req(slot_owner[i]);
            size_t lo, hi;
            window_range(r, H, &lo, &hi);
            if (hi > lo) {
                st.reqs.push_back(StepReq{r.id, lo, hi});
                r.windows_issued = hi;
                n_windows += hi - lo;
                st.min_t = std::min(st.min_t, r.t0 + (long long)(lo * r.S));
            }
        }
        if (H <= processed && n_windows == 0) return 0;
        st.ev_start = get_event();
        st.ev_done = get_event();
        // Every commit landing spectra in [processed, H): ingest copies are
        // ordered on one stream, but producer commits come from streams of
        // the caller's choosing.  Earlier spectra were covered by earlier
        // steps, which precede this one on the step stream.
        {
            auto it = std::upper_bound(commits.begin(), commits.end(), processed,
                                       [](long long v, const Commit& c) { return v < c.h1; });
            for (; it != commits.end() && it->h0 < H; ++it)
                hip_check(hipStreamWaitEvent(step_stream, it->ev, 0), "hipStreamWaitEvent");
        }
        // Bound the steps in flight so a head-ring entry is never reused
        // before its step has read it.
        while (steps.size() >= (size_t)kServeHeadRing / 2) {
            hip_check(hipEventSynchronize(steps.front().ev_done), "hipEventSynchronize");
            harvest(false);
        }
        h_heads[launched & (kServeHeadRing - 1)] = H;
        ++launched;
        hip_check(hipEventRecord(st.ev_start, step_stream), "hipEventRecord");
        if (use_graph) {
            if (graph_exec == nullptr) {
                hip_check(hipStreamBeginCapture(step_stream, hipStreamCaptureModeThreadLocal),
                          "hipStreamBeginCapture");
                launch_serve_step(step_stream, launch);
                hip_check(hipStreamEndCapture(step_stream, &graph), "hipStreamEndCapture");
                hip_check(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0),
                          "hipGraphInstantiate");
            }
            hip_check(hipGraphLaunch(graph_exec, step_stream), "hipGraphLaunch");
        } else {
            launch_serve_step(step_stream, launch);
        }
        hip_check(hipEventRecord(st.ev_done, step_stream), "hipEventRecord");
        steps.push_back(std::move(st));
        processed = H;
        return n_windows;
    }

    // The step that must finish before spectra [head, head + n) may be
    // written: spectra below X = head + n - R are overwritten, so every
    // window holding one must have been issued, i.e. by a step with
    // H >= X + max_window, and every step reading one must have finished
    // (a step that back-fills a request admitted with an early start reads
    // spectra far behind its head).  Steps finish in order, so the later of
    // the two is the one to wait for.  Null when both have been harvested.
    const Step* overrun_guard(size_t n) const {
        if (n + cfg.max_window > R)
            invalid("writing " + std::to_string(n) +
                    " spectra at once exceeds the ring capacity less the longest window");
        const long long X = head + (long long)n - (long long)R;
        if (X <= 0) return nullptr;
        const long long need = X + (long long)cfg.max_window;
        long long wait_at = -1;
        if (need > done_head) {
            for (size_t i = 0; i < steps.size() && wait_at < 0; ++i)
                if (steps[i].H >= need) wait_at = (long long)i;
            if (wait_at < 0)
                fail(BF_ERR_LOGIC, "ring overrun: step the server before ingesting more spectra");
        }
        for (size_t i = steps.size(); i-- > 0;)
            if (steps[i].min_t < X) {
                wait_at = std::max(wait_at, (long long)i);
                break;
            }
        return wait_at < 0 ? nullptr : &steps[(size_t)wait_at];
    }

    // Copies `rows` rows of `row` bytes into ring rows [at, at + rows),
    // split at the wrap; a null source fills with `fill`.
    void ring_copy(uint8_t* base, size_t row, size_t at, const uint8_t* src, size_t rows,
                   hipMemcpyKind kind, unsigned char fill) {
        const size_t first = std::min(rows, R - at);
        const size_t parts[2][2] = {{at, first}, {0, rows - first}};
        size_t done = 0;
        for (const auto& pr : parts) {
            if (pr[1] == 0) continue;
            if (src != nullptr)
                hip_check(hipMemcpyAsync(base + pr[0] * row, src + done * row, pr[1] * row, kind,
                                         ingest_stream),
                          "hipMemcpyAsync ring");
            else
                hip_check(hipMemsetAsync(base + pr[0] * row, fill, pr[1] * row, ingest_stream),
                          "hipMemsetAsync ring");
            done += pr[1];
        }
    }

    void ingest(const uint8_t* src, const uint8_t* mask, size_t n, bool on_device) {
        if (acquired != 0)
            fail(BF_ERR_LOGIC, "an acquired ring span must be committed before ingesting");
        if (mask != nullptr && d_valid == nullptr)
            invalid("loss masks need a server created with loss_masks = 1");
        if (n == 0) return;
        const Step* guard = overrun_guard(n);
        if (guard != nullptr)
            hip_check(hipStreamWaitEvent(ingest_stream, guard->ev_done, 0), "hipStreamWaitEvent");
        const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        const size_t at = (size_t)(head & (long long)(R - 1));
        ring_copy(d_ring, F * A, at, src, n, kind, 0);
        // Spectra ingested without a mask are valid in every channel.
        if (d_valid != nullptr) ring_copy(d_valid, F, at, mask, n, kind, 1);
        Commit c{head, head + (long long)n, get_event()};
        hip_check(hipEventRecord(c.ev, ingest_stream), "hipEventRecord");
        commits.push_back(c);
        head += (long long)n;
        // A host source may be reused by the caller as soon as this returns.
        if (!on_device) hip_check(hipEventSynchronize(c.ev), "hipEventSynchronize");
    }

    // Hands the producer the ring rows of spectra [head, head + n): one span,
    // or two when they wrap.  Blocks the host until no step still reads
    // them.  Spectra become visible to steps at commit.
    void acquire(size_t n, bf_ring_span spans[2], size_t* n_spans) {
        if (acquired != 0) fail(BF_ERR_LOGIC, "a ring span is already acquired");
        if (n == 0) invalid("acquire needs at least one spectrum");
        const Step* guard = overrun_guard(n);
        if (guard != nullptr) hip_check(hipEventSynchronize(guard->ev_done), "hipEventSynchronize");
        const size_t at = (size_t)(head & (long long)(R - 1));
        const size_t first = std::min(n, R - at);
        *n_spans = first < n ? 2 : 1;
        const size_t rows[2] = {first, n - first};
        const size_t starts[2] = {at, 0};
        long long t = head;
        for (size_t i = 0; i < *n_spans; ++i) {
            spans[i].first_spectrum = t;
            spans[i].n_spectra = rows[i];
            spans[i].voltages = d_ring + starts[i] * F * A;
            spans[i].valid = d_valid ? d_valid + starts[i] * F : nullptr;
            t += (long long)rows[i];
        }
        acquired = n;
    }

    // Publishes the first n acquired spectra once the producer's work
    // queued so far on `stream` has finished.
    void commit(size_t n, hipStream_t stream) {
        if (acquired == 0) fail(BF_ERR_LOGIC, "commit without an acquired ring span");
        if (n > acquired) invalid("commit of more spectra than were acquired");
        acquired = 0;
        if (n == 0) return;
        Commit c{head, head + (long long)n, get_event()};
        hip_check(hipEventRecord(c.ev, stream), "hipEventRecord");
        commits.push_back(c);
        head += (long long)n;
    }

    void release(uint64_t id) {
        Request& r = req(id);
        if (r.state == kPending) {
            pending.erase(std::find(pending.begin(), pending.end(), id));
        } else if (r.slot >= 0) {
            if (r.state != kComplete) {
                // Cancel: no step may still write into its outputs.
                hip_check(hipStreamSynchronize(step_stream), "hipStreamSynchronize");
                harvest(true);
                if (r.state == kActive) active_beams -= r.B;
            }
            arena.release(r.arena_off, r.arena_bytes);
            h_staging[r.slot].active = 0;
            dirty_slots.push_back(r.slot);
            slot_owner[(size_t)r.slot] = 0;
        }
        requests.erase(id);
    }

    void fill_stats(bf_server_stats* s) {
        std::memset(s, 0, sizeof(*s));
        s->windows_timed = lat_total;
        s->latency_p50_ms = percentile(lat, 0.50, lat_total);
        s->latency_p90_ms = percentile(lat, 0.90, lat_total);
        s->latency_p99_ms = percentile(lat, 0.99, lat_total);
        s->latency_max_ms = lat_max;
        s->latency_mean_ms = lat_total ? lat_sum / (double)lat_total : 0.0;
        s->steps = n_steps;
        s->windows_processed = windows_processed;
        s->beam_samples = beam_samples;
        s->gpu_busy_ms = busy_ms;
        s->requests_submitted = n_submitted;
        s->requests_admitted = n_admitted;
        s->requests_completed = n_completed;
        s->requests_rejected = n_rejected;
        s->requests_pending = pending.size();
        size_t active = 0;
        for (uint64_t o : slot_owner) active += o != 0;
        s->requests_in_slots = active;
        s->arena_bytes_in_use = arena.used();
        s->arena_peak_bytes = arena.peak();
        s->head = head;
        s->processed_head = processed;
        s->mean_admission_wait_spectra = n_admitted ? admit_wait_sum / (double)n_admitted : 0.0;
    }

    void reset_stats() {
        lat.clear();
        lat_total = 0;
        lat_sum = 0.0;
        lat_max = 0.0f;
        n_steps = windows_processed = 0;
        beam_samples = busy_ms = 0.0;
        n_submitted = n_admitted = n_completed = n_rejected = 0;
        admit_wait_sum = 0.0;
        arena.reset_peak();
    }
};

namespace {

void create_server(const bf_server_config* c, std::unique_ptr<bf_server>* out) {
    require_ptr(c, "config");
    if (c->n_ant != 32 && c->n_ant != 64) invalid("the serving loop supports 32 or 64 antennas");
    if (c->n_freq == 0 || c->n_freq > 65535) invalid("n_freq must be in 1..65535");
    const size_t R = c->ring_spectra;
    if (R < 32 || (R & (R - 1)) != 0) invalid("ring_spectra must be a power of two >= 32");
    if (c->max_window == 0 || c->max_window * 2 > R)
        invalid("max_window must be in 1..ring_spectra/2");
    if (c->max_step_spectra == 0) invalid("max_step_spectra must be positive");
    if (c->max_requests == 0 || c->max_requests > (size_t)kServeMaxSlots)
        invalid("max_requests must be in 1.." + std::to_string(kServeMaxSlots));
    if (c->hbm_budget_bytes < kArenaAlign) invalid("hbm_budget_bytes is too small");
    int kernel = c->kernel;
    if (kernel == BF_KERNEL_AUTO) kernel = BF_KERNEL_SPLIT_F16;
    if (kernel != BF_KERNEL_SPLIT_F16 && kernel != BF_KERNEL_F16)
        invalid("the serving loop runs the SplitF16 or F16 MFMA kernels");
    auto s = std::make_unique<bf_server>();
    s->cfg = *c;
    s->cfg.positions = nullptr;
    s->cfg.frequencies = nullptr;
    s->cfg.shard = nullptr;
    bf_shard shard{0, kFrequencyShardCount, kLocalFrequencyChannels, 0, 0, 1, 0, 1};
    if (c->shard != nullptr) {
        if (c->n_freq != kLocalFrequencyChannels)
            invalid("a shard descriptor needs n_freq = 336 (one local shard)");
        validate_shard(*c->shard);
        shard = *c->shard;
    }
    s->F = c->n_freq;
    s->A = c->n_ant;
    s->R = R;
    s->split = kernel == BF_KERNEL_SPLIT_F16;
    s->use_graph = c->enable_graph != 0;
    hip_check(hipStreamCreateWithFlags(&s->step_stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&s->ingest_stream, hipStreamNonBlocking),
              "hipStreamCreate");
    const size_t ring_bytes = R * s->F * s->A;
    hip_check(hipMalloc(reinterpret_cast<void**>(&s->d_ring), ring_bytes), "hipMalloc ring");
    hip_check(hipMemset(s->d_ring, 0, ring_bytes), "hipMemset ring");
    if (c->loss_masks != 0) {
        hip_check(hipMalloc(reinterpret_cast<void**>(&s->d_valid), R * s->F), "hipMalloc valid");
        hip_check(hipMemset(s->d_valid, 1, R * s->F), "hipMemset valid");
    }
    hip_check(hipMalloc(reinterpret_cast<void**>(&s->d_arena), c->hbm_budget_bytes),
              "hipMalloc arena");
    hip_check(hipMalloc(reinterpret_cast<void**>(&s->d_slots),
                        kServeMaxSlots * sizeof(ServeSlot)),
              "hipMalloc slots");
    hip_check(hipMemset(s->d_slots, 0, kServeMaxSlots * sizeof(ServeSlot)), "hipMemset slots");
    hip_check(hipMalloc(reinterpret_cast<void**>(&s->d_ctrl), sizeof(ServeCtrl)),
              "hipMalloc ctrl");
    hip_check(hipMemset(s->d_ctrl, 0, sizeof(ServeCtrl)), "hipMemset ctrl");
    hip_check(hipMalloc(reinterpret_cast<void**>(&s->d_work), sizeof(ServeWork)),
              "hipMalloc work");
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&s->h_staging),
                            kServeMaxSlots * sizeof(ServeSlot), hipHostMallocDefault),
              "hipHostMalloc staging");
    std::memset(s->h_staging, 0, kServeMaxSlots * sizeof(ServeSlot));
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&s->h_heads),
                            kServeHeadRing * sizeof(long long),
                            hipHostMallocMapped | hipHostMallocCoherent),
              "hipHostMalloc heads");
    std::memset(s->h_heads, 0, kServeHeadRing * sizeof(long long));
    long long* d_heads = nullptr;
    hip_check(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_heads), s->h_heads, 0),
              "hipHostGetDevicePointer heads");
    // Geometry tables; the defaults are the reference's regular grid of 8
    // columns at 0.6 m and channels 300 MHz + k * 300 kHz
    // (src/geometry.cpp:33-95), evaluated in float as the tracker does.
    std::vector<float> pos(s->A * 3), fr(s->F);
    for (size_t a = 0; a < s->A; ++a) {
        pos[3 * a] = c->positions ? c->positions[3 * a] : (float)(a % 8) * 0.6f;
        pos[3 * a + 1] = c->positions ? c->positions[3 * a + 1] : (float)(a / 8) * 0.6f;
        pos[3 * a + 2] = c->positions ? c->positions[3 * a + 2] : 0.0f;
    }
    // channels shard.absolute_frequency_start + f of the full-band plan
    for (size_t f = 0; f < s->F; ++f) {
        const float step = (float)(shard.absolute_frequency_start + f) * 300000.0f;
        fr[f] = c->frequencies ? c->frequencies[f] : 300000000.0f + step;
    }
    hip_check(hipMalloc(reinterpret_cast<void**>(&s->d_pos), pos.size() * sizeof(float)),
              "hipMalloc positions");
    hip_check(hipMemcpy(s->d_pos, pos.data(), pos.size() * sizeof(float), hipMemcpyHostToDevice),
              "hipMemcpy positions");
    hip_check(hipMalloc(reinterpret_cast<void**>(&s->d_freqs), fr.size() * sizeof(float)),
              "hipMalloc frequencies");
    hip_check(hipMemcpy(s->d_freqs, fr.data(), fr.size() * sizeof(float), hipMemcpyHostToDevice),
              "hipMemcpy frequencies");
    Geom g{};
    g.pos = s->d_pos;
    g.freqs = s->d_freqs;
    g.cols = 8;
    g.spacing = 0.6f;
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    s->arena.reset(c->hbm_budget_bytes);
    s->slot_owner.assign(c->max_requests, 0);
    ServeLaunch& L = s->launch;
    L.ring = s->d_ring;
    L.valid = s->d_valid;
    L.ring_mask = (long long)(R - 1);
    L.F = s->F;
    L.A = s->A;
    L.geo = g;
    L.slots = s->d_slots;
    L.ctrl = s->d_ctrl;
    L.work = s->d_work;
    L.heads = d_heads;
    hipDeviceProp_t prop{};
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    L.grid = std::max(1, prop.multiProcessorCount) * 2;  // resident at 2 workgroups per CU
    L.split = s->split;
    *out = std::move(s);
}

void submit_request(bf_server* s, const bf_server_request* q, uint64_t* id) {
    require_ptr(q, "request");
    require_ptr(id, "id");
    require_ptr(q->trajectories, "trajectories");
    if (q->n_beams == 0 || q->n_beams > (size_t)kServeMaxBeams)
        invalid("n_beams must be in 1.." + std::to_string(kServeMaxBeams));
    if (s->cfg.max_active_beams != 0 && q->n_beams > s->cfg.max_active_beams)
        invalid("n_beams exceeds the server's max_active_beams");
    if (q->window == 0 || q->window > s->cfg.max_window)
        invalid("window must be in 1..max_window (" + std::to_string(s->cfg.max_window) + ")");
    if (q->n_spectra == 0) invalid("n_spectra must be positive");
    if (q->start < -1) invalid("start must be -1 (the head at admission) or a spectrum index");
    const size_t W = (q->n_spectra + q->window - 1) / q->window;
    if (W > (size_t)INT32_MAX / 2) fail(BF_ERR_OVERFLOW, "request has too many windows");
    for (size_t b = 0; b < q->n_beams; ++b) validate_trajectory(q->trajectories + 4 * b, W, q->window, b);
    Request r;
    r.id = s->next_id++;
    r.B = q->n_beams;
    r.S = q->window;
    r.n_spectra = q->n_spectra;
    r.start = q->start;
    r.windows_total = W;
    r.out_windows = q->output_windows == 0 ? W : std::min(W, q->output_windows);
    const double bytes = (double)r.out_windows * (double)s->F * (double)r.B * sizeof(float);
    if (bytes > (double)s->cfg.hbm_budget_bytes)
        fail(BF_ERR_OVERFLOW, "request outputs exceed the server's HBM budget");
    r.arena_bytes = (size_t)bytes;
    r.traj.assign(q->trajectories, q->trajectories + 4 * q->n_beams);
    r.submit_head = s->head;
    *id = r.id;
    s->pending.push_back(r.id);
    s->requests.emplace(r.id, std::move(r));
    ++s->n_submitted;
}

}  // namespace

extern "C" {

bf_status bf_server_create(const bf_server_config* c, bf_server** out) {
    return guarded([&] {
        require_ptr(out, "out");
        *out = nullptr;
        std::unique_ptr<bf_server> s;
        create_server(c, &s);
        *out = s.release();
    });
}

bf_status bf_server_destroy(bf_server* s) {
    return guarded([&] { delete s; });
}

bf_status bf_server_submit(bf_server* s, const bf_server_request* q, uint64_t* id) {
    return guarded([&] {
        require_ptr(s, "server");
        submit_request(s, q, id);
    });
}

bf_status bf_server_ingest(bf_server* s, const uint8_t* src, size_t n_spectra, int on_device) {
    return guarded([&] {
        require_ptr(s, "server");
        require_ptr(src, "src");
        s->ingest(src, nullptr, n_spectra, on_device != 0);
    });
}

bf_status bf_server_ingest_masked(bf_server* s, const uint8_t* src, const uint8_t* valid,
                                  size_t n_spectra, int on_device) {
    return guarded([&] {
        require_ptr(s, "server");
        require_ptr(src, "src");
        require_ptr(valid, "valid");
        s->ingest(src, valid, n_spectra, on_device != 0);
    });
}

bf_status bf_server_acquire(bf_server* s, size_t n_spectra, bf_ring_span* spans,
                            size_t* n_spans) {
    return guarded([&] {
        require_ptr(s, "server");
        require_ptr(spans, "spans");
        require_ptr(n_spans, "n_spans");
        s->acquire(n_spectra, spans, n_spans);
    });
}

bf_status bf_server_commit(bf_server* s, size_t n_spectra, void* producer_stream) {
    return guarded([&] {
        require_ptr(s, "server");
        s->commit(n_spectra, static_cast<hipStream_t>(producer_stream));
    });
}

bf_status bf_server_step(bf_server* s, size_t* windows) {
    return guarded([&] {
        require_ptr(s, "server");
        const size_t n = s->step();
        if (windows) *windows = n;
    });
}

bf_status bf_server_request_status(bf_server* s, uint64_t id, bf_request_status* st) {
    return guarded([&] {
        require_ptr(s, "server");
        require_ptr(st, "status");
        s->harvest(false);
        const Request& r = s->req(id);
        std::memset(st, 0, sizeof(*st));
        st->state = r.state;
        st->start = r.state == kPending ? r.start : r.t0;
        st->windows_total = r.windows_total;
        st->windows_issued = r.windows_issued;
        st->windows_done = r.windows_done;
        st->output_windows = r.out_windows;
        st->output_bytes = r.arena_bytes;
    });
}

bf_status bf_server_read_output(bf_server* s, uint64_t id, size_t first_window,
                                size_t n_windows, float* host_out) {
    return guarded([&] {
        require_ptr(s, "server");
        require_ptr(host_out, "host_out");
        Request& r = s->req(id);
        if (r.state != kActive && r.state != kComplete)
            fail(BF_ERR_LOGIC, "request " + std::to_string(id) + " has no outputs");
        if (first_window + n_windows > r.windows_issued)
            fail(BF_ERR_OUT_OF_RANGE, "windows not yet processed");
        if (first_window + r.out_windows < r.windows_issued)
            fail(BF_ERR_OUT_OF_RANGE, "windows already overwritten in the output ring");
        if (first_window + n_windows > r.windows_done) {
            hip_check(hipStreamSynchronize(s->step_stream), "hipStreamSynchronize");
            s->harvest(true);
        }
        const size_t win_floats = s->F * r.B;
        const float* base = reinterpret_cast<const float*>(s->d_arena + r.arena_off);
        size_t w = first_window;
        float* dst = host_out;
        while (w < first_window + n_windows) {
            const size_t slot = w % r.out_windows;
            const size_t run = std::min(first_window + n_windows - w, r.out_windows - slot);
            hip_check(hipMemcpy(dst, base + slot * win_floats, run * win_floats * sizeof(float),
                                hipMemcpyDeviceToHost),
                      "hipMemcpy outputs");
            dst += run * win_floats;
            w += run;
        }
    });
}

bf_status bf_server_release(bf_server* s, uint64_t id) {
    return guarded([&] {
        require_ptr(s, "server");
        s->release(id);
    });
}

bf_status bf_server_synchronize(bf_server* s) {
    return guarded([&] {
        require_ptr(s, "server");
        hip_check(hipStreamSynchronize(s->ingest_stream), "hipStreamSynchronize");
        hip_check(hipStreamSynchronize(s->step_stream), "hipStreamSynchronize");
        s->harvest(true);
    });
}

bf_status bf_server_get_stats(bf_server* s, bf_server_stats* st) {
    return guarded([&] {
        require_ptr(s, "server");
        require_ptr(st, "stats");
        s->harvest(false);
        s->fill_stats(st);
    });
}

bf_status bf_server_reset_stats(bf_server* s) {
    return guarded([&] {
        require_ptr(s, "server");
        s->reset_stats();
    });
}

bf_status bf_server_simulate(const bf_server_sim_config* c, bf_server_sim_report* rep) {
    return guarded([&] {
        require_ptr(c, "config");
        require_ptr(rep, "report");
        std::memset(rep, 0, sizeof(*rep));
        if (!(c->cadence_s > 0.0) || !(c->run_seconds > 0.0) || c->chunk_spectra == 0 ||
            !(c->arrival_rate_hz >= 0.0) || c->min_beams == 0 || c->max_beams < c->min_beams ||
            !(c->min_duration_s > 0.0) || c->max_duration_s < c->min_duration_s ||
            c->window_a == 0 || c->window_b == 0 || c->pool_spectra < c->chunk_spectra)
            invalid("invalid simulation parameters");
        std::unique_ptr<bf_server> srv;
        create_server(&c->server, &srv);
        bf_server* s = srv.get();
        const size_t F = s->F, A = s->A, row = F * A;
        const size_t chunk = c->chunk_spectra;
        const size_t pool = c->pool_spectra / chunk * chunk;
        uint8_t* d_pool = nullptr;
        hip_check(hipMalloc(reinterpret_cast<void**>(&d_pool), pool * row), "hipMalloc pool");
        std::unique_ptr<uint8_t, void (*)(uint8_t*)> pool_guard(
            d_pool, [](uint8_t* p) { (void)hipFree(p); });
        {
            // One point source at broadside-ish (l, m) = (0.1, -0.05) in noise.
            const float l = 0.1f, m = -0.05f;
            launch_point_source_noise(s->step_stream, d_pool, pool, F, A, s->launch.geo.pos,
                                      s->launch.geo.freqs, l, m, std::sqrt(1.0f - l * l - m * m),
                                      c->source_amplitude, c->noise_sigma, c->seed);
            hip_check(hipStreamSynchronize(s->step_stream), "hipStreamSynchronize");
        }
        std::mt19937_64 rng(c->seed);
        std::exponential_distribution<double> gap(c->arrival_rate_hz > 0 ? c->arrival_rate_hz
                                                                         : 1.0);
        std::uniform_int_distribution<size_t> beams(c->min_beams, c->max_beams);
        std::uniform_real_distribution<double> dur(c->min_duration_s, c->max_duration_s);
        std::uniform_real_distribution<float> pos(-0.5f, 0.5f);
        std::uniform_real_distribution<float> rate(0.0f, 1.0f);
        std::bernoulli_distribution pick_a(0.5);
        const double deg_per_hour = 15.0;
        const float max_rate = (float)(deg_per_hour * M_PI / 180.0 / 3600.0 * c->cadence_s);
        const long long total = (long long)std::ceil(c->run_seconds / c->cadence_s);
        double next_arrival = c->arrival_rate_hz > 0 ? gap(rng) : 1e300;  // seconds
        std::vector<uint64_t> live;
        std::vector<float> traj;
        double active_sum = 0.0;
        size_t active_samples = 0, generated = 0;
        double max_lag_ms = 0.0;
        const auto t_start = std::chrono::steady_clock::now();
        for (long long k = 0; s->head < total; ++k) {
            const long long chunk_end = s->head + (long long)chunk;
            if (c->paced) {
                const auto due = t_start + std::chrono::duration<double>(
                                               (double)chunk_end * c->cadence_s);
                while (std::chrono::steady_clock::now() < due) {
                }
                const double lag =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                              due)
                        .count();
                max_lag_ms = std::max(max_lag_ms, lag);
            }
            // Requests that arrived while this chunk was being received.
            while (next_arrival * 1.0 <= (double)chunk_end * c->cadence_s) {
                bf_server_request q{};
                q.n_beams = beams(rng);
                q.window = pick_a(rng) ? c->window_a : c->window_b;
                q.n_spectra = (size_t)std::llround(dur(rng) / c->cadence_s);
                q.start = -1;
                q.output_windows = c->output_windows_cap;
                traj.resize(4 * q.n_beams);
                for (size_t b = 0; b < q.n_beams; ++b) {
                    traj[4 * b + 0] = pos(rng);
                    traj[4 * b + 1] = pos(rng);
                    traj[4 * b + 2] = max_rate * (2.0f * rate(rng) - 1.0f);
                    traj[4 * b + 3] = max_rate * (2.0f * rate(rng) - 1.0f);
                }
                q.trajectories = traj.data();
                uint64_t id = 0;
                submit_request(s, &q, &id);
                live.push_back(id);
                ++generated;
                next_arrival += gap(rng);
            }
            const size_t at = (size_t)((s->head) % (long long)pool);
            s->ingest(d_pool + at * row, nullptr, chunk, true);
            while (s->processed < s->head) s->step();
            // Release finished requests: their output memory returns to the
            // arena and their slots to the table.
            s->harvest(false);
            size_t in_slots = 0;
            for (size_t i = 0; i < live.size();) {
                auto it = s->requests.find(live[i]);
                if (it == s->requests.end() || it->second.state == kComplete ||
                    it->second.state == kRejected) {
                    if (it != s->requests.end()) s->release(live[i]);
                    live[i] = live.back();
                    live.pop_back();
                } else {
                    in_slots += it->second.state == kActive;
                    ++i;
                }
            }
            active_sum += (double)in_slots;
            ++active_samples;
        }
        hip_check(hipStreamSynchronize(s->ingest_stream), "hipStreamSynchronize");
        hip_check(hipStreamSynchronize(s->step_stream), "hipStreamSynchronize");
        s->harvest(true);
        const double wall =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        s->fill_stats(&rep->stats);
        rep->wall_seconds = wall;
        rep->stream_seconds = (double)s->head * c->cadence_s;
        rep->beam_samples_per_s = rep->stats.beam_samples / wall;
        rep->mean_requests_in_slots = active_samples ? active_sum / (double)active_samples : 0.0;
        rep->requests_generated = generated;
        rep->max_pacing_lag_ms = max_lag_ms;
        rep->graph_used = s->use_graph ? 1 : 0;
    });
}

}  // extern "C"
