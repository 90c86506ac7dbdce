// k_mfma.hip -- dispatch and autotuning of the MFMA fixed-grid beamformer.
//
// The kernel template lives in mfma_fused.hpp and is instantiated per
// antenna count in k_mfma_a*.hip for a small set of tile shapes (beams per
// wave, waves per workgroup, single or double LDS stage).  A launch takes
// the shape the autotuner measured best for its exact problem, or the
// default below.  The autotuner times every instantiated shape (and three
// chunk-sizing assumptions) on the caller's buffers with device events.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "mfma_fused.hpp"

namespace bfhip {
namespace {

using TuneKey = std::tuple<int, int, int, int, int, bool>;  // T, F, A, B, S, split

std::mutex g_tune_mu;
std::map<TuneKey, MfmaConfig>& tune_cache() {
    static std::map<TuneKey, MfmaConfig> cache;
    return cache;
}

// Untuned shape per antenna count, beam count and precision: the fastest
// measured on MI355X for representative problems (F = 336, S = 10 and 320,
// 1-3 ms launches).  With many beams at A = 64 the split kernel runs two
// N-tiles on each of eight waves (126 registers: two workgroups per CU)
// and the single-FP16 kernel four N-tiles on four; fewer beams use fewer
// N-tiles.  A single LDS stage wins from 64 antennas up, two stages at
// 32 and for the split kernel at 128.
MfmaConfig default_config(int A, int B, bool split) {
    const int need = (B + 7) / 8;  // 8-beam N-tiles
    MfmaConfig c;
    c.waves = 4;
    c.db = 0;
    c.resident = 512;
    if (A == 32) {
        c.nt = need <= 4 ? 1 : 2;
        c.waves = need <= 8 ? 4 : 8;
        c.db = 1;
        c.resident = 256;
    } else if (A == 64) {
        if (need <= 2) {
            c.nt = 1;
        } else if (need <= 8) {
            c.nt = 2;
        } else if (split) {
            c.nt = 2;
            c.waves = 8;
            c.resident = 256;
        } else {
            c.nt = 4;
            c.resident = 256;
        }
    } else if (A == 128) {
        c.nt = need <= 4 ? 1 : 2;
        c.db = split ? 1 : 0;
        c.resident = 256;
    } else {
        c.nt = 1;
        c.resident = 1024;
    }
    return c;
}

bool launch_cfg(hipStream_t stream, const MfmaLaunch& L, int A, const MfmaConfig& cfg) {
    switch (A) {
    case 32: return launch_mfma_a32(stream, L, cfg);
    case 64: return launch_mfma_a64(stream, L, cfg);
    case 128: return launch_mfma_a128(stream, L, cfg);
    case 256: return launch_mfma_a256(stream, L, cfg);
    default: invalid("n_ant must be 32, 64, 128, or 256");
    }
}

std::vector<std::pair<int, int>> shapes_for(int A) {
    if (A <= 64) return {{4, 4}, {2, 8}, {2, 4}, {1, 8}, {1, 4}};
    if (A == 128) return {{2, 4}, {1, 8}, {1, 4}};
    return {{1, 4}};
}

void check_problem(size_t T, size_t F, size_t A, size_t B, size_t S) {
    if (!mfma_supported(A, B, S))
        invalid("the MFMA beamformer supports integration of 10 spectra or a multiple of 4");
    if (T > (size_t)INT32_MAX / 8 || S > (size_t)INT32_MAX / 8 || F > 65535 ||
        T * F > ((size_t)1 << 40))
        fail(BF_ERR_OVERFLOW, "MFMA beamformer grid exceeds the supported size");
}

}  // namespace

bool mfma_supported(size_t A, size_t B, size_t S) {
    (void)B;
    const bool a_ok = A == 32 || A == 64 || A == 128 || A == 256;
    const bool s_ok = S == 0 || S == 10 || (S % 4 == 0);
    return a_ok && s_ok;
}

void launch_mfma_fused(hipStream_t stream, bool split, const uint8_t* x, const float* w,
                       float* out, size_t T, size_t F, size_t A, size_t B, size_t S,
                       const uint8_t* mask) {
    check_problem(T, F, A, B, S);
    MfmaConfig cfg = default_config((int)A, (int)B, split);
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = tune_cache().find(TuneKey((int)T, (int)F, (int)A, (int)B, (int)S, split));
        if (it != tune_cache().end()) cfg = it->second;
    }
    const MfmaLaunch L{x, w, out, mask, (int)T, (int)F, (int)B, (int)S, split};
    if (!launch_cfg(stream, L, (int)A, cfg)) invalid("MFMA tile shape is not instantiated");
    hip_check(hipGetLastError(), "mfma_fused_kernel launch");
}

std::vector<MfmaTuneResult> autotune_mfma_fused(hipStream_t stream, bool split,
                                                const uint8_t* x, const float* w, float* out,
                                                size_t T, size_t F, size_t A, size_t B,
                                                size_t S, int repeats) {
    check_problem(T, F, A, B, S);
    if (repeats < 1) invalid("repeats must be positive");
    const MfmaLaunch L{x, w, out, nullptr, (int)T, (int)F, (int)B, (int)S, split};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    std::vector<MfmaTuneResult> results;
    try {
        // Bring the clocks up before the first candidate is timed.
        const MfmaConfig warm = default_config((int)A, (int)B, split);
        for (int r = 0; r < 3; ++r) launch_cfg(stream, L, (int)A, warm);
        hip_check(hipGetLastError(), "mfma_fused_kernel launch");
        for (const auto& sh : shapes_for((int)A)) {
            for (int db = 0; db < 2; ++db) {
                for (int resident : {256, 512, 1024}) {
                    MfmaConfig cfg;
                    cfg.nt = sh.first;
                    cfg.waves = sh.second;
                    cfg.db = db;
                    cfg.resident = resident;
                    if (!launch_cfg(stream, L, (int)A, cfg)) continue;  // warm-up
                    hip_check(hipGetLastError(), "mfma_fused_kernel launch");
                    std::vector<float> ms;
                    for (int r = 0; r < repeats; ++r) {
                        hip_check(hipEventRecord(e0, stream), "hipEventRecord");
                        launch_cfg(stream, L, (int)A, cfg);
                        hip_check(hipEventRecord(e1, stream), "hipEventRecord");
                        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                        float t = 0.0f;
                        hip_check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
                        ms.push_back(t);
                    }
                    std::sort(ms.begin(), ms.end());
                    results.push_back(MfmaTuneResult{cfg, ms[ms.size() / 2]});
                }
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const auto best = std::min_element(
        results.begin(), results.end(),
        [](const MfmaTuneResult& a, const MfmaTuneResult& b) { return a.ms < b.ms; });
    if (best != results.end()) {
        {
            std::lock_guard<std::mutex> lock(g_tune_mu);
            tune_cache()[TuneKey((int)T, (int)F, (int)A, (int)B, (int)S, split)] = best->cfg;
        }
        // the output ends with the winner's result (every shape computes the same sums)
        launch_cfg(stream, L, (int)A, best->cfg);
        hip_check(hipGetLastError(), "mfma_fused_kernel launch");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    return results;
}

void clear_mfma_tuning() {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    tune_cache().clear();
}

}  // namespace bfhip
