// bf_kernels.hpp -- host-side launchers of the HIP kernels (internal).
#pragma once

#include <vector>

#include "bf_common.hpp"

namespace bfhip {

enum class WeightLayout { Canonical, Tiled };

// Plain FP32 fused beamformer (VALU FMA).  S == 0: per-spectrum output
// [T][F][B]; S >= 1: integrated [W][F][B].  Any A in {32,64,128,256},
// 1 <= B <= 128, any S.
void launch_fp32_fused(hipStream_t stream, const uint8_t* x, const float* w,
                       WeightLayout layout, float* out, size_t T, size_t F, size_t A,
                       size_t B, size_t S);

// FP16 MFMA fused beamformer (16x16x32 f16, FP32 accumulate); `split` adds
// the lo half of each weight.  Supported when mfma_supported() is true.
// The tile shape comes from the autotuner's cache for this exact problem,
// else from a default.
bool mfma_supported(size_t A, size_t B, size_t S);
// `mask`: optional [T][F] frame-validity bytes (0 = lost frame, whose
// spectrum then contributes nothing); null when every frame is valid.
void launch_mfma_fused(hipStream_t stream, bool split, const uint8_t* x, const float* w,
                       float* out, size_t T, size_t F, size_t A, size_t B, size_t S,
                       const uint8_t* mask = nullptr);

// Tile shape of one launch of the MFMA beamformer.
struct MfmaConfig {
    int nt = 4;          // 8-beam N-tiles per wave
    int waves = 4;       // waves per workgroup
    int db = 0;          // two LDS stages (one barrier per stage)
    int resident = 512;  // workgroups assumed resident when sizing chunks
};

struct MfmaTuneResult {
    MfmaConfig cfg;
    float ms;  // median launch time
};

// Times every instantiated shape on the caller's buffers (`repeats` timed
// launches each, after one warm-up), caches the fastest for this problem
// and leaves its output in `out`.  Synchronous on `stream`.
std::vector<MfmaTuneResult> autotune_mfma_fused(hipStream_t stream, bool split,
                                                const uint8_t* x, const float* w, float* out,
                                                size_t T, size_t F, size_t A, size_t B,
                                                size_t S, int repeats);
void clear_mfma_tuning();

// int8 quantiser over [Tw][F][B] floats.
void launch_quantize(hipStream_t stream, const float* in, int8_t* codes, float* params,
                     size_t Tw, size_t F, size_t B);

// Seeded synthetic generators.
void launch_uniform_int4(hipStream_t stream, uint8_t* out, size_t count, uint64_t seed);
void launch_point_source_noise(hipStream_t stream, uint8_t* out, size_t T, size_t F,
                               size_t A, const float* pos, const float* freqs,
                               float dx, float dy, float dz, float amplitude, float sigma,
                               uint64_t seed);

}  // namespace bfhip
