// k_weights.hip -- steering weights generated in HBM: w[b][f][a] =
// (float)cos(phase), (float)sin(phase) with phase = 2 pi f_hz (p_a . d_b) / c
// in FP64 (src/weights.cpp:11-122), canonical [B][F][A] or tiled
// [F][ceil(B/32)][A][32] (zero beyond B).  One workgroup per (beam or beam
// slot, channel), one antenna per thread: a write-once HBM-bound kernel,
// 8 bytes per weight.
#include "bf_common.hpp"

namespace bfhip {
namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void weights_kernel(const float* __restrict__ pos,
                                                           const float* __restrict__ freqs,
                                                           const float* __restrict__ dirs,
                                                           float2* __restrict__ out, int B,
                                                           int F, int A, int tiled) {
    const int b = blockIdx.x;  // beam (canonical) or beam slot (tiled)
    const int f = blockIdx.y;
    const double k = kTwoPi * (double)freqs[f] / kSpeedOfLight;
    for (int a = threadIdx.x; a < A; a += kThreads) {
        float2 w = make_float2(0.0f, 0.0f);
        if (b < B) {
            const float* p = pos + 3 * a;
            const float* q = dirs + 3 * b;
            const double delay = (double)p[0] * q[0] + (double)p[1] * q[1] + (double)p[2] * q[2];
            double s, c;
            sincos(k * delay, &s, &c);
            w = make_float2((float)c, (float)s);
        }
        size_t idx;
        if (tiled) {
            const int bt = b / 32, lb = b % 32;
            idx = (((size_t)f * (gridDim.x / 32) + bt) * A + a) * 32 + lb;
        } else {
            idx = ((size_t)b * F + f) * A + a;
        }
        out[idx] = w;
    }
}

}  // namespace

void launch_generate_weights(hipStream_t stream, const bf_dims& d, const float* pos,
                             const float* freqs, const float* dirs, float* out, bool tiled) {
    const size_t slots = tiled ? ceil_div(d.n_beams, 32) * 32 : d.n_beams;
    hipLaunchKernelGGL(weights_kernel, dim3((unsigned)slots, (unsigned)d.n_freq), dim3(kThreads),
                       0, stream, pos, freqs, dirs, reinterpret_cast<float2*>(out),
                       (int)d.n_beams, (int)d.n_freq, (int)d.n_ant, tiled ? 1 : 0);
    hip_check(hipGetLastError(), "weights_kernel launch");
}

}  // namespace bfhip
