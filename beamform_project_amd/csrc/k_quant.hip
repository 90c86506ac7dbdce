// k_quant.hip -- CHIME-style int8 quantiser (1 x 16 x 16 chunks) and the
// seeded synthetic-input generators.
//
// Quantiser semantics follow reference src/quantization.cpp:20-30 and
// :61-161 exactly: min/max over the chunk's finite values; any non-finite
// value zeroes the chunk's parameters; scale = (max-min)/255,
// offset = min + 127.5*scale, code = clamp(rint((v-offset)/scale), -127,
// 127); non-finite -> -128; scale == 0 -> 0.  Every operation is written
// as a separately rounded IEEE operation (no contraction, correctly rounded
// division) so the codes are bit-identical to the CPU quantiser's.
#include "bf_kernels.hpp"

#include <cmath>

namespace bfhip {
namespace {

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

// One workgroup per (window t, 16-channel tile fi): the 16 x B block of
// floats is contiguous in [t][f][b], so it is read and the codes written
// with fully coalesced accesses through LDS; wave w then quantises chunks
// w, w+4, ... of the block (lane l: channel l/4, beams 4*(l%4) .. +3).
__global__ __launch_bounds__(256) void quantize_kernel(const float* __restrict__ in,
                                                       int8_t* __restrict__ codes,
                                                       float2* __restrict__ params, int F,
                                                       int B, int ft, int bt) {
#pragma clang fp contract(off)
    __shared__ float tile[16 * 128];
    __shared__ int8_t ctile[16 * 128];
    const long blk = blockIdx.x;
    const int fi = (int)(blk % ft);
    const long t = blk / ft;
    const int f0 = fi * 16;
    const int nf = F - f0 < 16 ? F - f0 : 16;
    const size_t base = ((size_t)t * F + f0) * B;
    const int n = nf * B;
    for (int e = threadIdx.x; e < n; e += 256) tile[e] = in[base + e];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int lf = lane / 4;
    for (int bi = wave; bi < bt; bi += 4) {
        const int bb = bi * 16 + (lane % 4) * 4;
        float v[4];
        bool ok[4];
        float mn = INFINITY, mx = -INFINITY;
        bool nonfinite = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ok[k] = lf < nf && bb + k < B;
            v[k] = ok[k] ? tile[lf * B + bb + k] : 0.0f;
            if (ok[k]) {
                if (!isfinite(v[k])) {
                    nonfinite = true;
                } else {
                    mn = fminf(mn, v[k]);
                    mx = fmaxf(mx, v[k]);
                }
            }
        }
        nonfinite = __any(nonfinite);
        mn = wave_min(mn);
        mx = wave_max(mx);
        constexpr float out_min = -127.0f - 0.5f;
        constexpr float out_max = 127.0f + 0.5f;
        constexpr float out_range = out_max - out_min;
        float offset = 0.0f, scale = 0.0f;
        if (!nonfinite && isfinite(mn) && isfinite(mx)) {
            // This translation unit is compiled without contraction.
            const float span = mx - mn;
            scale = span / out_range;
            const float shift = out_min * scale;
            offset = mn - shift;
            if (!isfinite(offset) || !isfinite(scale)) {
                offset = 0.0f;
                scale = 0.0f;
            }
        }
        if (lane == 0) params[(t * ft + fi) * bt + bi] = make_float2(offset, scale);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!ok[k]) continue;
            int code = 0;
            if (!isfinite(v[k])) {
                code = -128;
            } else if (scale != 0.0f) {
                const float d = v[k] - offset;
                const float q = d / scale;
                float r = __builtin_rintf(q);
                r = r < -127.0f ? -127.0f : r;
                r = r > 127.0f ? 127.0f : r;
                code = (int)r;
            }
            ctile[lf * B + bb + k] = (int8_t)code;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += 256) codes[base + e] = ctile[e];
}

__device__ __forceinline__ uint64_t splitmix(uint64_t key, uint64_t seed) {
    uint64_t z = key + (seed + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void uniform_int4_kernel(uint8_t* __restrict__ out, size_t count, uint64_t seed) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        out[i] = (uint8_t)(splitmix(i, seed) & 0xFFu);
}

// Box-Muller normal pair from two hashed uniforms in (0, 1].
__device__ __forceinline__ float2 gauss2(uint64_t key, uint64_t seed) {
    const uint64_t h = splitmix(key, seed);
    const float u1 = ((float)(uint32_t)(h >> 40) + 1.0f) * (1.0f / 16777216.0f);
    const float u2 = (float)(uint32_t)(h & 0xFFFFFFu) * (1.0f / 16777216.0f);
    const float r = sqrtf(-2.0f * logf(u1));
    float s, c;
    sincosf(6.283185307f * u2, &s, &c);
    return make_float2(r * c, r * s);
}

__device__ __forceinline__ int clip4(float v) {
    float r = rintf(v);
    r = r < -8.0f ? -8.0f : (r > 7.0f ? 7.0f : r);
    return (int)r;
}

__global__ void point_source_noise_kernel(uint8_t* __restrict__ out, size_t T, int F, int A,
                                          const float* __restrict__ pos,
                                          const float* __restrict__ freqs, float dx, float dy,
                                          float dz, float amp, float sigma, uint64_t seed) {
    const size_t count = T * (size_t)F * A;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const int a = (int)(i % A);
        const int f = (int)((i / A) % F);
        const double k = kTwoPi * (double)freqs[f] / kSpeedOfLight;
        const double delay =
            (double)pos[3 * a] * dx + (double)pos[3 * a + 1] * dy + (double)pos[3 * a + 2] * dz;
        double s, c;
        sincos(k * delay, &s, &c);
        const float2 n = gauss2(i, seed);
        const int re = clip4(amp * (float)c + sigma * n.x);
        const int im = clip4(-amp * (float)s + sigma * n.y);
        out[i] = (uint8_t)(((unsigned)re & 0xFu) | (((unsigned)im & 0xFu) << 4));
    }
}

}  // namespace

void launch_quantize(hipStream_t stream, const float* in, int8_t* codes, float* params,
                     size_t Tw, size_t F, size_t B) {
    const size_t ft = ceil_div(F, 16), bt = ceil_div(B, 16);
    const size_t n_blocks = Tw * ft;
    if (B > kMaximumBeams) invalid("n_beams must be between 1 and 128");
    if (n_blocks > (size_t)UINT32_MAX)
        fail(BF_ERR_OVERFLOW, "quantization grid exceeds the supported size");
    hipLaunchKernelGGL(quantize_kernel, dim3((unsigned)n_blocks), dim3(256), 0, stream, in,
                       codes, reinterpret_cast<float2*>(params), (int)F, (int)B, (int)ft,
                       (int)bt);
    hip_check(hipGetLastError(), "quantize_kernel launch");
}

void launch_uniform_int4(hipStream_t stream, uint8_t* out, size_t count, uint64_t seed) {
    hipLaunchKernelGGL(uniform_int4_kernel, dim3(4096), dim3(256), 0, stream, out, count, seed);
    hip_check(hipGetLastError(), "uniform_int4_kernel launch");
}

void launch_point_source_noise(hipStream_t stream, uint8_t* out, size_t T, size_t F,
                               size_t A, const float* pos, const float* freqs, float dx,
                               float dy, float dz, float amplitude, float sigma,
                               uint64_t seed) {
    hipLaunchKernelGGL(point_source_noise_kernel, dim3(8192), dim3(256), 0, stream, out, T,
                       (int)F, (int)A, pos, freqs, dx, dy, dz, amplitude, sigma, seed);
    hip_check(hipGetLastError(), "point_source_noise_kernel launch");
}

}  // namespace bfhip
