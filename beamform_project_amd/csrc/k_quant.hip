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

// Min, max and a non-finite flag reduced over the 16 lanes that share
// (lane & 3), i.e. over one chunk's 16 channels.
__device__ __forceinline__ void chunk_reduce(float& mn, float& mx, int& bad) {
#pragma unroll
    for (int off = 4; off <= 32; off <<= 1) {
        mn = fminf(mn, __shfl_xor(mn, off));
        mx = fmaxf(mx, __shfl_xor(mx, off));
        bad |= __shfl_xor(bad, off);
    }
}

// Vector form for B a multiple of 16 (16-byte aligned rows): one wave per
// (window t, 16-channel tile fi, group of four 16-beam chunks).  Lane l
// holds channel f0 + l/4 of chunk 4g + l%4: its 16 beams are one 64-byte
// run, read with four 16-byte loads, and its 16 codes leave in one 16-byte
// store.  The chunk's min/max is the lane's own over its 16 values, then
// a butterfly over the 16 lanes of the chunk; every value is quantised
// with the same separately rounded operations as quantize_kernel.
__global__ __launch_bounds__(256) void quantize_vec16_kernel(const float* __restrict__ in,
                                                             int8_t* __restrict__ codes,
                                                             float2* __restrict__ params,
                                                             int F, int B, int ft, int bt,
                                                             int groups, long n_waves) {
#pragma clang fp contract(off)
    const long wid = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wid >= n_waves) return;  // whole waves only
    const int lane = threadIdx.x & 63;
    const int g = (int)(wid % groups);
    const long rest = wid / groups;
    const int fi = (int)(rest % ft);
    const long t = rest / ft;
    const int f = fi * 16 + (lane >> 2);
    const int bi = g * 4 + (lane & 3);
    const bool ok = f < F && bi < bt;
    const size_t off = ((size_t)t * F + f) * B + (size_t)bi * 16;
    float v[16];
    float mn = INFINITY, mx = -INFINITY;
    int bad = 0;
    if (ok) {
        // Plain (cached) loads: non-temporal ones measured 12% slower here.
        typedef float f32x4 __attribute__((ext_vector_type(4)));
        const f32x4* p = reinterpret_cast<const f32x4*>(in + off);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const f32x4 u = p[q];
            v[4 * q] = u.x;
            v[4 * q + 1] = u.y;
            v[4 * q + 2] = u.z;
            v[4 * q + 3] = u.w;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (!isfinite(v[k])) {
                bad = 1;
            } else {
                mn = fminf(mn, v[k]);
                mx = fmaxf(mx, v[k]);
            }
        }
    }
    chunk_reduce(mn, mx, bad);
    constexpr float out_min = -127.0f - 0.5f;
    constexpr float out_max = 127.0f + 0.5f;
    constexpr float out_range = out_max - out_min;
    float offset = 0.0f, scale = 0.0f;
    if (!bad && isfinite(mn) && isfinite(mx)) {
        const float span = mx - mn;
        scale = span / out_range;
        const float shift = out_min * scale;
        offset = mn - shift;
        if (!isfinite(offset) || !isfinite(scale)) {
            offset = 0.0f;
            scale = 0.0f;
        }
    }
    if (!ok) return;
    if ((lane >> 2) == 0) params[((size_t)t * ft + fi) * bt + bi] = make_float2(offset, scale);
    // The code is rint of the correctly rounded quotient d / scale.  With
    // scale >= 2^-100 the quotient is formed as d * rcp(scale) (relative
    // error below 2^-22.2; |d / scale| <= 128, so within 2.3e-5 of the
    // exact quotient, and the correctly rounded one is within 1.5e-5 of
    // that).  Where this q lies at least 2^-12 from every half-integer, the
    // correctly rounded quotient lies strictly on the same side of each
    // half-integer and is not a tie, so both round to the same integer;
    // elsewhere, and for smaller scales, the correctly rounded division is
    // taken.
    const bool fast = scale >= 0x1p-100f;
    const float rcp = __builtin_amdgcn_rcpf(fast ? scale : 1.0f);
    uint32_t packed[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const float d = v[k] - offset;
        float q = d * rcp;
        const float fr = q - floorf(q);
        if (!fast || fabsf(fr - 0.5f) < 0x1p-12f) q = d / scale;
        const float r = fminf(fmaxf(__builtin_rintf(q), -127.0f), 127.0f);
        int code = scale != 0.0f ? (int)r : 0;
        code = isfinite(v[k]) ? code : -128;
        packed[k >> 2] |= ((uint32_t)code & 0xFFu) << (8 * (k & 3));
    }
    *reinterpret_cast<uint4*>(codes + off) = make_uint4(packed[0], packed[1], packed[2], packed[3]);
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
    const bool vec = B % 16 == 0 && reinterpret_cast<uintptr_t>(in) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(codes) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(params) % 8 == 0;
    if (vec) {
        const size_t groups = ceil_div(bt, 4);
        const size_t n_waves = Tw * ft * groups;
        hipLaunchKernelGGL(quantize_vec16_kernel, dim3((unsigned)ceil_div(n_waves, 4)),
                           dim3(256), 0, stream, in, codes, reinterpret_cast<float2*>(params),
                           (int)F, (int)B, (int)ft, (int)bt, (int)groups, (long)n_waves);
        hip_check(hipGetLastError(), "quantize_vec16_kernel launch");
        return;
    }
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
