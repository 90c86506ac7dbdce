// k_fp32.hip -- plain FP32 fused fixed-grid beamformer for gfx950.
//
// out[w][f][b] = sum_{t in window w} | sum_a W[b][f][a] * X[t][f][a] |^2
// (no conjugate, reference src/cpu_beamformer.cpp:56-78 and the integration
// of src/temporal_integration.cpp:95-109).
//
// Work split: a workgroup owns one frequency f and one block of BB beams;
// its BB x A weights are staged once in LDS as [a][BB] so that every lane of
// a wave reads the same LDS address (broadcast).  Each lane owns one window
// part: it walks its time rows, loading the A packed bytes of a row as
// 16-byte vectors, sign-extending nibbles in registers (v_bfe_i32) and doing
// 4 FMAs per complex MAC for each of its BB beams.  Windows longer than one
// lane's share are split over P adjacent lanes and folded with a butterfly
// of cross-lane adds.
#include "bf_kernels.hpp"

namespace bfhip {
namespace {

constexpr int kThreads = 256;

template <int A, int BB, bool TILED>
__global__ __launch_bounds__(kThreads) void fp32_fused_kernel(
    const uint8_t* __restrict__ x, const float2* __restrict__ w, float* __restrict__ out,
    int T, int F, int B, int S, int W, int P, int chunk) {
    __shared__ float2 ws[A * BB];
    const int f = blockIdx.y;
    const int b0 = blockIdx.z * BB;
    const int tiles = (B + 31) / 32;
    for (int i = threadIdx.x; i < A * BB; i += kThreads) {
        const int a = i / BB;
        const int j = i % BB;
        const int b = b0 + j;
        float2 v = make_float2(0.0f, 0.0f);
        if (b < B) {
            const size_t idx = TILED ? ((size_t)(f * tiles + b / 32) * A + a) * 32 + (b % 32)
                                     : ((size_t)b * F + f) * A + a;
            v = w[idx];
        }
        ws[i] = v;
    }
    __syncthreads();

    const long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    const int win = (int)(gid / P);
    const int part = (int)(gid % P);
    float acc[BB];
#pragma unroll
    for (int j = 0; j < BB; ++j) acc[j] = 0.0f;

    if (win < W) {
        const int wstart = win * S;
        const int t0 = wstart + part * chunk;
        int t1 = t0 + chunk;
        t1 = t1 < wstart + S ? t1 : wstart + S;
        t1 = t1 < T ? t1 : T;
        for (int t = t0; t < t1; ++t) {
            const uint8_t* row = x + ((size_t)t * F + f) * A;
            float sr[BB], si[BB];
#pragma unroll
            for (int j = 0; j < BB; ++j) {
                sr[j] = 0.0f;
                si[j] = 0.0f;
            }
#pragma unroll 1
            for (int a0 = 0; a0 < A; a0 += 16) {
                const uint4 v = *reinterpret_cast<const uint4*>(row + a0);
                const unsigned words[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const unsigned wd = words[q / 4];
                    const unsigned sh = (q % 4) * 8;
                    const float xr = (float)bf_nibble(wd, sh);
                    const float xi = (float)bf_nibble(wd, sh + 4);
                    const float2* wa = ws + (a0 + q) * BB;
#pragma unroll
                    for (int j = 0; j < BB; ++j) {
                        const float2 c = wa[j];
                        sr[j] = fmaf(c.x, xr, sr[j]);
                        sr[j] = fmaf(-c.y, xi, sr[j]);
                        si[j] = fmaf(c.x, xi, si[j]);
                        si[j] = fmaf(c.y, xr, si[j]);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < BB; ++j) {
                const float p = sr[j] * sr[j] + si[j] * si[j];
                acc[j] += p;
            }
        }
    }
    // Fold the P parts of a window (adjacent lanes, P a power of two <= 64).
    for (int off = 1; off < P; off <<= 1) {
#pragma unroll
        for (int j = 0; j < BB; ++j) acc[j] += __shfl_xor(acc[j], off);
    }
    if (win < W && part == 0) {
        float* o = out + ((size_t)win * F + f) * B + b0;
#pragma unroll
        for (int j = 0; j < BB; ++j) {
            if (b0 + j < B) o[j] = acc[j];
        }
    }
}

template <int A, int BB>
void launch_ab(hipStream_t stream, const uint8_t* x, const float* w, bool tiled, float* out,
               int T, int F, int B, int S, int W, int P, int chunk) {
    const long threads = (long)W * P;
    const dim3 grid((unsigned)ceil_div((size_t)threads, kThreads), (unsigned)F,
                    (unsigned)ceil_div((size_t)B, BB));
    const float2* w2 = reinterpret_cast<const float2*>(w);
    if (tiled) {
        hipLaunchKernelGGL((fp32_fused_kernel<A, BB, true>), grid, dim3(kThreads), 0, stream,
                           x, w2, out, T, F, B, S, W, P, chunk);
    } else {
        hipLaunchKernelGGL((fp32_fused_kernel<A, BB, false>), grid, dim3(kThreads), 0, stream,
                           x, w2, out, T, F, B, S, W, P, chunk);
    }
}

template <int A>
void launch_a(hipStream_t stream, const uint8_t* x, const float* w, bool tiled, float* out,
              int T, int F, int B, int S, int W, int P, int chunk) {
    if (B >= 16) {
        launch_ab<A, 16>(stream, x, w, tiled, out, T, F, B, S, W, P, chunk);
    } else if (B >= 4) {
        launch_ab<A, 4>(stream, x, w, tiled, out, T, F, B, S, W, P, chunk);
    } else {
        launch_ab<A, 1>(stream, x, w, tiled, out, T, F, B, S, W, P, chunk);
    }
}

}  // namespace

void launch_fp32_fused(hipStream_t stream, const uint8_t* x, const float* w,
                       WeightLayout layout, float* out, size_t T, size_t F, size_t A,
                       size_t B, size_t S) {
    // S == 0: one "window" per spectrum.
    const size_t s_eff = S == 0 ? 1 : S;
    const size_t W = ceil_div(T, s_eff);
    const size_t bb = B >= 16 ? 16 : (B >= 4 ? 4 : 1);
    const size_t cells = W * F * ceil_div(B, bb);
    // Split long windows over P lanes until ~1M lanes are in flight.
    size_t P = 1;
    while (P < 64 && P * 2 <= s_eff && cells * P < (size_t(1) << 20)) P *= 2;
    const size_t chunk = ceil_div(s_eff, P);
    if (W * P > (size_t)INT32_MAX || T > (size_t)INT32_MAX)
        fail(BF_ERR_OVERFLOW, "FP32 beamformer grid exceeds the supported size");
    const bool tiled = layout == WeightLayout::Tiled;
    const int iT = (int)T, iF = (int)F, iB = (int)B, iS = (int)s_eff, iW = (int)W;
    switch (A) {
    case 32: launch_a<32>(stream, x, w, tiled, out, iT, iF, iB, iS, iW, (int)P, (int)chunk); break;
    case 64: launch_a<64>(stream, x, w, tiled, out, iT, iF, iB, iS, iW, (int)P, (int)chunk); break;
    case 128: launch_a<128>(stream, x, w, tiled, out, iT, iF, iB, iS, iW, (int)P, (int)chunk); break;
    case 256: launch_a<256>(stream, x, w, tiled, out, iT, iF, iB, iS, iW, (int)P, (int)chunk); break;
    default: invalid("n_ant must be 32, 64, 128, or 256");
    }
    hip_check(hipGetLastError(), "fp32_fused_kernel launch");
}

}  // namespace bfhip
