// bf_common.hpp -- shared internals of libbeamform_hip.so (gfx950 only).
//
// Host side: the error type every entry point converts into a bf_status,
// HIP error checking, and dimension helpers.  Device side: int4 decoding,
// the MFMA operand vector types and the wave64 lane helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/bf_capi.h"

namespace bfhip {

// ------------------------------------------------------------------ errors
// Internal exception carrying the C-ABI status; converted at the boundary.
struct Error : std::runtime_error {
    bf_status code;
    Error(bf_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(bf_status code, const std::string& msg) {
    throw Error(code, msg);
}
[[noreturn]] inline void invalid(const std::string& msg) {
    fail(BF_ERR_INVALID_ARGUMENT, msg);
}

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        fail(BF_ERR_RUNTIME, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// Records `msg` as this thread's last error and returns `code`.
bf_status set_error(bf_status code, const char* msg);

// Runs `fn`, mapping every exception onto a status and a thread-local
// message.  Used by every extern "C" entry point.
template <typename Fn>
bf_status guarded(Fn&& fn) {
    try {
        fn();
        set_error(BF_OK, "");
        return BF_OK;
    } catch (const Error& e) {
        return set_error(e.code, e.what());
    } catch (const std::invalid_argument& e) {
        return set_error(BF_ERR_INVALID_ARGUMENT, e.what());
    } catch (const std::out_of_range& e) {
        return set_error(BF_ERR_OUT_OF_RANGE, e.what());
    } catch (const std::overflow_error& e) {
        return set_error(BF_ERR_OVERFLOW, e.what());
    } catch (const std::logic_error& e) {
        return set_error(BF_ERR_LOGIC, e.what());
    } catch (const std::exception& e) {
        return set_error(BF_ERR_RUNTIME, e.what());
    } catch (...) {
        return set_error(BF_ERR_RUNTIME, "unknown error");
    }
}

// --------------------------------------------------------------- contract
inline constexpr size_t kLocalFrequencyChannels = 336;   // config.hpp:12
inline constexpr size_t kMaximumBeams = 128;              // config.hpp:18
inline constexpr double kTwoPi = 6.283185307179586476925286766559;
inline constexpr double kSpeedOfLight = 299792458.0;

// Reference include/beamformer/config.hpp:81-94.
inline void validate_dimensions(const bf_dims& d) {
    if (d.n_time == 0) invalid("n_time must be positive");
    if (d.n_freq != kLocalFrequencyChannels)
        invalid("the local-shard contract requires exactly 336 frequency channels");
    if (d.n_ant != 32 && d.n_ant != 64 && d.n_ant != 128 && d.n_ant != 256)
        invalid("n_ant must be 32, 64, 128, or 256");
    if (d.n_beams == 0 || d.n_beams > kMaximumBeams)
        invalid("n_beams must be between 1 and 128");
}

inline constexpr size_t kFrequencyShardCount = 2;  // config.hpp:13

// Reference include/beamformer/config.hpp:43-63 (validate_shard_descriptor).
inline void validate_shard(const bf_shard& s) {
    if (s.shard_count != kFrequencyShardCount) invalid("shard_count must be exactly 2");
    if (s.shard_id >= s.shard_count) invalid("shard_id is outside the two-shard band");
    if (s.local_frequency_count != kLocalFrequencyChannels)
        invalid("local_frequency_count must be exactly 336");
    if (s.absolute_frequency_start > (kFrequencyShardCount - 1) * kLocalFrequencyChannels)
        invalid("absolute frequency start is outside the full band");
    if (s.timestamp_step == 0) invalid("timestamp_step must be positive");
    if (!s.loss_mask_independent) invalid("loss masks must be independent per shard");
}

inline const bf_dims& deref_dims(const bf_dims* d) {
    if (d == nullptr) invalid("dims must not be null");
    return *d;
}

// Reference src/temporal_integration.cpp:16-26.
inline size_t integrated_time_count(size_t n_time, size_t spectra) {
    if (spectra == 0) invalid("integration_spectra must be positive");
    if (n_time == 0) invalid("n_time must be positive");
    return n_time / spectra + (n_time % spectra == 0 ? 0u : 1u);
}

inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }

// Reference src/quantization.cpp:34-47.
inline size_t quantization_parameter_count(const bf_dims& d) {
    return d.n_time * ceil_div(d.n_freq, 16) * ceil_div(d.n_beams, 16);
}

inline size_t tiled_weight_count(const bf_dims& d) {   // formats.hpp:58-61
    return d.n_freq * ceil_div(d.n_beams, 32) * d.n_ant * 32;
}

}  // namespace bfhip

// ------------------------------------------------------------ device side
typedef _Float16 bf_half8 __attribute__((ext_vector_type(8)));
typedef float bf_float4 __attribute__((ext_vector_type(4)));
typedef int bf_int4v __attribute__((ext_vector_type(4)));
typedef unsigned int bf_uint4v __attribute__((ext_vector_type(4)));

// Signed nibble of `v` at bit offset `bit` (two's complement in [-8, 7]);
// one v_bfe_i32.  Reference include/beamformer/int4.hpp:19-30.
__device__ __forceinline__ int bf_nibble(unsigned v, unsigned bit) {
    return __builtin_amdgcn_sbfe((int)v, bit, 4);
}

// Converts the 4 packed bytes of `d` (4 antennas) into 4 fp16x2 words,
// each holding (re, im) of one antenna as exact fp16 values.  The nibble u
// is placed in the mantissa of 1024.0 (0x6400): 0x6400 | (u ^ 8) is the
// fp16 value 1024 + (u ^ 8) = 1032 + x for the signed value x, so one
// packed subtraction of 1032 yields x exactly.
__device__ __forceinline__ void bf_int4x4_to_half2x4(unsigned d, unsigned out[4]) {
    const unsigned lo = d & 0x0F0F0F0Fu;         // real nibbles
    const unsigned hi = (d >> 4) & 0x0F0F0F0Fu;  // imaginary nibbles
    // byte j of lo -> bits 0..7, byte j of hi -> bits 16..23 (v_perm_b32).
    // Selector byte values: 0..3 pick from the second operand (lo), 4..7 from
    // the first (hi), 0x0C yields zero.
    const unsigned sel[4] = {0x0C040C00u, 0x0C050C01u, 0x0C060C02u, 0x0C070C03u};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        unsigned w = __builtin_amdgcn_perm(hi, lo, sel[j]);
        w ^= 0x64086408u;
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        h2 v = __builtin_bit_cast(h2, w);
        const h2 bias = {(_Float16)1032.0f, (_Float16)1032.0f};
        v = v - bias;
        out[j] = __builtin_bit_cast(unsigned, v);
    }
}
