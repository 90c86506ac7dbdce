// mfma_stage.hpp -- staging map shared by the MFMA beamformer and tracker
// kernels: how the packed int4 dwords of a stage of 16-row tiles are dealt
// to the threads of a workgroup and where each lands in LDS.
//
// LDS layout of a tile (A-operand layout of v_mfma_f32_16x16x32_f16): the
// 16-byte slot s*64 + row + 16*g holds the FP16 pairs of packed dword
// j = g*(A/16) + s of that row (antennas g*(A/4) + 4s .. +3), so lane l
// reads its k-step s operand from slot s*64 + l.  Rows are fastest among
// the threads: 16 consecutive lanes write 16 consecutive slots (no bank
// conflicts) while each global load still reads 16 contiguous bytes of
// every row it touches.
#pragma once

namespace bfhip {

constexpr int kStageWaveSize = 64;

// Dword k of a thread is q = tid + k*kThreads of the stage, i.e. tile
// tl0 + tld(k) and in-tile dword rem0 + remd(k), where tl0, rem0 depend on
// the thread only (tl0 on its wave) and tld, remd are compile-time (both
// kThreads and the tile's dword count 4A are powers of two).
template <int A, int kThreads>
struct StageMap {
    static constexpr int DW_TILE = 4 * A;
    static constexpr int TILE_BYTES = (A / 16) * kStageWaveSize * 16;
    static_assert(kThreads % DW_TILE == 0 || DW_TILE % kThreads == 0, "power-of-two staging");
    static constexpr int tld(int k) {
        return kThreads >= DW_TILE ? k * (kThreads / DW_TILE) : (k * kThreads) / DW_TILE;
    }
    static constexpr int remd(int k) {
        return kThreads >= DW_TILE ? 0 : (k * kThreads) % DW_TILE;
    }
    // remd is a multiple of 256, so it moves the lane group g only.
    static constexpr int lds_delta(int k) {
        return tld(k) * TILE_BYTES + 16 * 16 * ((remd(k) >> 4) / (A / 16));
    }
    static constexpr int byte_delta(int k) { return 4 * (remd(k) >> 4); }
    // Per-thread parts for thread tid of wave `wave`.
    __device__ static int tile0(int wave) {
        if (kThreads <= DW_TILE) return 0;  // a tile spans the whole workgroup
        return __builtin_amdgcn_readfirstlane(wave * kStageWaveSize / DW_TILE);
    }
    __device__ static int rem0(int tid) { return kThreads >= DW_TILE ? tid % DW_TILE : tid; }
    // LDS byte offset of dword k = 0 of a thread whose tile is tl0, in-tile dword rem0.
    __device__ static int lds_offset(int tl0, int rem0) {
        const int rho = rem0 & 15;
        const int j0 = rem0 >> 4;
        return tl0 * TILE_BYTES + ((j0 % (A / 16)) * kStageWaveSize + rho + 16 * (j0 / (A / 16))) * 16;
    }
};

}  // namespace bfhip
