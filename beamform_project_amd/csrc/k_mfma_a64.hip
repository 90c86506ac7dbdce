// k_mfma_a64.hip -- instantiations of the MFMA fixed-grid beamformer for
// 64 antennas (one translation unit per antenna count keeps builds parallel).
#include "mfma_fused.hpp"

namespace bfhip {
BF_DEFINE_MFMA_FOR_A(64, BF_MFMA_SHAPES_32_64(64))
}  // namespace bfhip
