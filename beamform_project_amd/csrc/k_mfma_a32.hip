// k_mfma_a32.hip -- instantiations of the MFMA fixed-grid beamformer for
// 32 antennas (one translation unit per antenna count keeps builds parallel).
#include "mfma_fused.hpp"

namespace bfhip {
BF_DEFINE_MFMA_FOR_A(32, BF_MFMA_SHAPES_32_64(32))
}  // namespace bfhip
