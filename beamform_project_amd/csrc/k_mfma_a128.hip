// k_mfma_a128.hip -- instantiations of the MFMA fixed-grid beamformer for
// 128 antennas (one translation unit per antenna count keeps builds parallel).
#include "mfma_fused.hpp"

namespace bfhip {
BF_DEFINE_MFMA_FOR_A(128, BF_MFMA_SHAPES_128)
}  // namespace bfhip
