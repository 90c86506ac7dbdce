// k_mfma_a256.hip -- instantiations of the MFMA fixed-grid beamformer for
// 256 antennas (one translation unit per antenna count keeps builds parallel).
#include "mfma_fused.hpp"

namespace bfhip {
BF_DEFINE_MFMA_FOR_A(256, BF_MFMA_SHAPES_256)
}  // namespace bfhip
