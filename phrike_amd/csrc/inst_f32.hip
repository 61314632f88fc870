// Explicit instantiation of the FP32 pipelines (separate translation unit so
// the two precisions compile in parallel).
#include "pipeline.hpp"

namespace ph {
template struct Plan3D<float>;
template struct Plan2D<float>;
}  // namespace ph
