// Explicit instantiation of the FP64 pipelines (separate translation unit so
// the two precisions compile in parallel).
#include "pipeline.hpp"

namespace ph {
template struct Plan3D<double>;
template struct Plan2D<double>;
}  // namespace ph
