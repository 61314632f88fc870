// Matrix-core kernels have no CPU emulation; the emulated library reports so.
#include <hip/hip_runtime.h>

#include "../../phrike_amd/csrc/pipeline.hpp"

namespace ph {
void gll_apply(int, int, int, int, const void*, const void*, const void*, const void*, const void*, void*,
               hipStream_t) {
  throw Error(-9, "MFMA kernels are not emulated on the CPU");
}
}  // namespace ph
