// CPU stand-in for the matrix-core apply (tests only): the definition
// out[b,i,x] = sum_j D[i,j] F[b,j,x], with the hi/lo split forms evaluated
// in float products and sums as on the device (summation order differs).
#include <hip/hip_runtime.h>

#include "../../include/phrike_hip.h"
#include "../../phrike_amd/csrc/pipeline.hpp"

namespace ph {
void gll_apply(int mode, int ny, int nx, int nbatch, const void* D, const void* Dhi, const void* Dlo,
               const void* F, const void* add, void* out, hipStream_t, const int* active, int group) {
  const bool offset = (mode & PH_GLL_OFFSET) != 0;
  mode &= ~PH_GLL_OFFSET;
  const long long plane = (long long)ny * nx;
  for (int b = 0; b < nbatch; ++b) {
    if (active != nullptr && active[b / group] == 0) continue;
    for (int i = 0; i < ny; ++i)
      for (int x = 0; x < nx; ++x) {
        const long long o = b * plane + (long long)i * nx + x;
        if (mode == PH_GLL_F64) {
          const double* d = static_cast<const double*>(D);
          const double* f = static_cast<const double*>(F) + b * plane;
          double s = 0.0;
          for (int j = 0; j < ny; ++j) s += d[i * ny + j] * f[(long long)j * nx + x];
          static_cast<double*>(out)[o] = add ? -(static_cast<const double*>(add)[o] + s) : s;
        } else if (mode == PH_GLL_SPLIT_F32) {
          const float* dh = static_cast<const float*>(Dhi);
          const float* dl = static_cast<const float*>(Dlo);
          const float* f = static_cast<const float*>(F) + b * plane;
          const float c0 = offset ? f[x] : 0.f;
          float s1 = 0.f, s2 = 0.f;
          for (int j = 0; j < ny; ++j) {
            s1 += dh[i * ny + j] * (f[(long long)j * nx + x] - c0);
            s2 += dl[i * ny + j] * (f[(long long)j * nx + x] - c0);
          }
          const float s = s1 + s2;
          static_cast<float*>(out)[o] = add ? -(static_cast<const float*>(add)[o] + s) : s;
        } else {
          const float* dh = static_cast<const float*>(Dhi);
          const float* dl = static_cast<const float*>(Dlo);
          const double* f = static_cast<const double*>(F) + b * plane;
          const double c0 = offset ? f[x] : 0.0;
          float c1 = 0.f, c2 = 0.f, c3 = 0.f;
          for (int j = 0; j < ny; ++j) {
            const double v = f[(long long)j * nx + x] - c0;
            const float hi = float(v), lo = float(v - double(hi));
            c1 += dh[i * ny + j] * hi;
            c2 += dl[i * ny + j] * hi;
            c3 += dh[i * ny + j] * lo;
          }
          const double s = double(c1) + double(c2) + double(c3);
          static_cast<double*>(out)[o] = add ? -(static_cast<const double*>(add)[o] + s) : s;
        }
      }
  }
}
}  // namespace ph
