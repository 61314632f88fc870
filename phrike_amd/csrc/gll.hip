// Dense Legendre-Gauss-Lobatto collocation derivative on matrix cores:
//   out[b, i, x] = sum_j D[i, j] * F[b, j, x]     (b over fields x runs)
// i.e. D (ny x ny) applied along the second-to-last axis of a (nb, ny, nx)
// stack, as LegendreLobattoBasis2D.dy / LegendreLobattoBasis1D.dx do with a
// tensordot/matmul (phrike/basis/legendre.py:787-802, 1020-1028).
//
// One wave64 owns a 16 x 16 output tile and walks K = ny in steps of 4 with
// v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32 (split FP32).
// Operand layout (wave64): lane l holds A[l%16][l/16] and B[l/16][l%16] for
// both instructions; accumulator register v of lane l is D[4*(l/16)+v][l%16]
// for the FP32 instruction and D[(l/16)+4*v][l%16] for the FP64 one
// (checked on the device with single-entry matrices).
//
// Split precision: D = Dhi + Dlo with Dhi = fp32(D), Dlo = fp32(D - Dhi)
// (built on the host from the FP64 table).  The operand is G = F - F[row 0]
// per column: the rows of a collocation derivative sum to zero, so D G = D F,
// and FP32 rounding then scales with the variation of F along the line, not
// with its size (a uniform pressure inside a momentum flux, times entries of
// order ny^2, would otherwise dominate the error).  For FP32 data the product
// is Dhi*G + Dlo*G; for FP64 data G = Ghi + Glo as well and the product is
// Dhi*Ghi + Dhi*Glo + Dlo*Ghi, each term accumulated in its own FP32 register
// set and summed in FP64 at the end.  The filter matrix (rows summing to one)
// is applied with offset=0.
//
// Each lane keeps the operands of up to KC k-steps in registers and issues
// their loads together, so one global-load latency is paid per KC steps.
// Large batches (ny, nx multiples of 64) use a block-tiled form that stages
// operands in LDS and reuses each fragment for two MFMAs; both forms walk k in
// the same order, so they give identical results.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/phrike_hip.h"
#include "pipeline.hpp"

namespace ph {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int GLL_WAVES = 4;  // a block: 4 waves stacked along i (64 rows)

constexpr int GLL_KC = 32;  // k-steps (of 4) whose operands are loaded together

// EPI: 0 out = acc; 1 out = -(acc + add)  (hybrid RHS epilogue)
template <int EPI>
__global__ __launch_bounds__(64 * GLL_WAVES) void k_gll_f64(const double* __restrict__ D,
                                                           const double* __restrict__ F,
                                                           const double* __restrict__ add,
                                                           double* __restrict__ out, int ny, int nx,
                                                           const int* __restrict__ active, int group) {
  if (active != nullptr && active[blockIdx.z / group] == 0) return;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int i0 = (blockIdx.y * GLL_WAVES + wave) * 16;
  const int x0 = blockIdx.x * 16;
  const long long bo = (long long)blockIdx.z * ny * nx;
  if (i0 >= ny) return;
  const int r = lane & 15, q = lane >> 4;
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  const double* Drow = D + (long long)(i0 + r) * ny;
  const double* Fc = F + bo + x0 + r;
  const int nk = ny / 4;
  for (int s0 = 0; s0 < nk; s0 += GLL_KC) {
    double a[GLL_KC], b[GLL_KC];
#pragma unroll
    for (int s = 0; s < GLL_KC; ++s) {
      const int k = 4 * (s0 + s) + q;
      a[s] = s0 + s < nk ? Drow[k] : 0.0;
      b[s] = s0 + s < nk ? Fc[(long long)k * nx] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < GLL_KC; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const long long o = bo + (long long)(i0 + q + 4 * v) * nx + x0 + r;
    double val = acc[v];
    if constexpr (EPI == 1) val = -(add[o] + val);
    out[o] = val;
  }
}

// split-precision FP32; FP32 data (T = float) or FP64 data (T = double)
template <typename T, int EPI>
__global__ __launch_bounds__(64 * GLL_WAVES) void k_gll_split(const float* __restrict__ Dhi,
                                                             const float* __restrict__ Dlo,
                                                             const T* __restrict__ F,
                                                             const T* __restrict__ add,
                                                             T* __restrict__ out, int ny, int nx,
                                                             int offset, const int* __restrict__ active,
                                                             int group) {
  if (active != nullptr && active[blockIdx.z / group] == 0) return;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int i0 = (blockIdx.y * GLL_WAVES + wave) * 16;
  const int x0 = blockIdx.x * 16;
  const long long bo = (long long)blockIdx.z * ny * nx;
  if (i0 >= ny) return;
  const int r = lane & 15, q = lane >> 4;
  f4 c1 = {0.f, 0.f, 0.f, 0.f}, c2 = {0.f, 0.f, 0.f, 0.f}, c3 = {0.f, 0.f, 0.f, 0.f};
  const long long drow = (long long)(i0 + r) * ny;
  const T* Fc = F + bo + x0 + r;
  const T c0 = offset ? Fc[0] : T(0);  // this lane's column at row 0
  const int nk = ny / 4;
  for (int s0 = 0; s0 < nk; s0 += GLL_KC) {
    float ah[GLL_KC], al[GLL_KC];
    T bv[GLL_KC];
#pragma unroll
    for (int s = 0; s < GLL_KC; ++s) {
      const int k = 4 * (s0 + s) + q;
      const bool in = s0 + s < nk;
      ah[s] = in ? Dhi[drow + k] : 0.f;
      al[s] = in ? Dlo[drow + k] : 0.f;
      bv[s] = in ? Fc[(long long)k * nx] : T(0);
    }
#pragma unroll
    for (int s = 0; s < GLL_KC; ++s) {
      const T g = bv[s] - c0;
      const float bh = float(g);
      c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[s], bh, c1, 0, 0, 0);
      c2 = __builtin_amdgcn_mfma_f32_16x16x4f32(al[s], bh, c2, 0, 0, 0);
      if constexpr (sizeof(T) == 8) {
        const float bl = float(g - double(bh));
        c3 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[s], bl, c3, 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const long long o = bo + (long long)(i0 + 4 * q + v) * nx + x0 + r;
    double val = double(c1[v]) + (double(c2[v]) + double(c3[v]));
    if constexpr (EPI == 1) val = -(double(add[o]) + val);
    out[o] = T(val);
  }
}

// ---------------------------------------------------------------------------
// Block-tiled form (ny and nx multiples of 64): a block of 4 waves computes a
// 64 x 64 output tile of one batch; each wave owns a 32 x 32 quarter as 2 x 2
// MFMA tiles, so every operand fragment read from LDS feeds two MFMAs.  K runs
// in chunks of 16: the chunk's D rows (64 x 16) and F rows (16 x 64) are
// staged in LDS ([k][i] and [k][x], unit stride across the 16 lanes of a
// fragment), the next chunk's global loads are issued before the MFMAs of the
// current one.  Same operand/accumulator layouts as above.
// ---------------------------------------------------------------------------
constexpr int GT = 64;   // output tile edge
constexpr int GKC = 16;  // k per staged chunk

template <int EPI>
__global__ __launch_bounds__(256) void k_gll_f64_tiled(const double* __restrict__ D, const double* __restrict__ F,
                                                      const double* __restrict__ add, double* __restrict__ out,
                                                      int ny, int nx, const int* __restrict__ active, int group) {
  if (active != nullptr && active[blockIdx.z / group] == 0) return;
  __shared__ double As[GKC][GT];
  __shared__ double Bs[GKC][GT];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = blockIdx.y * GT, x0 = blockIdx.x * GT;
  const long long bo = (long long)blockIdx.z * ny * nx;
  const int r = lane & 15, q = lane >> 4;
  const int wi = (wave >> 1) * 32, wx = (wave & 1) * 32;
  // staging: thread -> (A row ar, k ak..ak+3), (B row bk, x bx..bx+3)
  const int ar = tid >> 2, ak = (tid & 3) * 4;
  const int bk = tid >> 4, bx = (tid & 15) * 4;
  const double* Ag = D + (long long)(i0 + ar) * ny + ak;
  const double* Bg = F + bo + (long long)bk * nx + x0 + bx;
  double ra[4], rb[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    ra[u] = Ag[u];
    rb[u] = Bg[u];
  }
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = d4{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < ny; k0 += GKC) {
    lds_barrier();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      As[ak + u][ar] = ra[u];
      Bs[bk][bx + u] = rb[u];
    }
    lds_barrier();
    if (k0 + GKC < ny) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        ra[u] = Ag[k0 + GKC + u];
        rb[u] = Bg[(long long)(k0 + GKC) * nx + u];
      }
    }
#pragma unroll
    for (int s = 0; s < GKC / 4; ++s) {
      const int kk = 4 * s + q;
      const double a0 = As[kk][wi + r], a1 = As[kk][wi + 16 + r];
      const double b0 = Bs[kk][wx + r], b1 = Bs[kk][wx + 16 + r];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long long o = bo + (long long)(i0 + wi + 16 * a + q + 4 * v) * nx + x0 + wx + 16 * c + r;
        double val = acc[a][c][v];
        if constexpr (EPI == 1) val = -(add[o] + val);
        out[o] = val;
      }
}

template <typename T, int EPI>
__global__ __launch_bounds__(256) void k_gll_split_tiled(const float* __restrict__ Dhi, const float* __restrict__ Dlo,
                                                        const T* __restrict__ F, const T* __restrict__ add,
                                                        T* __restrict__ out, int ny, int nx, int offset,
                                                        const int* __restrict__ active, int group) {
  if (active != nullptr && active[blockIdx.z / group] == 0) return;
  constexpr bool LO = sizeof(T) == 8;  // FP64 data: G = Ghi + Glo
  __shared__ float Ah[GKC][GT], Al[GKC][GT], Bh[GKC][GT];
  __shared__ float Bl[LO ? GKC : 1][LO ? GT : 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = blockIdx.y * GT, x0 = blockIdx.x * GT;
  const long long bo = (long long)blockIdx.z * ny * nx;
  const int r = lane & 15, q = lane >> 4;
  const int wi = (wave >> 1) * 32, wx = (wave & 1) * 32;
  const int ar = tid >> 2, ak = (tid & 3) * 4;
  const int bk = tid >> 4, bx = (tid & 15) * 4;
  const float* Ahg = Dhi + (long long)(i0 + ar) * ny + ak;
  const float* Alg = Dlo + (long long)(i0 + ar) * ny + ak;
  const T* Bg = F + bo + (long long)bk * nx + x0 + bx;
  T c0[4];  // this thread's columns at row 0 (the offset of D's zero row sums)
#pragma unroll
  for (int u = 0; u < 4; ++u) c0[u] = offset ? F[bo + x0 + bx + u] : T(0);
  float rah[4], ral[4];
  T rb[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    rah[u] = Ahg[u];
    ral[u] = Alg[u];
    rb[u] = Bg[u];
  }
  f4 c1[2][2], c2[2][2], c3[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      c1[a][c] = f4{0.f, 0.f, 0.f, 0.f};
      c2[a][c] = c1[a][c];
      c3[a][c] = c1[a][c];
    }
  for (int k0 = 0; k0 < ny; k0 += GKC) {
    lds_barrier();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      Ah[ak + u][ar] = rah[u];
      Al[ak + u][ar] = ral[u];
      const T g = rb[u] - c0[u];
      const float gh = float(g);
      Bh[bk][bx + u] = gh;
      if constexpr (LO) Bl[bk][bx + u] = float(g - double(gh));
    }
    lds_barrier();
    if (k0 + GKC < ny) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        rah[u] = Ahg[k0 + GKC + u];
        ral[u] = Alg[k0 + GKC + u];
        rb[u] = Bg[(long long)(k0 + GKC) * nx + u];
      }
    }
#pragma unroll
    for (int s = 0; s < GKC / 4; ++s) {
      const int kk = 4 * s + q;
      const float ah0 = Ah[kk][wi + r], ah1 = Ah[kk][wi + 16 + r];
      const float al0 = Al[kk][wi + r], al1 = Al[kk][wi + 16 + r];
      const float bh0 = Bh[kk][wx + r], bh1 = Bh[kk][wx + 16 + r];
      c1[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah0, bh0, c1[0][0], 0, 0, 0);
      c1[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah0, bh1, c1[0][1], 0, 0, 0);
      c1[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah1, bh0, c1[1][0], 0, 0, 0);
      c1[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah1, bh1, c1[1][1], 0, 0, 0);
      c2[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(al0, bh0, c2[0][0], 0, 0, 0);
      c2[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(al0, bh1, c2[0][1], 0, 0, 0);
      c2[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(al1, bh0, c2[1][0], 0, 0, 0);
      c2[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(al1, bh1, c2[1][1], 0, 0, 0);
      if constexpr (LO) {
        const float bl0 = Bl[kk][wx + r], bl1 = Bl[kk][wx + 16 + r];
        c3[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah0, bl0, c3[0][0], 0, 0, 0);
        c3[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah0, bl1, c3[0][1], 0, 0, 0);
        c3[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah1, bl0, c3[1][0], 0, 0, 0);
        c3[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah1, bl1, c3[1][1], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long long o = bo + (long long)(i0 + wi + 16 * a + 4 * q + v) * nx + x0 + wx + 16 * c + r;
        double val = double(c1[a][c][v]) + (double(c2[a][c][v]) + double(c3[a][c][v]));
        if constexpr (EPI == 1) val = -(double(add[o]) + val);
        out[o] = T(val);
      }
}

}  // namespace ph

namespace ph {

void gll_apply(int mode, int ny, int nx, int nbatch, const void* D, const void* Dhi, const void* Dlo,
               const void* F, const void* add, void* out, hipStream_t st, const int* active, int group) {
  if (active != nullptr && (group < 1 || nbatch % group != 0))
    throw Error(PH_ERR_ARG, "active groups must divide the batch");
  if (ny < 16 || nx < 16 || ny % 16 || nx % 16 || nbatch < 1 || nbatch > 65535)
    throw Error(PH_ERR_ARG, "ny, nx must be positive multiples of 16; nbatch in [1, 65535]");
  if (!F || !out) throw Error(PH_ERR_ARG, "null buffer");
  const int offset = (mode & PH_GLL_OFFSET) ? 1 : 0;
  mode &= ~PH_GLL_OFFSET;
  const bool epi = add != nullptr;
  // the tiled form when it yields at least one block per CU; smaller batches
  // keep the 16-column blocks of the per-wave form, which spread wider
  if (ny % GT == 0 && nx % GT == 0 && (long long)(nx / GT) * (ny / GT) * nbatch >= 256) {
    dim3 grid(nx / GT, ny / GT, nbatch), block(256);
    switch (mode) {
      case PH_GLL_F64:
        if (!D) throw Error(PH_ERR_ARG, "null D");
        if (epi) k_gll_f64_tiled<1><<<grid, block, 0, st>>>((const double*)D, (const double*)F, (const double*)add,
                                                            (double*)out, ny, nx, active, group);
        else k_gll_f64_tiled<0><<<grid, block, 0, st>>>((const double*)D, (const double*)F, nullptr, (double*)out,
                                                        ny, nx, active, group);
        break;
      case PH_GLL_SPLIT_F32:
        if (!Dhi || !Dlo) throw Error(PH_ERR_ARG, "null Dhi/Dlo");
        if (epi) k_gll_split_tiled<float, 1><<<grid, block, 0, st>>>((const float*)Dhi, (const float*)Dlo,
                                                                     (const float*)F, (const float*)add,
                                                                     (float*)out, ny, nx, offset, active, group);
        else k_gll_split_tiled<float, 0><<<grid, block, 0, st>>>((const float*)Dhi, (const float*)Dlo,
                                                                 (const float*)F, nullptr, (float*)out, ny, nx,
                                                                 offset, active, group);
        break;
      case PH_GLL_SPLIT_F64IO:
        if (!Dhi || !Dlo) throw Error(PH_ERR_ARG, "null Dhi/Dlo");
        if (epi) k_gll_split_tiled<double, 1><<<grid, block, 0, st>>>((const float*)Dhi, (const float*)Dlo,
                                                                      (const double*)F, (const double*)add,
                                                                      (double*)out, ny, nx, offset, active, group);
        else k_gll_split_tiled<double, 0><<<grid, block, 0, st>>>((const float*)Dhi, (const float*)Dlo,
                                                                  (const double*)F, nullptr, (double*)out, ny,
                                                                  nx, offset, active, group);
        break;
      default:
        throw Error(PH_ERR_ARG, "unknown mode");
    }
    check_launch("k_gll(tiled)");
    return;
  }
  dim3 block(64 * GLL_WAVES);
  dim3 grid(nx / 16, (ny + 16 * GLL_WAVES - 1) / (16 * GLL_WAVES), nbatch);
  switch (mode) {
    case PH_GLL_F64:
      if (!D) throw Error(PH_ERR_ARG, "null D");
      if (epi)
        k_gll_f64<1><<<grid, block, 0, st>>>((const double*)D, (const double*)F, (const double*)add,
                                             (double*)out, ny, nx, active, group);
      else
        k_gll_f64<0><<<grid, block, 0, st>>>((const double*)D, (const double*)F, nullptr, (double*)out,
                                             ny, nx, active, group);
      break;
    case PH_GLL_SPLIT_F32:
      if (!Dhi || !Dlo) throw Error(PH_ERR_ARG, "null Dhi/Dlo");
      if (epi)
        k_gll_split<float, 1><<<grid, block, 0, st>>>((const float*)Dhi, (const float*)Dlo,
                                                      (const float*)F, (const float*)add, (float*)out,
                                                      ny, nx, offset, active, group);
      else
        k_gll_split<float, 0><<<grid, block, 0, st>>>((const float*)Dhi, (const float*)Dlo,
                                                      (const float*)F, nullptr, (float*)out, ny, nx,
                                                      offset, active, group);
      break;
    case PH_GLL_SPLIT_F64IO:
      if (!Dhi || !Dlo) throw Error(PH_ERR_ARG, "null Dhi/Dlo");
      if (epi)
        k_gll_split<double, 1><<<grid, block, 0, st>>>((const float*)Dhi, (const float*)Dlo,
                                                       (const double*)F, (const double*)add,
                                                       (double*)out, ny, nx, offset, active, group);
      else
        k_gll_split<double, 0><<<grid, block, 0, st>>>((const float*)Dhi, (const float*)Dlo,
                                                       (const double*)F, nullptr, (double*)out, ny, nx,
                                                       offset, active, group);
      break;
    default:
      throw Error(PH_ERR_ARG, "unknown mode");
  }
  check_launch("k_gll");
}

}  // namespace ph
