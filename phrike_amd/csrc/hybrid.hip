// 2D Kelvin-Helmholtz on a hybrid grid: Fourier along x (periodic), Legendre-
// Gauss-Lobatto collocation along y.  The reference's own hybrid Grid2D is not
// usable for this (Gauss nodes, an ad-hoc D, FFTs across the non-periodic axis
// and a transposed dy1: phrike/grid.py:738-772, 852-891, 1231-1256), so the
// composition is defined from its working pieces:
//
//   rhs = -(dx Fx + D Fy)
//     dx  = Re IFFT_x( i kx * mask_x * sigma_x * FFT_x(.) )   (Grid1D.dx1,
//           phrike/grid.py:512-519, per row)
//     D   = LegendreLobattoBasis1D.D_x along y (phrike/basis/legendre.py:
//           714-717, 1020-1028), on matrix cores (gll.hip)
//   RK4 with the positivity clamp after every stage and at the end
//   (phrike/solver.py:467-513); post-step filter: x mask*sigma, then the LGL
//   nodal filter along y (phrike/basis/legendre.py:817-872); CFL
//   cfl*min(dx, dy_min)/max(|ux|+a) (phrike/solver.py:613-620).
//   y walls (bc = 1): the 1D Legendre grid's "wall" rule
//   (phrike/grid.py:621-738) along y: normal momentum my = 0 on the two
//   boundary rows, rho and E copied from the adjacent row, mx left as is.  As
//   in the reference's 1D RHS (phrike/solver.py:64-68, where the state is
//   modified in place), it applies to every RHS input, so the step's base
//   state U is taken with walls applied, and again after the filter
//   (phrike/solver.py:120-130).  bc = 0 is the reference hybrid grid's no-op
//   (phrike/grid.py:1082-1084); collocation without it is unstable for KHI.
//
// Per stage: k_hx (fluxes, x derivative of Fx with both rows of a pair in one
// complex transform; Fy written as is) -> D.Fy with the epilogue
// rhs = -(dFx + D.Fy).  The RK4 combine + clamp of stage s runs inside the
// k_hx of stage s+1 (and the stage-4 combine inside the filter's k_hx), so
// the stage states never go to HBM; k_hrk does it as a separate pass when the
// step has no filter.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/phrike_hip.h"
#include "pipeline.hpp"

namespace ph {


enum HxMode { HX_RHS = 0, HX_FILTER = 1 };

template <typename T>
struct HxArgs {
  const T* in;      // (runs, 4, ny, nx)
  T* dfx;           // RHS: d/dx of Fx; FILTER: filtered output
  T* fy;            // RHS: Fy
  long long fs, ss; // field / run strides
  int ny;
  const cx<T>* tw;
  const T* fac;     // by natural x index: mask*sigma/nx
  const T* kx;      // by natural x index: kx (0 at the Nyquist index)
  T gamma;
  const int* active;
  int bc;
  // combine mode: the input is the RK4 combine of stage `stage` from its
  // derivative rhs, the step's base state un and the accumulator acc
  const T* rhs;
  const T* un;
  T* acc;
  int stage;
  double dt;
  const double* dt_dev;  // per run, or null
};

// wall rule on a boundary row: my = 0, rho and E from the adjacent row
// (nb points at the same x on that row)
template <typename T>
__device__ __forceinline__ void wall_row(T (&q)[4], const T* nb, long long fs) {
  q[0] = nb[0];
  q[2] = T(0);
  q[3] = nb[3 * fs];
}

// RK4 combine of stage a.stage at point o (row y) with k = rhs, clamp
// (phrike/solver.py:467-513): stages 1-3 give the next stage state and, with
// write_acc, update the accumulator (k1, then + 2 k2, + 2 k3); stage 4 gives
// U + dt/6 (acc + k4).  The base state takes the wall rule on boundary rows.
template <typename T>
__device__ __forceinline__ void hyb_combine(const HxArgs<T>& a, long long o, int y, int nx, double d,
                                            bool write_acc, T (&u)[4]) {
  const T half = T(0.5 * d), full = T(d), sixth = T(d / 6.0);
  T b[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) b[c] = a.un[c * a.fs + o];
  if (a.bc && (y == 0 || y == a.ny - 1)) wall_row(b, a.un + o + (y == 0 ? nx : -nx), a.fs);
  if (a.stage < 4) {
    const T coef = a.stage == 3 ? full : half;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const T k = a.rhs[c * a.fs + o];
      if (write_acc) a.acc[c * a.fs + o] = a.stage == 1 ? k : a.acc[c * a.fs + o] + T(2) * k;
      u[c] = b[c] + coef * k;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c) u[c] = b[c] + sixth * (a.acc[c * a.fs + o] + a.rhs[c * a.fs + o]);
  }
  clamp2d(u, a.gamma, T(1e-8), T(1e-10));
}

template <typename T, int NX, int MODE, bool COMB>
__global__ __launch_bounds__((XThreads<T, NX>::value)) void k_hx(HxArgs<T> a) {
  constexpr int NT = XThreads<T, NX>::value;
  constexpr int PPT = (2 * NX + NT - 1) / NT;
  PH_XTW_DECL(T, NX);
  __shared__ cx<T> buf[xbuf_size<NX, 4, T>()];
  T* bufr = reinterpret_cast<T*>(buf);
  const int tid = threadIdx.x;
  const int run = blockIdx.y;
  if (a.active != nullptr && a.active[run] == 0) return;
  const int y0 = 2 * blockIdx.x;
  PH_XTW_FILL(a.tw, NX, NT, tid);
  const long long base = (long long)run * a.ss + (long long)y0 * NX;
  const double d = COMB ? (a.dt_dev != nullptr ? a.dt_dev[run] : a.dt) : 0.0;
#pragma unroll
  for (int pi = 0; pi < PPT; ++pi) {
    const int p = tid + pi * NT;
    if (p >= 2 * NX) continue;
    const int line = p / NX, x = p - line * NX;
    const int y = y0 + line;
    T q[4];
    if constexpr (COMB) {
      hyb_combine(a, base + p, y, NX, d, true, q);
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) q[c] = a.in[c * a.fs + base + p];
    }
    if constexpr (MODE == HX_RHS) {
      if (a.bc && (y == 0 || y == a.ny - 1)) {
        const int off = y == 0 ? NX : -NX;
        if constexpr (COMB) {
          // the adjacent (interior) row's stage state, recomputed: stages
          // 1-3 do not read the accumulator, so no other thread's write races
          T nb[4];
          hyb_combine(a, base + p + off, y + (y == 0 ? 1 : -1), NX, d, false, nb);
          q[0] = nb[0];
          q[2] = T(0);
          q[3] = nb[3];
        } else {
          wall_row(q, a.in + base + p + off, a.fs);
        }
      }
      // phrike/equations.py:104-134
      const T rho = q[0];
      const T ux = q[1] / rho, uy = q[2] / rho;
      const T kin = (T(0.5) * rho) * (ux * ux + uy * uy);
      const T pr = (a.gamma - T(1)) * (q[3] - kin);
      const T mx = rho * ux, my = rho * uy, ep = q[3] + pr;
      const T fx[4] = {mx, mx * ux + pr, mx * uy, ep * ux};
      const T fy[4] = {my, my * ux, my * uy + pr, ep * uy};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        bufr[2 * xslot<4, T>(x, c) + line] = fx[c];
        a.fy[c * a.fs + base + p] = fy[c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) bufr[2 * xslot<4, T>(x, c) + line] = q[c];
    }
  }
  lds_barrier();
  fft_lds<T, NX, 4, -1, NT, xstride(4, sizeof(cx<T>)), xpadb(4, sizeof(cx<T>)), XPlan<NX>, 1, PH_XTW_SWZ>(buf, tw, tid);
  // i k (or 1) * mask * sigma / nx; both rows of the pair share the multiplier,
  // which keeps the spectrum Hermitian per row (k(-m) = -k(m), k = 0 at Nyquist)
  for (int i = tid; i < 4 * NX; i += NT) {
    const int n = i >> 2, c = i & 3;
    cx<T> z = buf[xslot<4, T>(n, c)];
    z = cscale(z, a.fac[n]);
    if constexpr (MODE == HX_RHS) z = mul_ik(z, a.kx[n]);
    buf[xslot<4, T>(n, c)] = z;
  }
  lds_barrier();
  fft_lds<T, NX, 4, +1, NT, xstride(4, sizeof(cx<T>)), xpadb(4, sizeof(cx<T>)), XPlan<NX>, 1, PH_XTW_SWZ>(buf, tw, tid);
#pragma unroll
  for (int pi = 0; pi < PPT; ++pi) {
    const int p = tid + pi * NT;
    if (p >= 2 * NX) continue;
    const int line = p / NX, x = p - line * NX;
#pragma unroll
    for (int c = 0; c < 4; ++c) a.dfx[c * a.fs + base + p] = bufr[2 * xslot<4, T>(x, c) + line];
  }
}

// RK4 stage combine with clamp, pointwise over (runs, ny, nx); k_s = rhs
template <typename T>
__global__ __launch_bounds__(NT) void k_hrk(const T* rhs, const T* un, T* acc_out, T* us, long long ncell,
                                            long long fs, long long ss, int stage, double dt,
                                            const double* dt_dev, const int* active, T gamma, int nx,
                                            int ny, int bc) {
  const int run = blockIdx.y;
  if (active != nullptr && active[run] == 0) return;
  const long long i = (long long)blockIdx.x * NT + threadIdx.x;
  if (i >= ncell) return;
  const double d = dt_dev != nullptr ? dt_dev[run] : dt;
  const long long o = (long long)run * ss + i;
  const T half = T(0.5 * d), full = T(d), sixth = T(d / 6.0);
  // the step's base state, with the walls applied
  T b[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) b[c] = un[c * fs + o];
  const int y = int(i / nx);
  if (bc && (y == 0 || y == ny - 1)) wall_row(b, un + o + (y == 0 ? nx : -nx), fs);
  T u[4];
  if (stage < 4) {
    const T coef = stage == 3 ? full : half;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const T k = rhs[c * fs + o];
      if (stage == 1) acc_out[c * fs + o] = k;
      else acc_out[c * fs + o] = acc_out[c * fs + o] + T(2) * k;
      u[c] = b[c] + coef * k;
    }
    clamp2d(u, gamma, T(1e-8), T(1e-10));
#pragma unroll
    for (int c = 0; c < 4; ++c) us[c * fs + o] = u[c];
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c) u[c] = b[c] + sixth * (acc_out[c * fs + o] + rhs[c * fs + o]);
    clamp2d(u, gamma, T(1e-8), T(1e-10));
#pragma unroll
    for (int c = 0; c < 4; ++c) acc_out[c * fs + o] = u[c];
  }
}

// walls on the two boundary rows of each run, in place (after the filter)
template <typename T>
__global__ __launch_bounds__(NT) void k_hbc(T* U, long long fs, long long ss, int nx, int ny, const int* active) {
  const int run = blockIdx.y;
  if (active != nullptr && active[run] == 0) return;
  const int x = blockIdx.x * NT + threadIdx.x;
  if (x >= nx) return;
  T* u = U + (long long)run * ss + x;
  const long long top = (long long)(ny - 1) * nx;
  u[0] = u[nx];
  u[2 * fs] = T(0);
  u[3 * fs] = u[3 * fs + nx];
  u[top] = u[top - nx];
  u[2 * fs + top] = T(0);
  u[3 * fs + top] = u[3 * fs + top - nx];
}

template <typename T>
struct PlanHyb : PlanBase {
  int nx, ny, batch, gll_mode, bc;
  double lx, dx, dy_min, gamma;
  AxisTables ax;
  cx<T>* tw = nullptr;
  T* fac = nullptr;
  T* kx = nullptr;
  double* D = nullptr;   // FP64 tables (row-major ny x ny)
  float* Dhi = nullptr;
  float* Dlo = nullptr;
  double* Fn = nullptr;
  float* Fnhi = nullptr;
  float* Fnlo = nullptr;
  bool filter;
  long long ncell;
  size_t wsBytes;
  static constexpr int NPART = 256;

  PlanHyb(int nx_, int ny_, double lx_, double dy_min_, bool dealias, bool filt, int p, double alpha,
          double gamma_, int batch_, const double* Dh, const double* Fnh, int mode, int ybc)
      : nx(nx_), ny(ny_), batch(batch_), gll_mode(mode), bc(ybc), lx(lx_), dy_min(dy_min_), gamma(gamma_),
        filter(filt) {
    if (ybc != 0 && ybc != 1) throw Error(-2, "y_bc must be 0 (none) or 1 (wall)");
    dtype = sizeof(T) == 8 ? 0 : 1;
    if (!supported_len(nx) || nx < 16) throw Error(-2, "nx must be a power of two in [16, 1024]");
    if (ny < 16 || ny % 16) throw Error(-2, "ny must be a positive multiple of 16");
    if (batch < 1 || batch > 65535) throw Error(-2, "batch must be in [1, 65535]");
    if (sizeof(T) == 4 && mode == PH_GLL_F64) throw Error(-2, "FP32 data needs the split FP32 MFMA mode");
    if (sizeof(T) == 8 && mode == PH_GLL_SPLIT_F32) throw Error(-2, "FP64 data: use F64 or SPLIT_F64IO");
    dx = lx / nx;
    ncell = (long long)nx * ny;
    // full-length x tables by natural index (Grid1D conventions)
    ax = AxisTables(nx, lx, dealias, filt, p, alpha, false);
    std::vector<cx<T>> twh(nx);
    std::vector<T> fh(nx, T(0)), kh(nx, T(0));
    for (int m = 0; m < nx; ++m) {
      twh[m].x = T(ax.twr[m]);
      twh[m].y = T(ax.twi[m]);
    }
    for (int j = 0; j < ax.nr; ++j) {
      const int n = j <= ax.kp ? j : j + (nx - ax.nr);
      fh[n] = T(ax.sig[j] / nx);
      kh[n] = T(ax.k[j]);
    }
    up(&tw, twh.data(), nx);
    up(&fac, fh.data(), nx);
    up(&kx, kh.data(), nx);
    std::vector<float> hi(ny * ny), lo(ny * ny);
    auto split = [&](const double* m, double** d64, float** dh, float** dl) {
      for (int i = 0; i < ny * ny; ++i) {
        hi[i] = float(m[i]);
        lo[i] = float(m[i] - double(hi[i]));
      }
      up(d64, m, ny * ny);
      up(dh, hi.data(), ny * ny);
      up(dl, lo.data(), ny * ny);
    };
    split(Dh, &D, &Dhi, &Dlo);
    split(Fnh, &Fn, &Fnhi, &Fnlo);
    const size_t fld = 4ull * batch * ncell * sizeof(T);
    auto al = [](size_t v) { return (v + 255) / 256 * 256; };
    wsBytes = 4 * al(fld) + al(sizeof(double) * NPART * batch);
  }
  template <typename V>
  static void up(V** d, const V* h, size_t n) {
    hip_check(hipMalloc(d, sizeof(V) * n), "hipMalloc(hybrid table)");
    hip_check(hipMemcpy(*d, h, sizeof(V) * n, hipMemcpyHostToDevice), "copy hybrid table");
  }
  ~PlanHyb() override {
    for (void* p : {(void*)tw, (void*)fac, (void*)kx, (void*)D, (void*)Dhi, (void*)Dlo, (void*)Fn,
                    (void*)Fnhi, (void*)Fnlo})
      if (p) (void)hipFree(p);
  }
  size_t fld() const { return (4ull * batch * ncell * sizeof(T) + 255) / 256 * 256; }
  T* W(void* ws, int k) const { return reinterpret_cast<T*>((char*)ws + k * fld()); }
  double* Part(void* ws) const { return reinterpret_cast<double*>((char*)ws + 4 * fld()); }

  HxArgs<T> hx_args(const T* in, T* dfx, T* fy, const int* active) const {
    HxArgs<T> a{};
    a.in = in;
    a.dfx = dfx;
    a.fy = fy;
    a.fs = ncell;
    a.ss = 4 * ncell;
    a.ny = ny;
    a.tw = tw;
    a.fac = fac;
    a.kx = kx;
    a.gamma = T(gamma);
    a.active = active;
    a.bc = bc;
    return a;
  }
  template <int MODE, bool COMB = false>
  void hx(int nb, const HxArgs<T>& a, hipStream_t st) const {
    dim3 grid(ny / 2, nb);
    KScope ks(MODE == HX_RHS ? (COMB ? "h_xderiv_rk" : "h_xderiv") : (COMB ? "h_xfilter_rk" : "h_xfilter"), st);
    PH_SWITCH_N(nx, hipLaunchKernelGGL((k_hx<T, N_, MODE, COMB>), grid, dim3(XThreads<T, N_>::value), 0, st, a));
    check_launch("k_hx");
  }
  // the combine of stage s (from its derivative r) as the input of hx
  HxArgs<T> hx_comb_args(int s, const T* r, const T* U, T* acc, double dt, const double* dt_dev, T* dfx,
                         T* fy, const int* active) const {
    HxArgs<T> a = hx_args(nullptr, dfx, fy, active);
    a.rhs = r;
    a.un = U;
    a.acc = acc;
    a.stage = s;
    a.dt = dt;
    a.dt_dev = dt_dev;
    return a;
  }
  // runs with active[run] == 0 are skipped: their out is left untouched
  void dy_apply(int nb, bool filt_matrix, const T* F, const T* add, T* out, const int* active,
                hipStream_t st) const {
    // the derivative's rows sum to zero: split modes may apply it to F - F[row 0]
    const int mode = (gll_mode != PH_GLL_F64 && !filt_matrix) ? (gll_mode | PH_GLL_OFFSET) : gll_mode;
    KScope ks(filt_matrix ? "h_gll_filter" : "h_gll_d", st);
    gll_apply(mode, ny, nx, 4 * nb, filt_matrix ? (const void*)Fn : (const void*)D,
              filt_matrix ? (const void*)Fnhi : (const void*)Dhi, filt_matrix ? (const void*)Fnlo : (const void*)Dlo,
              F, add, out, st, active, 4);
  }
  // rhs = -(dx Fx + D Fy)
  void rhs(int nb, const T* U, T* out, const int* active, void* ws, hipStream_t st) const {
    T* dfx = W(ws, 0);
    T* fy = W(ws, 1);
    hx<HX_RHS>(nb, hx_args(U, dfx, fy, active), st);
    dy_apply(nb, false, fy, dfx, out, active, st);
  }
  void step(int nb, const T* U, T* out, double dt, const double* dt_dev, const int* active, bool filt,
            void* ws, hipStream_t st) const {
    T* r = W(ws, 2);
    T* us = W(ws, 3);
    T* dfx = W(ws, 0);
    T* fy = W(ws, 1);
    // stage 1 from U; stages 2-4 from the combine of the previous stage,
    // formed inside their x-derivative kernel (accumulator in out)
    hx<HX_RHS>(nb, hx_args(U, dfx, fy, active), st);
    dy_apply(nb, false, fy, dfx, r, active, st);
    for (int s = 2; s <= 4; ++s) {
      hx<HX_RHS, true>(nb, hx_comb_args(s - 1, r, U, out, dt, dt_dev, dfx, fy, active), st);
      dy_apply(nb, false, fy, dfx, r, active, st);
    }
    if (filt) {
      // stage-4 combine -> x mask*sigma, then the nodal filter along y
      T* tmp = W(ws, 0);
      hx<HX_FILTER, true>(nb, hx_comb_args(4, r, U, out, dt, dt_dev, tmp, nullptr, active), st);
      dy_apply(nb, true, tmp, nullptr, out, active, st);
    } else {
      dim3 grid((ncell + NT - 1) / NT, nb);
      KScope ks("h_rk", st);
      hipLaunchKernelGGL((k_hrk<T>), grid, dim3(NT), 0, st, (const T*)r, U, out, us, ncell, ncell, 4 * ncell, 4,
                         dt, dt_dev, active, T(gamma), nx, ny, bc);
      check_launch("k_hrk");
    }
    if (bc) {
      KScope ks("h_wall", st);
      hipLaunchKernelGGL((k_hbc<T>), dim3((nx + NT - 1) / NT, nb), dim3(NT), 0, st, out, ncell, 4 * ncell, nx, ny,
                         active);
      check_launch("k_hbc");
    }
  }
  void max_speed(int nb, const T* U, double* mx, double* dt_out, double cfl, void* ws, hipStream_t st) const {
    double* part = Part(ws);
    const int npb = (int)std::min<long long>(NPART, (ncell + NT - 1) / NT);
    hipLaunchKernelGGL((k_maxspeed_partial<T, 2>), dim3(npb, nb), dim3(NT), 0, st, U, ncell, 4 * ncell, ncell,
                       T(gamma), part, (const int*)nullptr);
    check_launch("k_maxspeed_partial(hyb)");
    hipLaunchKernelGGL((k_max_final<2>), dim3(nb), dim3(NT), 0, st, (const double*)part, npb, mx, dt_out,
                       cfl * std::min(dx, dy_min), (const int*)nullptr);
    check_launch("k_max_final(hyb)");
  }
};

template struct PlanHyb<double>;
template struct PlanHyb<float>;

}  // namespace ph

namespace ph {
void set_last_error(const char* msg);
}

namespace {
template <typename F>
int hguard(F&& f) {
  try {
    f();
    return PH_OK;
  } catch (const ph::Error& e) {
    ph::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    ph::set_last_error(e.what());
    return PH_ERR_INTERNAL;
  } catch (...) {
    ph::set_last_error("unknown error");
    return PH_ERR_INTERNAL;
  }
}
template <typename F>
void withh(ph_planh_t p, F&& f) {
  if (!p) throw ph::Error(PH_ERR_ARG, "null plan");
  ph::PlanBase* b = reinterpret_cast<ph::PlanBase*>(p);
  if (auto* d = dynamic_cast<ph::PlanHyb<double>*>(b)) f(*d);
  else if (auto* s = dynamic_cast<ph::PlanHyb<float>*>(b)) f(*s);
  else throw ph::Error(PH_ERR_ARG, "not a hybrid plan");
}
}  // namespace

extern "C" {

int ph_planh_create(ph_planh_t* plan, int nx, int ny, double lx, double dy_min, int dtype, int dealias,
                    int filter_enabled, int filter_p, double filter_alpha, double gamma, int batch,
                    const double* D_host, const double* Fn_host, int gll_mode, int y_bc) {
  return hguard([&] {
    if (!plan || !D_host || !Fn_host) throw ph::Error(PH_ERR_ARG, "null argument");
    if (lx <= 0 || dy_min <= 0 || gamma <= 1.0) throw ph::Error(PH_ERR_ARG, "bad lx, dy_min or gamma");
    ph::PlanBase* p;
    if (dtype == PH_F64)
      p = new ph::PlanHyb<double>(nx, ny, lx, dy_min, dealias != 0, filter_enabled != 0, filter_p, filter_alpha,
                                  gamma, batch, D_host, Fn_host, gll_mode, y_bc);
    else if (dtype == PH_F32)
      p = new ph::PlanHyb<float>(nx, ny, lx, dy_min, dealias != 0, filter_enabled != 0, filter_p, filter_alpha,
                                 gamma, batch, D_host, Fn_host, gll_mode, y_bc);
    else
      throw ph::Error(PH_ERR_ARG, "dtype must be PH_F64 or PH_F32");
    *plan = reinterpret_cast<ph_planh_t>(p);
  });
}

int ph_planh_destroy(ph_planh_t plan) {
  return hguard([&] { delete reinterpret_cast<ph::PlanBase*>(plan); });
}

int ph_planh_workspace_bytes(ph_planh_t plan, size_t* bytes) {
  return hguard([&] {
    if (!bytes) throw ph::Error(PH_ERR_ARG, "null output");
    withh(plan, [&](auto& p) { *bytes = p.wsBytes; });
  });
}

#define PH_RH using R = std::conditional_t<std::is_same_v<std::decay_t<decltype(p)>, ph::PlanHyb<double>>, double, float>

int ph_rhs2dh(ph_planh_t plan, int nbatch, const void* U, void* rhs, void* ws, void* stream) {
  return hguard([&] {
    withh(plan, [&](auto& p) {
      PH_RH;
      if (nbatch < 1 || nbatch > p.batch || !U || !rhs || !ws) throw ph::Error(PH_ERR_ARG, "bad arguments");
      ph::chain_registry().forget(ws);
      p.rhs(nbatch, static_cast<const R*>(U), static_cast<R*>(rhs), nullptr, ws,
            reinterpret_cast<hipStream_t>(stream));
    });
  });
}

int ph_step2dh(ph_planh_t plan, int nbatch, const void* U, void* out, double dt, const double* dt_dev,
               const int* active_dev, int apply_filter, void* ws, void* stream) {
  return hguard([&] {
    withh(plan, [&](auto& p) {
      PH_RH;
      if (nbatch < 1 || nbatch > p.batch || !U || !out || !ws) throw ph::Error(PH_ERR_ARG, "bad arguments");
      ph::chain_registry().forget(ws);
      const size_t n = 4ull * size_t(nbatch) * size_t(p.ncell) * sizeof(R);
      if (ph::ranges_overlap(U, n, out, n)) throw ph::Error(PH_ERR_ARG, "U and out must not overlap");
      p.step(nbatch, static_cast<const R*>(U), static_cast<R*>(out), dt, dt_dev, active_dev, apply_filter != 0,
             ws, reinterpret_cast<hipStream_t>(stream));
    });
  });
}

int ph_max_wave_speed2dh(ph_planh_t plan, int nbatch, const void* U, double* max_dev, double* dt_dev,
                         double cfl, void* ws, void* stream) {
  return hguard([&] {
    withh(plan, [&](auto& p) {
      PH_RH;
      if (nbatch < 1 || nbatch > p.batch || !U || !ws || !(max_dev || dt_dev))
        throw ph::Error(PH_ERR_ARG, "bad arguments");
      p.max_speed(nbatch, static_cast<const R*>(U), max_dev, dt_dev, cfl, ws,
                  reinterpret_cast<hipStream_t>(stream));
    });
  });
}

}  // extern "C"
