// 3D ideal MHD on the Fourier pencil pipeline (SpectralSolverMHD3D with the
// reference defaults: resistivity and viscosity 0, Helmholtz projection of B
// after every RK substage, spectral filter then projection after every step;
// phrike/solver.py:840-1050, 1173; MHDEquations3D, phrike/equations.py:
// 911-1100; Grid3D.project_solenoidal, phrike/grid.py:1648-1684).
//
// State U (8, Nz, Ny, Nx): rho, rho ux, rho uy, rho uz, E, Bx, By, Bz.
//   RHS    -(dx Fx + dy Fy + dz Fz): the Euler pencil pipeline with 8 fields
//          (k_xfwd XF_MHD3: fluxes of the three directions in registers, x
//          transforms; y pass 24 -> 16; z pass 16 -> 8; y and x inverses).
//   stage  the RHS's x inverse (XI_MHDSTAGE) also does the RK4 combine in
//          the reference's order and the clamp rho >= 1e-10, p >= 1e-12
//          through prim -> cons, so k never goes to HBM.
//   proj   an unpruned plan (every mode kept, no filter): B to spectral space
//          (x, y, z forward), k_mhd_proj zeroes the Nyquist planes and removes
//          k (k.B^)/k^2, back along z, y, x; the x inverse (XI_PROJE) writes B
//          and E = p/(gamma-1) + KE + |B|^2/2 at the thermal pressure p of the
//          unprojected state.
//   filter the main plan's spectral filter of all 8 fields.
//   fused  for a solenoidal U (PH_STEP_SOLENOIDAL): the stage boundaries as
//          one k_xmid each, one projection per step, inside the filter's z
//          pass (step_fused).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/phrike_hip.h"
#include "pipeline.hpp"

namespace ph {

constexpr double MHD_RHO_MIN = 1e-10;  // phrike/solver.py:877
constexpr double MHD_P_MIN = 1e-12;    // phrike/solver.py:878

// Leray projection of the three B spectra (all modes kept), in place:
// Nyquist planes zeroed, B^ -= k (k.B^)/k^2 with k^2 = 1 at k = 0.  Element
// (f, mz, my, c) at f*fs + mz*zs + my*ys + c, c a column-pass element (one
// FP64 complex, or two FP32 columns).
template <typename CE>
__global__ __launch_bounds__(NT) void k_mhd_proj(cx<CE>* B, long long fs, long long zs, long long ys, int ncol,
                                                int ny, int nz, const CE* kxc, const CE* nyqc, const CE* ky,
                                                const CE* kz) {
  const long long total = (long long)nz * ny * ncol;
  for (long long e = (long long)blockIdx.x * NT + threadIdx.x; e < total; e += (long long)gridDim.x * NT) {
    const int c = int(e % ncol);
    const long long r = e / ncol;
    const int my = int(r % ny), mz = int(r / ny);
    const long long o = (long long)mz * zs + (long long)my * ys + c;
    cx<CE> bx = B[o], by = B[fs + o], bz = B[2 * fs + o];
    const CE z0 = CE(0.0);
    if (2 * my == ny || 2 * mz == nz) {
      bx = mk(z0, z0);
      by = bx;
      bz = bx;
    } else {
      const CE m = nyqc[c];  // 0 at the x Nyquist column, 1 elsewhere
      bx = cscale(bx, m);
      by = cscale(by, m);
      bz = cscale(bz, m);
      const CE kx = kxc[c], kyv = ky[my], kzv = kz[mz];
      const CE k2 = (kx * kx + kyv * kyv) + kzv * kzv;
      const cx<CE> kb = mk((kx * bx.x + kyv * by.x) + kzv * bz.x, (kx * bx.y + kyv * by.y) + kzv * bz.y);
      CE inv;
      if constexpr (sizeof(CE) == 8 && !std::is_same<CE, double>::value) {
        inv = CE(k2.a == 0.f ? 1.f : 1.f / k2.a, k2.b == 0.f ? 1.f : 1.f / k2.b);
        const cx<CE> cf = mk(kb.x * inv, kb.y * inv);
        bx = bx - cscale(cf, kx);
        by = by - cscale(cf, kyv);
        bz = bz - cscale(cf, kzv);
      } else {
        const CE k2s = k2 == CE(0.0) ? CE(1.0) : k2;
        const cx<CE> cf = mk(kb.x / k2s, kb.y / k2s);
        bx = bx - cscale(cf, kx);
        by = by - cscale(cf, kyv);
        bz = bz - cscale(cf, kzv);
      }
    }
    B[o] = bx;
    B[fs + o] = by;
    B[2 * fs + o] = bz;
  }
}

// CFL speed max(|ux|,|uy|,|uz|) + sqrt(max(a^2 + b^2, 0)), a^2 = gamma p/rho,
// b^2 = |B|^2/rho (phrike/equations.py:1041-1061): per-block partials
template <typename T>
__global__ __launch_bounds__(NT) void k_mhd_maxspeed(const T* u, long long n, T gamma, double* part) {
  __shared__ double red[NT];
  double m = 0.0;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT) {
    const T rho = u[i];
    const T ux = u[n + i] / rho, uy = u[2 * n + i] / rho, uz = u[3 * n + i] / rho;
    const T Bx = u[5 * n + i], By = u[6 * n + i], Bz = u[7 * n + i];
    const T kin = (T(0.5) * rho) * ((ux * ux + uy * uy) + uz * uz);
    const T b2s = (Bx * Bx + By * By) + Bz * Bz;
    const T p = (gamma - T(1)) * ((u[4 * n + i] - kin) - T(0.5) * b2s);
    const T a2 = gamma * p / rho, b2 = b2s / rho;
    const T s = a2 + b2;
    const T fast = sqrt(s > T(0) ? s : T(0));
    T c = fabs(ux) > fabs(uy) ? fabs(ux) : fabs(uy);
    c = c > fabs(uz) ? c : fabs(uz);
    m = dmax_nan(m, double(c + fast));
  }
  red[threadIdx.x] = m;
  lds_barrier();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = dmax_nan(red[threadIdx.x], red[threadIdx.x + w]);
    lds_barrier();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// conserved quantities (phrike/equations.py:1083-1100): mass, momenta,
// total energy, magnetic energy, cross helicity; FP64 partials per block
template <typename T>
__global__ __launch_bounds__(NT) void k_mhd_sums(const T* u, long long n, T gamma, double* part) {
  constexpr int NS = 7;
  __shared__ double red[NS][NT];
  double acc[NS] = {0, 0, 0, 0, 0, 0, 0};
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT) {
    const T rho = u[i];
    const T ux = u[n + i] / rho, uy = u[2 * n + i] / rho, uz = u[3 * n + i] / rho;
    const T Bx = u[5 * n + i], By = u[6 * n + i], Bz = u[7 * n + i];
    const T q2 = (ux * ux + uy * uy) + uz * uz;
    const T b2 = (Bx * Bx + By * By) + Bz * Bz;
    const T p = (gamma - T(1)) * ((u[4 * n + i] - (T(0.5) * rho) * q2) - T(0.5) * b2);
    acc[0] += double(rho);
    acc[1] += double(rho * ux);
    acc[2] += double(rho * uy);
    acc[3] += double(rho * uz);
    acc[4] += double((p / (gamma - T(1)) + (T(0.5) * rho) * q2) + T(0.5) * b2);
    acc[5] += double(T(0.5) * b2);
    acc[6] += double((ux * Bx + uy * By) + uz * Bz);
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) red[s][threadIdx.x] = acc[s];
  lds_barrier();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
#pragma unroll
      for (int s = 0; s < NS; ++s) red[s][threadIdx.x] += red[s][threadIdx.x + w];
    }
    lds_barrier();
  }
  if (threadIdx.x < NS) part[(long long)blockIdx.x * NS + threadIdx.x] = red[threadIdx.x][0];
}

template <typename T>
struct PlanMHD3D : PlanBase {
  using CE = typename ColElem<T>::type;
  static constexpr int PW = ColElem<T>::PW;
  static constexpr int NPART = 1024;
  Plan3D<T> main;  // RHS and filter (2/3 rule, filter as configured), 24/16 spectral fields
  Plan3D<T> prj;   // projection: every mode kept, no filter, 3/3 spectral fields
  int nx, ny, nz;
  long long ncell;
  double gamma;
  CE *kxc = nullptr, *nyqc = nullptr, *kyv = nullptr, *kzv = nullptr;
  size_t offMain, offPrj, offAcc, offUs, offPart, wsBytes;

  PlanMHD3D(int nx_, int ny_, int nz_, double lx, double ly, double lz, bool dealias, bool filt, int p,
            double alpha, double gamma_)
      : main(nx_, ny_, nz_, lx, ly, lz, dealias, filt, p, alpha, gamma_, PH_MHD_SYM ? 15 : 24, 16),
        prj(nx_, ny_, nz_, lx, ly, lz, false, false, p, alpha, gamma_, 3, 3),
        nx(nx_), ny(ny_), nz(nz_), gamma(gamma_) {
    dtype = sizeof(T) == 8 ? 0 : 1;
    ncell = (long long)nx * ny * nz;
    // projection wavenumbers by column-pass element (fftfreq, as the
    // reference's kx, ky, kz; Nyquist planes are zeroed, so their k is moot)
    const int ncol = (prj.mx + PW - 1) / PW;
    std::vector<CE> hk(ncol), hm(ncol), hy(ny), hz(nz);
    auto kval = [](int m, int n, double L) {
      const int f = m < (n + 1) / 2 ? m : m - n;
      return 2.0 * M_PI * (double(f) * (1.0 / (n * (L / n))));
    };
    for (int c = 0; c < ncol; ++c) {
      if constexpr (PW == 2) {
        const int m0 = 2 * c, m1 = 2 * c + 1;
        hk[c] = CE(float(kval(m0, nx, lx)), float(kval(m1, nx, lx)));
        hm[c] = CE(2 * m0 == nx ? 0.f : 1.f, 2 * m1 == nx ? 0.f : 1.f);
      } else {
        hk[c] = CE(kval(c, nx, lx));
        hm[c] = CE(2 * c == nx ? 0.0 : 1.0);
      }
    }
    for (int m = 0; m < ny; ++m) hy[m] = CE(kval(m, ny, ly));
    for (int m = 0; m < nz; ++m) hz[m] = CE(kval(m, nz, lz));
    auto up = [](CE*& d, const std::vector<CE>& h, const char* what) {
      hip_check(hipMalloc(&d, sizeof(CE) * h.size()), what);
      hip_check(hipMemcpy(d, h.data(), sizeof(CE) * h.size(), hipMemcpyHostToDevice), what);
    };
    up(kxc, hk, "projection kx");
    up(nyqc, hm, "projection Nyquist mask");
    up(kyv, hy, "projection ky");
    up(kzv, hz, "projection kz");
    auto al = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t st = sizeof(T) * 8ull * ncell;
    offMain = 0;
    offPrj = al(main.wsBytes);
    offAcc = al(offPrj + prj.wsBytes);
    offUs = al(offAcc + st);
    offPart = al(offUs + st);
    wsBytes = al(offPart + sizeof(double) * NPART * MON_NRED);
  }
  ~PlanMHD3D() override {
    for (CE* p : {kxc, nyqc, kyv, kzv})
      if (p) (void)hipFree(p);
  }
  void* wsMain(void* ws) const { return (char*)ws + offMain; }
  void* wsPrj(void* ws) const { return (char*)ws + offPrj; }
  T* Acc(void* ws) const { return reinterpret_cast<T*>((char*)ws + offAcc); }
  T* Us(void* ws) const { return reinterpret_cast<T*>((char*)ws + offUs); }
  double* Part(void* ws) const { return reinterpret_cast<double*>((char*)ws + offPart); }

  static int blocks(long long n) { return (int)std::min<long long>((n + NT - 1) / NT, NPART); }

  // rhs = -(dx Fx + dy Fy + dz Fz) (phrike/solver.py:932-954)
  void rhs(const T* U, T* out, void* ws, hipStream_t st) const {
    void* w = wsMain(ws);
    main.xfwd(XF_MHD3, U, 8, 0xFFu, w, st);
    main.yfwd(true, 8, 0, w, st);
    main.zpass(true, 8, 0, w, st);
    main.yinv(8, w, st);
    XInvArgs<T> a = main.xinv_args(w);
    a.nf = 8;
    a.out = out;
    KScope ks("xinv_rhs8", st);
    launch_xinv<T, 8, XI_RHS>(nx, ny, nz, a, st);
    check_launch("k_xinv(mhd rhs)");
  }
  // B <- P B, E at fixed thermal pressure, in place in U (phrike/solver.py:966-986)
  void project(T* U, void* ws, hipStream_t st) const {
    void* w = wsPrj(ws);
    prj.xfwd(XF_COPY, U + 5 * ncell, 3, 0u, w, st);
    prj.yfwd(false, 3, 0, w, st);
    prj.zline(true, 3, w, st);
    {
      KScope ks("mhd_proj", st);
      const long long fs = prj.b_fs() / PW, zs = prj.b_zs() / PW, ys = prj.b_ys() / PW;
      const int ncol = (prj.mx + PW - 1) / PW;
      const long long total = (long long)nz * ny * ncol;
      const int nb = (int)std::min<long long>((total + NT - 1) / NT, 65535);
      hipLaunchKernelGGL((k_mhd_proj<CE>), dim3(nb), dim3(NT), 0, st,
                         reinterpret_cast<cx<CE>*>(prj.B(w)), fs, zs, ys, ncol, ny, nz, kxc, nyqc, kyv, kzv);
      check_launch("k_mhd_proj");
    }
    prj.zline(false, 3, w, st);
    prj.yinv(3, w, st);
    XInvArgs<T> a = prj.xinv_args(w);
    a.nf = 3;
    a.out = U;
    KScope ks("xinv_proje", st);
    launch_xinv<T, 3, XI_PROJE>(nx, ny, nz, a, st);
    check_launch("k_xinv(mhd projection)");
  }
  // RHS of stage state Uin fused with the stage-s combine: the accumulator in
  // Acc(ws), the next (clamped) state to out (phrike/solver.py:932-964, 988-1006)
  void rhs_stage(int s, const T* Uin, const T* U, T* out, double dt, void* ws, hipStream_t st) const {
    void* w = wsMain(ws);
    main.xfwd(XF_MHD3, Uin, 8, 0xFFu, w, st);
    main.yfwd(true, 8, 0, w, st);
    main.zpass(true, 8, 0, w, st);
    main.yinv(8, w, st);
    XInvArgs<T> a = main.xinv_args(w);
    a.nf = 8;
    a.out = Acc(ws);
    a.un = U;
    a.us = out;
    a.stage = s;
    a.dt = dt;
    a.rho_min = T(MHD_RHO_MIN);
    a.p_min = T(MHD_P_MIN);
    KScope ks("xinv_mhdstage", st);
    launch_xinv<T, 8, XI_MHDSTAGE>(nx, ny, nz, a, st);
    check_launch("k_xinv(mhd stage)");
  }
  // the main plan's spectral filter (mask and sigma) of all 8 fields, in place
  void filter(T* U, void* ws, hipStream_t st) const {
    for (int f0 = 0; f0 < 8; f0 += 5) {
      const int nf = std::min(5, 8 - f0);
      main.deriv(-1, nf, U + f0 * ncell, U + f0 * ncell, wsMain(ws), st);
    }
  }
  // Fused step for a state whose B is solenoidal (the output of a previous
  // step, or a projected initial state).  The ideal-MHD induction flux is
  // antisymmetric, F_g[B_i] = u_g B_i - u_i B_g, and every spectral factor of
  // the RHS (i k_g, the 2/3 mask, sigma) is the same for all components, so
  // sum_i k_i rhs^_B,i = -i sigma sum_{g,i} k_g k_i F^_g,i = 0: the RHS of B is
  // solenoidal, every stage state U + c k keeps the solenoidal B of U (the
  // clamp leaves B alone), and the projection after each substage is the
  // identity up to round-off.  The stages therefore run as the Euler
  // pipeline does: one k_xmid per stage boundary (x inverse, RK4 accumulator
  // combine, clamp, next fluxes and x forward), stage states never in HBM;
  // then the spectral filter of the 8 fields and one projection, which
  // removes the round-off divergence (phrike/solver.py:988-1006, 1041-1050).
  void step_fused(const T* U, T* out, double dt, void* ws, hipStream_t st) const {
    void* w = wsMain(ws);
    main.xfwd(XF_MHD3, U, 8, 0xFFu, w, st);
    for (int s = 1; s <= 4; ++s) {
      main.yfwd(true, 8, 0, w, st);
      main.zpass(true, 8, 0, w, st);
      main.yinv(8, w, st);
      XMidArgs<T> m{};
      m.inv = main.xinv_args(w);
      m.inv.nf = 8;
      m.inv.out = out;  // the RK4 accumulator until stage 4
      m.inv.un = U;
      m.inv.stage = s;
      m.inv.dt = dt;
      m.inv.rho_min = T(MHD_RHO_MIN);
      m.inv.p_min = T(MHD_P_MIN);
      m.fwd = main.xfwd_args(nullptr, 8, 0xFFu, w);
      m.fwd_last = main.xfwd_args(nullptr, 8, 0u, w);
      KScope ks(s == 4 ? "xmid_mhd_last" : "xmid_mhd", st);
      launch_xmid<T, 8, XF_MHD3>(nx, ny, nz, m, st);
      check_launch("k_xmid(mhd)");
    }
    // the filter's y, z and y-inverse passes of the 8 fields; the x inverse
    // writes the filtered state.  With the 2/3 rule (no Nyquist mode kept)
    // the final projection runs inside the filter's z pass: B^ of the kept
    // modes is projected while its three components are in LDS, and E keeps
    // its filtered value -- the reference recomputes E at fixed thermal
    // pressure, E + (|P B|^2 - |B|^2)/2, and B here is the filtered state of
    // a solenoidal step, so P B - B and that difference are round-off.
    main.yfwd(false, 8, 0, w, st);
    if (proj_in_filter()) {
      main.zpass(false, 5, 0, w, st);
      main.zproj(5, w, st);
    } else {
      main.zpass(false, 8, 0, w, st);
    }
    main.yinv(8, w, st);
    {
      XInvArgs<T> a = main.xinv_args(w);
      a.nf = 8;
      a.out = out;
      KScope ks("xinv_out8", st);
      launch_xinv<T, 8, XI_OUT>(nx, ny, nz, a, st);
      check_launch("k_xinv(mhd filter)");
    }
    if (!proj_in_filter()) project(out, ws, st);
  }
  // the main plan keeps no Nyquist mode (2/3 rule along every axis)
  bool proj_in_filter() const { return main.ax.kp < nx / 2 && main.ay.kp < ny / 2 && main.az.kp < nz / 2; }
  // phrike/solver.py:988-1006, 1041-1050
  void step(const T* U, T* out, double dt, void* ws, hipStream_t st) const {
    T* us = Us(ws);
    for (int s = 1; s <= 4; ++s) {
      // stages 2, 3 overwrite us, which their forward x pass has read by then
      rhs_stage(s, s == 1 ? U : us, U, s < 4 ? us : out, dt, ws, st);
      project(s < 4 ? us : out, ws, st);
    }
    filter(out, ws, st);
    project(out, ws, st);
  }
  void max_speed(const T* U, double* out, double* dt_out, double cfl, void* ws, hipStream_t st) const {
    const int nb = blocks(ncell);
    hipLaunchKernelGGL((k_mhd_maxspeed<T>), dim3(nb), dim3(NT), 0, st, U, ncell, T(gamma), Part(ws));
    check_launch("k_mhd_maxspeed");
    const double h = std::min({main.lx / nx, main.ly / ny, main.lz / nz});
    hipLaunchKernelGGL((k_max_final<3>), dim3(1), dim3(NT), 0, st, Part(ws), nb, out, dt_out, cfl * h,
                       (const int*)nullptr);
    check_launch("k_max_final");
  }
  void sums(const T* U, double* out, void* ws, hipStream_t st) const {
    const int nb = blocks(ncell);
    hipLaunchKernelGGL((k_mhd_sums<T>), dim3(nb), dim3(NT), 0, st, U, ncell, T(gamma), Part(ws));
    check_launch("k_mhd_sums");
    hipLaunchKernelGGL((k_sums_final<7>), dim3(1), dim3(NT), 0, st, Part(ws), nb, out);
    check_launch("k_sums_final");
  }
  void monitor(const T* U, double* out, void* ws, hipStream_t st) const {
    const int nb = blocks(ncell);
    hipLaunchKernelGGL((k_monitor_partial<T, 8>), dim3(nb, 1), dim3(NT), 0, st, U, ncell, 0ll, ncell, T(gamma),
                       Part(ws));
    check_launch("k_monitor_partial(mhd)");
    hipLaunchKernelGGL((k_monitor_final<T>), dim3(1), dim3(NT), 0, st, (const double*)Part(ws), nb, U, 0ll, out);
    check_launch("k_monitor_final(mhd)");
  }
};

template struct PlanMHD3D<double>;
template struct PlanMHD3D<float>;

void set_last_error(const char* msg);

}  // namespace ph

namespace {
template <typename F>
int mguard(F&& f) {
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
void withm(ph_planmhd3d_t p, F&& f) {
  if (!p) throw ph::Error(PH_ERR_ARG, "null plan");
  auto* b = reinterpret_cast<ph::PlanBase*>(p);
  if (auto* d = dynamic_cast<ph::PlanMHD3D<double>*>(b)) f(*d);
  else if (auto* s = dynamic_cast<ph::PlanMHD3D<float>*>(b)) f(*s);
  else throw ph::Error(PH_ERR_ARG, "not an MHD plan");
}
hipStream_t MS(void* s) { return reinterpret_cast<hipStream_t>(s); }
template <typename P>
using Elem = std::conditional_t<std::is_same_v<std::decay_t<P>, ph::PlanMHD3D<double>>, double, float>;
}  // namespace

extern "C" {

int ph_planmhd3d_create(ph_planmhd3d_t* plan, int nx, int ny, int nz, double lx, double ly, double lz, int dtype,
                        int dealias, int filter_enabled, int filter_p, double filter_alpha, double gamma) {
  return mguard([&] {
    if (!plan) throw ph::Error(PH_ERR_ARG, "null argument");
    if (!(lx > 0 && ly > 0 && lz > 0) || !(gamma > 1.0)) throw ph::Error(PH_ERR_ARG, "bad lengths or gamma");
    ph::PlanBase* p;
    if (dtype == PH_F64)
      p = new ph::PlanMHD3D<double>(nx, ny, nz, lx, ly, lz, dealias != 0, filter_enabled != 0, filter_p,
                                    filter_alpha, gamma);
    else if (dtype == PH_F32)
      p = new ph::PlanMHD3D<float>(nx, ny, nz, lx, ly, lz, dealias != 0, filter_enabled != 0, filter_p,
                                   filter_alpha, gamma);
    else
      throw ph::Error(PH_ERR_ARG, "dtype must be PH_F64 or PH_F32");
    *plan = reinterpret_cast<ph_planmhd3d_t>(p);
  });
}

int ph_planmhd3d_destroy(ph_planmhd3d_t plan) {
  return mguard([&] { delete reinterpret_cast<ph::PlanBase*>(plan); });
}

int ph_planmhd3d_workspace_bytes(ph_planmhd3d_t plan, size_t* bytes) {
  return mguard([&] {
    if (!bytes) throw ph::Error(PH_ERR_ARG, "null output");
    withm(plan, [&](auto& p) { *bytes = p.wsBytes; });
  });
}

int ph_rhs_mhd3d(ph_planmhd3d_t plan, const void* U, void* rhs, void* ws, void* stream) {
  return mguard([&] {
    if (!U || !rhs || !ws) throw ph::Error(PH_ERR_ARG, "null buffer");
    ph::chain_registry().forget(ws);
    withm(plan, [&](auto& p) {
      using R = Elem<decltype(p)>;
      p.rhs(static_cast<const R*>(U), static_cast<R*>(rhs), ws, MS(stream));
    });
  });
}

int ph_project_mhd3d(ph_planmhd3d_t plan, void* U, void* ws, void* stream) {
  return mguard([&] {
    if (!U || !ws) throw ph::Error(PH_ERR_ARG, "null buffer");
    ph::chain_registry().forget(ws);
    withm(plan, [&](auto& p) {
      using R = Elem<decltype(p)>;
      p.project(static_cast<R*>(U), ws, MS(stream));
    });
  });
}

int ph_step_mhd3d_ex(ph_planmhd3d_t plan, const void* U, void* out, double dt, int flags, void* ws,
                     void* stream) {
  return mguard([&] {
    if (!U || !out || !ws) throw ph::Error(PH_ERR_ARG, "null buffer");
    ph::chain_registry().forget(ws);
    if (flags & ~PH_STEP_SOLENOIDAL) throw ph::Error(PH_ERR_ARG, "unknown step flags");
    withm(plan, [&](auto& p) {
      using R = Elem<decltype(p)>;
      const size_t n = 8ull * size_t(p.ncell) * sizeof(R);
      if (ph::ranges_overlap(U, n, out, n)) throw ph::Error(PH_ERR_ARG, "U and out must not overlap");
      if (flags & PH_STEP_SOLENOIDAL) p.step_fused(static_cast<const R*>(U), static_cast<R*>(out), dt, ws, MS(stream));
      else p.step(static_cast<const R*>(U), static_cast<R*>(out), dt, ws, MS(stream));
    });
  });
}

int ph_step_mhd3d(ph_planmhd3d_t plan, const void* U, void* out, double dt, void* ws, void* stream) {
  return ph_step_mhd3d_ex(plan, U, out, dt, 0, ws, stream);
}

int ph_max_wave_speed_mhd3d(ph_planmhd3d_t plan, const void* U, double* max_dev, double* dt_dev, double cfl,
                            void* ws, void* stream) {
  return mguard([&] {
    if (!U || !ws || (!max_dev && !dt_dev)) throw ph::Error(PH_ERR_ARG, "null buffer");
    withm(plan, [&](auto& p) {
      using R = Elem<decltype(p)>;
      p.max_speed(static_cast<const R*>(U), max_dev, dt_dev, cfl, ws, MS(stream));
    });
  });
}

int ph_monitor_mhd3d(ph_planmhd3d_t plan, const void* U, double* out_dev, void* ws, void* stream) {
  return mguard([&] {
    if (!U || !ws || !out_dev) throw ph::Error(PH_ERR_ARG, "null buffer");
    withm(plan, [&](auto& p) {
      using R = Elem<decltype(p)>;
      p.monitor(static_cast<const R*>(U), out_dev, ws, MS(stream));
    });
  });
}

int ph_conserved_sums_mhd3d(ph_planmhd3d_t plan, const void* U, double* out7_dev, void* ws, void* stream) {
  return mguard([&] {
    if (!U || !ws || !out7_dev) throw ph::Error(PH_ERR_ARG, "null buffer");
    withm(plan, [&](auto& p) {
      using R = Elem<decltype(p)>;
      p.sums(static_cast<const R*>(U), out7_dev, ws, MS(stream));
    });
  });
}

}  // extern "C"
