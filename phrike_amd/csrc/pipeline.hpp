// Host-side plans and launch sequences for the 3D and batched-2D pipelines.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"

#ifndef PH_B_MYZ
#define PH_B_MYZ 0
#endif
// 3D Euler steps keep the RK4 sum in spectral space (0: in the state's
// layout, in the x kernels, with the post-step filter as separate passes)
#ifndef PH_SPECACC
#define PH_SPECACC 1
#endif
// z passes of fewer tiles than this take one field per block (blockIdx.z),
// so that few blocks do not leave a nearly empty last round on the CUs
#ifndef PH_COL_SPLIT_BLOCKS
#define PH_COL_SPLIT_BLOCKS 4096
#endif
// the same for the 3D y inverse pass (1: on)
#ifndef PH_YINV_SPLIT
#define PH_YINV_SPLIT 1
#endif

namespace ph {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(-5, std::string(what) + ": " + hipGetErrorString(e));
}

inline bool is_pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }

// do the byte ranges [a, a + na) and [b, b + nb) overlap?
inline bool ranges_overlap(const void* a, size_t na, const void* b, size_t nb) {
  const char* pa = static_cast<const char*>(a);
  const char* pb = static_cast<const char*>(b);
  return pa < pb + nb && pb < pa + na;
}

// Retained modes of one axis and its FP64 tables (phrike/grid.py:87-141,
// 1381-1392): k = 2*pi*(m*(1/(N*d))) with d = L/N in fftfreq order,
// 2/3 rule |m| <= N//3, sigma = exp(-alpha*(|m|/max(N/2,1))^p).
struct AxisTables {
  int n = 0;
  int kp = 0, kn = 0, nr = 0;  // kept positions n <= kp and n >= N-kn
  std::vector<double> k, sig, sm1;  // by compact index (x: m = 0..kp); sm1 = sig - 1
  std::vector<double> twr, twi;

  AxisTables() = default;
  AxisTables(int N, double L, bool dealias, bool filt, int p, double alpha, bool half) : n(N) {
    if (dealias) {
      kp = N / 3;
      kn = N / 3;
    } else {
      kp = N / 2;
      kn = N / 2 - 1;
    }
    if (half) kn = 0;
    nr = kp + 1 + kn;
    const double d = L / N;
    const double inv = 1.0 / (N * d);
    const double mmax = std::max(double(N / 2), 1.0);
    for (int j = 0; j < nr; ++j) {
      const int nat = j <= kp ? j : j + (N - nr);
      const int m = nat < (N + 1) / 2 ? nat : nat - N;  // fftfreq ordering
      double kv = 2.0 * M_PI * (double(m) * inv);
      if (2 * nat == N) kv = 0.0;  // Nyquist: i*k gives no real part
      k.push_back(kv);
      double s = 1.0, s1 = 0.0;
      if (filt) {
        const double g = -alpha * std::pow(std::fabs(double(m)) / mmax, double(p));
        s = std::exp(g);
        s1 = std::expm1(g);
      }
      sig.push_back(s);
      sm1.push_back(s1);
    }
    twr.resize(N);
    twi.resize(N);
    for (int m = 0; m < N; ++m) {
      const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / N;
      twr[m] = (double)cosl(a);
      twi[m] = (double)sinl(a);
    }
  }
  Keep keep() const { return Keep{kp, kn, nr}; }
};

template <typename T>
struct DevAxis {
  cx<T>* tw = nullptr;
  T* sig = nullptr;
  T* k = nullptr;
  T* one = nullptr;  // the scale alone (no filter factor), by compact index
  T* sm1 = nullptr;  // sigma - 1, by compact index
  void upload(const AxisTables& a, double scale) {
    std::vector<cx<T>> tw_h(a.n);
    for (int m = 0; m < a.n; ++m) {
      tw_h[m].x = T(a.twr[m]);
      tw_h[m].y = T(a.twi[m]);
    }
    std::vector<T> s_h(a.nr), k_h(a.nr), o_h(a.nr), m_h(a.nr);
    for (int j = 0; j < a.nr; ++j) {
      s_h[j] = T(a.sig[j] * scale);
      k_h[j] = T(a.k[j]);
      o_h[j] = T(scale);
      m_h[j] = T(a.sm1[j]);
    }
    hip_check(hipMalloc(&tw, sizeof(cx<T>) * a.n), "hipMalloc(twiddles)");
    hip_check(hipMalloc(&sig, sizeof(T) * a.nr), "hipMalloc(sigma)");
    hip_check(hipMalloc(&k, sizeof(T) * a.nr), "hipMalloc(k)");
    hip_check(hipMalloc(&one, sizeof(T) * a.nr), "hipMalloc(scale)");
    hip_check(hipMalloc(&sm1, sizeof(T) * a.nr), "hipMalloc(sigma-1)");
    hip_check(hipMemcpy(one, o_h.data(), sizeof(T) * a.nr, hipMemcpyHostToDevice), "copy scale");
    hip_check(hipMemcpy(sm1, m_h.data(), sizeof(T) * a.nr, hipMemcpyHostToDevice), "copy sigma-1");
    hip_check(hipMemcpy(tw, tw_h.data(), sizeof(cx<T>) * a.n, hipMemcpyHostToDevice), "copy twiddles");
    hip_check(hipMemcpy(sig, s_h.data(), sizeof(T) * a.nr, hipMemcpyHostToDevice), "copy sigma");
    hip_check(hipMemcpy(k, k_h.data(), sizeof(T) * a.nr, hipMemcpyHostToDevice), "copy k");
  }
  void release() {
    if (tw) (void)hipFree(tw);
    if (sig) (void)hipFree(sig);
    if (k) (void)hipFree(k);
    if (one) (void)hipFree(one);
    if (sm1) (void)hipFree(sm1);
    tw = nullptr;
    sig = nullptr;
    k = nullptr;
    one = nullptr;
    sm1 = nullptr;
  }
};

// ---------------------------------------------------------------------------
// launch helpers with compile-time lengths
// ---------------------------------------------------------------------------
#define PH_SWITCH_N(NV, BODY)                                     \
  switch (NV) {                                                   \
    case 8: { constexpr int N_ = 8; BODY; } break;                \
    case 16: { constexpr int N_ = 16; BODY; } break;              \
    case 32: { constexpr int N_ = 32; BODY; } break;              \
    case 64: { constexpr int N_ = 64; BODY; } break;              \
    case 128: { constexpr int N_ = 128; BODY; } break;            \
    case 256: { constexpr int N_ = 256; BODY; } break;            \
    case 512: { constexpr int N_ = 512; BODY; } break;            \
    case 1024: { constexpr int N_ = 1024; BODY; } break;          \
    default: throw Error(-2, "unsupported axis length " + std::to_string(NV)); \
  }

inline bool supported_len(int n) { return n >= 8 && n <= 1024 && is_pow2(n); }

// the x kernels are instantiated for both kept-mode counts: N/3 + 1 (2/3 rule)
// and N/2 + 1
// the x-forward stores cover the row pitch the kernel assumes (XRGeo::MCP chunks)
template <typename T, int N, bool DA>
void check_x_pitch(int mxp) {
  using G = XRGeo<T, N, DA>;
  if (mxp != G::MCP * G::K) throw Error(-9, "spectral row pitch does not match the x kernels");
}
template <typename T, int N, int MODE>
void launch_xfwd_n(int mx, dim3 grid, const XFwdArgs<T>& a, hipStream_t st) {
  if (mx == N / 3 + 1) {
    check_x_pitch<T, N, true>(a.mxp);
    hipLaunchKernelGGL((k_xfwd<T, N, true, MODE>), grid, dim3(XLaunch<T, N>::threads), 0, st, a);
  } else if (mx == N / 2 + 1) {
    check_x_pitch<T, N, false>(a.mxp);
    hipLaunchKernelGGL((k_xfwd<T, N, false, MODE>), grid, dim3(XLaunch<T, N>::threads), 0, st, a);
  } else {
    throw Error(-2, "unsupported x mode count");
  }
}
template <typename T, int MODE>
void launch_xfwd(int nx, int nrows, int nslab, const XFwdArgs<T>& a, hipStream_t st) {
  PH_SWITCH_N(nx, (launch_xfwd_n<T, N_, MODE>(a.mx, dim3(nrows / XLaunch<T, N_>::rows, nslab), a, st)));
}

template <typename T, int MODE>
void launch_col(int n, int nblocky, const ColArgs<T>& a0, hipStream_t st) {
  PH_SWITCH_N(n, ({
    constexpr int TM = ColTile<T, N_, MODE>::TM;
    ColArgs<T> a = a0;
    const int gx = (a.ncol + TM - 1) / TM;
    // cpb < 0: one field per block when the launch has few tiles
    if (a.cpb < 0) a.cpb = (long long)gx * nblocky < PH_COL_SPLIT_BLOCKS ? 1 : 0;
    dim3 grid(gx, nblocky, a.cpb > 0 ? (a.nc + a.cpb - 1) / a.cpb : 1);
    hipLaunchKernelGGL((k_col<T, N_, MODE>), grid, dim3(ColTile<T, N_, MODE>::NT), 0, st, a);
  }));
}

template <typename T, int N, int NC, int MODE>
void launch_xinv_n(int mx, dim3 grid, const XInvArgs<T>& a, hipStream_t st) {
  if (mx == N / 3 + 1) hipLaunchKernelGGL((k_xinv<T, N, true, NC, MODE>), grid, dim3(XLaunch<T, N>::threads), 0, st, a);
  else if (mx == N / 2 + 1) hipLaunchKernelGGL((k_xinv<T, N, false, NC, MODE>), grid, dim3(XLaunch<T, N>::threads), 0, st, a);
  else throw Error(-2, "unsupported x mode count");
}
template <typename T, int NC, int MODE>
void launch_xinv(int nx, int nrows, int nslab, const XInvArgs<T>& a, hipStream_t st) {
  PH_SWITCH_N(nx, (launch_xinv_n<T, N_, NC, MODE>(a.mx, dim3(nrows / XLaunch<T, N_>::rows, nslab), a, st)));
}

template <typename T, int N, int NC, int FMODE>
void launch_xmid_n(int mx, dim3 grid, const XMidArgs<T>& a, hipStream_t st) {
  if (mx == N / 3 + 1) {
    check_x_pitch<T, N, true>(a.fwd.mxp);
    check_x_pitch<T, N, true>(a.fwd_last.mxp);
    hipLaunchKernelGGL((k_xmid<T, N, true, NC, FMODE>), grid, dim3(XLaunch<T, N>::threads), 0, st, a);
  } else if (mx == N / 2 + 1) {
    check_x_pitch<T, N, false>(a.fwd.mxp);
    check_x_pitch<T, N, false>(a.fwd_last.mxp);
    hipLaunchKernelGGL((k_xmid<T, N, false, NC, FMODE>), grid, dim3(XLaunch<T, N>::threads), 0, st, a);
  } else {
    throw Error(-2, "unsupported x mode count");
  }
}
template <typename T, int NC, int FMODE>
void launch_xmid(int nx, int nrows, int nslab, const XMidArgs<T>& a, hipStream_t st) {
  PH_SWITCH_N(nx, (launch_xmid_n<T, N_, NC, FMODE>(a.inv.mx, dim3(nrows / XLaunch<T, N_>::rows, nslab), a, st)));
}

template <typename T, int N, int NC, int MODE>
void launch_xend_n(int mx, dim3 grid, const XMidArgs<T>& a, hipStream_t st) {
  if (mx == N / 3 + 1) {
    check_x_pitch<T, N, true>(a.fwd.mxp);
    hipLaunchKernelGGL((k_xend<T, N, true, NC, MODE>), grid, dim3(XLaunch<T, N>::threads), 0, st, a);
  } else if (mx == N / 2 + 1) {
    check_x_pitch<T, N, false>(a.fwd.mxp);
    hipLaunchKernelGGL((k_xend<T, N, false, NC, MODE>), grid, dim3(XLaunch<T, N>::threads), 0, st, a);
  } else {
    throw Error(-2, "unsupported x mode count");
  }
}
template <typename T, int NC, int MODE>
void launch_xend(int nx, int nrows, int nslab, const XMidArgs<T>& a, hipStream_t st) {
  PH_SWITCH_N(nx, (launch_xend_n<T, N_, NC, MODE>(a.inv.mx, dim3(nrows / XLaunch<T, N_>::rows, nslab), a, st)));
}

inline void check_launch(const char* what) { hip_check(hipGetLastError(), what); }

// Chained steps.  A step asked to chain (PH_STEP_CHAIN) ends with k_xend,
// which leaves the x-transformed fluxes of the new state in the workspace's
// spectral buffer; the next chained step of the same plan on that state and
// workspace then skips its first x pass.  What a workspace holds is recorded
// here, by workspace address: a chained step records (plan, state), every
// other operation that writes a workspace's spectral buffers forgets it.
struct ChainRegistry {
  struct Entry {
    const void* plan;
    const void* state;
    const void* alt;  // a run loop's other buffer (its step count picks the state), or null
  };
  std::mutex mu;
  std::map<const void*, Entry> held;  // ws -> entry
  void set(const void* ws, const void* plan, const void* state, const void* alt = nullptr) {
    std::lock_guard<std::mutex> g(mu);
    held[ws] = {plan, state, alt};
  }
  void forget(const void* ws) {
    std::lock_guard<std::mutex> g(mu);
    held.erase(ws);
  }
  bool holds(const void* ws, const void* plan, const void* state) {
    std::lock_guard<std::mutex> g(mu);
    auto it = held.find(ws);
    return it != held.end() && it->second.plan == plan &&
           (it->second.state == state || (it->second.alt != nullptr && it->second.alt == state));
  }
};
inline ChainRegistry& chain_registry() {
  static ChainRegistry r;
  return r;
}

// D.U on matrix cores (gll.hip): out[b] = D F[b] for b < nbatch, or the
// hybrid-RHS epilogue out = -(add + D F) when add is not null.  With `active`,
// batch element b is skipped (out untouched) when active[b / group] == 0.
void gll_apply(int mode, int ny, int nx, int nbatch, const void* D, const void* Dhi, const void* Dlo,
               const void* F, const void* add, void* out, hipStream_t st, const int* active = nullptr,
               int group = 1);

// Optional per-kernel timing: when enabled, every pipeline launch is
// bracketed by two HIP events on its own stream; collect() waits for them and
// accumulates elapsed milliseconds per kernel tag.
struct KTimer {
  struct Rec {
    std::string tag;
    hipEvent_t a, b;
  };
  bool on = false;
  std::mutex mu;
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, long>> acc;

  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
  }
  void collect() {
    std::lock_guard<std::mutex> g(mu);
    for (auto& r : pending) {
      hip_check(hipEventSynchronize(r.b), "hipEventSynchronize");
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
      auto& v = acc[r.tag];
      v.first += ms;
      v.second += 1;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
  }
};
inline KTimer& ktimer() {
  static KTimer t;
  return t;
}
struct KScope {
  hipEvent_t a = nullptr, b = nullptr;
  const char* tag;
  hipStream_t st;
  KScope(const char* t, hipStream_t s) : tag(t), st(s) {
    KTimer& k = ktimer();
    if (!k.on) return;
    std::lock_guard<std::mutex> g(k.mu);
    a = k.get();
    b = k.get();
    hip_check(hipEventRecord(a, st), "hipEventRecord");
  }
  ~KScope() {
    if (!a) return;
    KTimer& k = ktimer();
    std::lock_guard<std::mutex> g(k.mu);
    if (hipEventRecord(b, st) == hipSuccess) k.pending.push_back({tag, a, b});
  }
};

// ---------------------------------------------------------------------------
// 3D plan
// ---------------------------------------------------------------------------
struct PlanBase {
  virtual ~PlanBase() = default;
  int dtype = 0;  // 0 = float64, 1 = float32
};

template <typename T>
struct Plan3D : PlanBase {
  int nx, ny, nz;
  double lx, ly, lz, gamma;
  AxisTables ax, ay, az;
  using CE = typename ColElem<T>::type;  // column-pass element
  static constexpr int PW = ColElem<T>::PW;
  DevAxis<T> dx;
  DevAxis<CE> dy, dz;
  int mx, mxp;
  long long ncell;
  CE* xcol_sm1 = nullptr;  // sigma_x - 1 per column of the column passes
  CE* xcol_k = nullptr;    // kx per column of the column passes (zero past the kept modes)
  T* unit_x = nullptr;     // ones by kept x mode (the unscaled forward transform)
  bool full = false;       // every mode kept and no filter: fft3d / ifft3d allowed
  bool symflux = true;     // Euler fluxes in the 12-field symmetric layout (x pass without i kx)
  bool symmhd = false;     // ideal-MHD fluxes in the 15-field symmetric layout (x pass without i kx)
  // the Euler step keeps its RK4 sum in 3D spectral space (S, updated by the
  // z passes; the 2/3 rule on): no accumulator in the x kernels, and the
  // post-step filter is the stage-4 z pass's factor on S
  bool specacc = false;
  CE* xcol_sig = nullptr;  // sigma_x per column of the column passes (zero past the kept modes)
  // workspace layout (bytes)
  size_t offA, offB, offUs, offS, offPart, offCfl, wsBytes;
  static constexpr int NPART = 1024;

  // na / nb: fields of the spectral buffers A (x pass output) and B (y pass
  // output): 12 / 10 for the Euler RHS (the 12 distinct fluxes of the
  // symmetric layout, euler3_flux_sym), 24 / 16 for ideal MHD
  Plan3D(int nx_, int ny_, int nz_, double lx_, double ly_, double lz_, bool dealias, bool filt,
         int p, double alpha, double gamma_, int na = 12, int nb = 10)
      : nx(nx_), ny(ny_), nz(nz_), lx(lx_), ly(ly_), lz(lz_), gamma(gamma_), symflux(na == 12),
        symmhd(na == 15), specacc(na == 12 && dealias && PH_SPECACC != 0) {
    dtype = sizeof(T) == 8 ? 0 : 1;
    if (!supported_len(nx) || !supported_len(ny) || !supported_len(nz))
      throw Error(-2, "grid sizes must be powers of two in [8, 1024]");
    ax = AxisTables(nx, lx, dealias, filt, p, alpha, true);
    ay = AxisTables(ny, ly, dealias, filt, p, alpha, false);
    az = AxisTables(nz, lz, dealias, filt, p, alpha, false);
    mx = ax.nr;
    const int align = sizeof(T) == 8 ? 8 : 16;  // 128-byte rows (cache lines)
    mxp = (mx + align - 1) / align * align;
    ncell = (long long)nx * ny * nz;
    const double scale = 1.0 / (double(nx) * double(ny) * double(nz));
    dx.upload(ax, scale);
    dy.upload(ay, 1.0);
    dz.upload(az, 1.0);
    {
      // by column of the row pitch (the column passes visit whole tiles),
      // zero past the kept modes
      const int ncol = (mx + PW - 1) / PW, ncolw = mxp / PW;
      std::vector<CE> h(ncolw, CE(0.0));
      for (int c = 0; c < ncol; ++c) {
        if constexpr (PW == 2) {
          const double v0 = ax.sm1[2 * c], v1 = 2 * c + 1 < mx ? ax.sm1[2 * c + 1] : 0.0;
          h[c] = CE(float(v0), float(v1));
        } else {
          h[c] = CE(ax.sm1[c]);
        }
      }
      hip_check(hipMalloc(&xcol_sm1, sizeof(CE) * ncolw), "hipMalloc(column sigma-1)");
      hip_check(hipMemcpy(xcol_sm1, h.data(), sizeof(CE) * ncolw, hipMemcpyHostToDevice), "copy column sigma-1");
      std::vector<CE> hk(ncolw, CE(0.0));
      for (int c = 0; c < ncol; ++c) {
        if constexpr (PW == 2) {
          hk[c] = CE(float(ax.k[2 * c]), float(2 * c + 1 < mx ? ax.k[2 * c + 1] : 0.0));
        } else {
          hk[c] = CE(ax.k[c]);
        }
      }
      hip_check(hipMalloc(&xcol_k, sizeof(CE) * ncolw), "hipMalloc(column k)");
      hip_check(hipMemcpy(xcol_k, hk.data(), sizeof(CE) * ncolw, hipMemcpyHostToDevice), "copy column k");
      std::vector<CE> hs(ncolw, CE(0.0));
      for (int c = 0; c < ncol; ++c) {
        if constexpr (PW == 2) {
          hs[c] = CE(float(ax.sig[2 * c]), float(2 * c + 1 < mx ? ax.sig[2 * c + 1] : 0.0));
        } else {
          hs[c] = CE(ax.sig[c]);
        }
      }
      hip_check(hipMalloc(&xcol_sig, sizeof(CE) * ncolw), "hipMalloc(column sigma)");
      hip_check(hipMemcpy(xcol_sig, hs.data(), sizeof(CE) * ncolw, hipMemcpyHostToDevice), "copy column sigma");
    }
    {
      std::vector<T> ones(mx, T(1));
      hip_check(hipMalloc(&unit_x, sizeof(T) * mx), "hipMalloc(unit x table)");
      hip_check(hipMemcpy(unit_x, ones.data(), sizeof(T) * mx, hipMemcpyHostToDevice), "copy unit x table");
    }
    full = !dealias && !filt;
    const size_t cs = sizeof(cx<T>);
    const size_t szA = (size_t)na * nz * ny * mxp * cs;
    const size_t szB = (size_t)nb * nz * ay.nr * mxp * cs;
    auto al = [](size_t v) { return (v + 255) / 256 * 256; };
    offA = 0;
    offB = al(offA + szA);
    offUs = offB;  // unused: the fused step keeps stage states in registers
    const size_t szS = specacc ? (size_t)5 * az.nr * ay.nr * mxp * cs : 0;
    offS = al(offB + szB);
    offPart = al(offS + szS);
    offCfl = al(offPart + sizeof(double) * NPART * MON_NRED);
    wsBytes = al(offCfl + sizeof(double) * (size_t)ny * nz);
  }
  ~Plan3D() override {
    if (xcol_sm1) (void)hipFree(xcol_sm1);
    if (xcol_k) (void)hipFree(xcol_k);
    if (xcol_sig) (void)hipFree(xcol_sig);
    if (unit_x) (void)hipFree(unit_x);
    dx.release();
    dy.release();
    dz.release();
  }

  cx<T>* A(void* ws) const { return reinterpret_cast<cx<T>*>((char*)ws + offA); }
  cx<T>* B(void* ws) const { return reinterpret_cast<cx<T>*>((char*)ws + offB); }
  T* Us(void* ws) const { return reinterpret_cast<T*>((char*)ws + offUs); }
  // the step's spectral RK4 sum: [5][kept z][kept y][mxp]
  cx<T>* S(void* ws) const { return reinterpret_cast<cx<T>*>((char*)ws + offS); }
  long long s_fs() const { return (long long)az.nr * ay.nr * mxp; }
  long long s_zs() const { return (long long)ay.nr * mxp; }
  double* Part(void* ws) const { return reinterpret_cast<double*>((char*)ws + offPart); }
  // per-row wave-speed maxima of the state the last step wrote (ny*nz)
  double* CflRows(void* ws) const { return reinterpret_cast<double*>((char*)ws + offCfl); }

  long long a_fs() const { return (long long)nz * ny * mxp; }
  long long b_fs() const { return (long long)nz * ay.nr * mxp; }
  // B strides (complex values) between z planes and between kept y modes:
  // [field][z][my][mx] by default, [field][my][z][mx] with PH_B_MYZ (z lines
  // then contiguous rows of mxp)
  long long b_zs() const { return PH_B_MYZ ? (long long)mxp : (long long)ay.nr * mxp; }
  long long b_ys() const { return PH_B_MYZ ? (long long)nz * mxp : (long long)mxp; }

  XFwdArgs<T> xfwd_args(const T* in, int nf, unsigned ikmask, void* ws, const int* gate = nullptr) const {
    XFwdArgs<T> a{};
    a.active = gate;
    a.act_all = 1;
    a.in = in;
    a.in_fs = ncell;
    a.in_ss = (long long)ny * nx;
    a.out = A(ws);
    a.out_fs = a_fs();
    a.out_ss = (long long)ny * mxp;
    a.nf = nf;
    a.mx = mx;
    a.mxp = mxp;
    a.ikmask = ikmask;
    a.tw = dx.tw;
    a.sig = dx.sig;
    a.kk = dx.k;
    a.gamma = T(gamma);
    return a;
  }
  // forward x pass
  void xfwd(int mode, const T* in, int nf, unsigned ikmask, void* ws, hipStream_t st,
            const int* gate = nullptr) const {
    XFwdArgs<T> a{};
    a.active = gate;
    a.act_all = 1;
    a.in = in;
    a.in_fs = ncell;
    a.in_ss = (long long)ny * nx;
    a.out = A(ws);
    a.out_fs = a_fs();
    a.out_ss = (long long)ny * mxp;
    a.nf = nf;
    a.mx = mx;
    a.mxp = mxp;
    a.ikmask = ikmask;
    a.tw = dx.tw;
    a.sig = dx.sig;
    a.kk = dx.k;
    a.gamma = T(gamma);
    KScope ks(mode == XF_FLUX3 ? "xfwd_flux" : (mode == XF_MHD3 ? "xfwd_mhd" : "xfwd_copy"), st);
    if (mode == XF_FLUX3) launch_xfwd<T, XF_FLUX3>(nx, ny, nz, a, st);
    else if (mode == XF_MHD3) launch_xfwd<T, XF_MHD3>(nx, ny, nz, a, st);
    else launch_xfwd<T, XF_COPY>(nx, ny, nz, a, st);
    check_launch("k_xfwd");
  }
  // y forward: A -> B
  void yfwd(bool div, int nc, int ikflag, void* ws, hipStream_t st, bool unit = false,
            const int* gate = nullptr) const {
    ColArgs<CE> a{};
    a.active = gate;
    a.act_all = 1;
    a.in = reinterpret_cast<cx<CE>*>(A(ws));
    a.in_fs = (a_fs()) / PW;
    a.in_bs = ((long long)ny * mxp) / PW;
    a.in_ls = (mxp) / PW;
    a.out = reinterpret_cast<cx<CE>*>(B(ws));
    a.out_fs = (b_fs()) / PW;
    a.out_bs = b_zs() / PW;
    a.out_ls = b_ys() / PW;
    a.ncol = (mx + PW - 1) / PW;
    a.ncolw = mxp / PW;
    a.nc = nc;
    a.keep = ay.keep();
    a.ikflag = ikflag;
    a.tw = dy.tw;
    a.sig = unit ? dy.one : dy.sig;
    a.kk = dy.k;
    a.symflux = div && symflux ? 1 : 0;
    if (a.symflux && nc != 5) throw Error(-2, "the symmetric flux layout takes the 5 Euler fields");
    if (div && symmhd && nc != 8) throw Error(-2, "the symmetric MHD flux layout takes the 8 MHD fields");
    a.colk = xcol_k;
    KScope ks(div ? "yfwd_div" : (unit ? "yfwd_unit" : "yfwd_copy"), st);
    if (div && symmhd) launch_col<CE, CM_FWD_DIV3M>(ny, nz, a, st);
    else if (div) launch_col<CE, CM_FWD_DIV3>(ny, nz, a, st);
    else launch_col<CE, CM_FWD_COPY>(ny, nz, a, st);
    check_launch("k_col(y fwd)");
  }
  // z forward + combine + inverse, in place in B
  void zpass(bool div, int nc, int ikflag, void* ws, hipStream_t st, const int* gate = nullptr) const {
    ColArgs<CE> a{};
    a.active = gate;
    a.act_all = 1;
    a.in = reinterpret_cast<cx<CE>*>(B(ws));
    a.in_fs = (b_fs()) / PW;
    a.in_bs = b_ys() / PW;
    a.in_ls = b_zs() / PW;
    a.out = reinterpret_cast<cx<CE>*>(B(ws));
    a.out_fs = a.in_fs;
    a.out_bs = a.in_bs;
    a.out_ls = a.in_ls;
    a.ncol = (mx + PW - 1) / PW;
    a.ncolw = mxp / PW;
    a.nc = nc;
    a.keep = az.keep();
    a.ikflag = ikflag;
    a.tw = dz.tw;
    a.sig = dz.sig;
    a.kk = dz.k;
    a.cpb = -1;
    KScope ks(div ? "z_div" : "z_copy", st);
    if (div) launch_col<CE, CM_FWDINV_DIV>(nz, ay.nr, a, st);
    else launch_col<CE, CM_FWDINV_COPY>(nz, ay.nr, a, st);
    check_launch("k_col(z)");
  }
  // forward (fwd) or inverse transform along z of nc fields of B, in place;
  // only for a plan that keeps every z mode (in place needs compact = natural)
  void zline(bool fwd, int nc, void* ws, hipStream_t st) const {
    if (az.nr != nz) throw Error(-9, "zline: the plan must keep every z mode");
    ColArgs<CE> a{};
    a.in = reinterpret_cast<cx<CE>*>(B(ws));
    a.in_fs = (b_fs()) / PW;
    a.in_bs = b_ys() / PW;
    a.in_ls = b_zs() / PW;
    a.out = reinterpret_cast<cx<CE>*>(B(ws));
    a.out_fs = a.in_fs;
    a.out_bs = a.in_bs;
    a.out_ls = a.in_ls;
    a.ncol = (mx + PW - 1) / PW;
    a.ncolw = mxp / PW;
    a.nc = nc;
    a.keep = az.keep();
    a.tw = dz.tw;
    a.sig = dz.sig;
    a.kk = dz.k;
    KScope ks(fwd ? "z_fwd" : "z_inv", st);
    if (fwd) launch_col<CE, CM_FWD_COPY>(nz, ay.nr, a, st);
    else launch_col<CE, CM_INV>(nz, ay.nr, a, st);
    check_launch("k_col(z line)");
  }
  // fields f0..f0+2 of B (a solenoidal field's three components) along z:
  // forward, sigma_z and the Leray projection on kept modes, inverse, in place
  void zproj(int f0, void* ws, hipStream_t st) const {
    ColArgs<CE> a{};
    cx<CE>* b = reinterpret_cast<cx<CE>*>(B(ws)) + (long long)f0 * (b_fs() / PW);
    a.in = b;
    a.in_fs = (b_fs()) / PW;
    a.in_bs = b_ys() / PW;
    a.in_ls = b_zs() / PW;
    a.out = b;
    a.out_fs = a.in_fs;
    a.out_bs = a.in_bs;
    a.out_ls = a.in_ls;
    a.ncol = (mx + PW - 1) / PW;
    a.ncolw = mxp / PW;
    a.nc = 3;
    a.keep = az.keep();
    a.tw = dz.tw;
    a.sig = dz.sig;
    a.kk = dz.k;
    a.colk = xcol_k;
    a.rowk = dy.k;
    KScope ks("z_proj", st);
    launch_col<CE, CM_FWDINV_PROJ>(nz, ay.nr, a, st);
    check_launch("k_col(z projection)");
  }
  // filter correction along z: kept modes times sx sy sz - 1 (after unit x
  // and y passes), in place in B
  void zcorr(int nc, void* ws, hipStream_t st, const int* gate = nullptr) const {
    ColArgs<CE> a{};
    a.active = gate;
    a.act_all = 1;
    a.in = reinterpret_cast<cx<CE>*>(B(ws));
    a.in_fs = (b_fs()) / PW;
    a.in_bs = b_ys() / PW;
    a.in_ls = b_zs() / PW;
    a.out = reinterpret_cast<cx<CE>*>(B(ws));
    a.out_fs = a.in_fs;
    a.out_bs = a.in_bs;
    a.out_ls = a.in_ls;
    a.ncol = (mx + PW - 1) / PW;
    a.ncolw = mxp / PW;
    a.nc = nc;
    a.keep = az.keep();
    a.tw = dz.tw;
    a.sig = dz.sig;
    a.kk = dz.k;
    a.sm1 = dz.sm1;
    a.colsm1 = xcol_sm1;
    a.rowsm1 = dy.sm1;
    a.cpb = -1;
    KScope ks("z_corr", st);
    launch_col<CE, CM_FWDINV_CORR>(nz, ay.nr, a, st);
    check_launch("k_col(z corr)");
  }
  // z pass of an RK4 stage with the spectral sum (CM_FWDINV_DIVACC): as
  // zpass(true, 5, ...), and S -= w D (stages 1-3) or S = sigma (S - dt/6 D)
  // transformed back in place of D (stage 4)
  void zpass_acc(int stage, double dt, const double* dt_dev, void* ws, hipStream_t st,
                 const int* gate = nullptr) const {
    // the kernel's sum loop covers the 2N/3 + 1 kept z modes of the 2/3 rule
    if (!specacc || az.nr > 2 * (nz / 3) + 1) throw Error(-9, "spectral RK4 sum needs the 2/3 rule");
    ColArgs<CE> a{};
    a.active = gate;
    a.act_all = 1;
    a.in = reinterpret_cast<cx<CE>*>(B(ws));
    a.in_fs = (b_fs()) / PW;
    a.in_bs = b_ys() / PW;
    a.in_ls = b_zs() / PW;
    a.out = reinterpret_cast<cx<CE>*>(B(ws));
    a.out_fs = a.in_fs;
    a.out_bs = a.in_bs;
    a.out_ls = a.in_ls;
    a.ncol = (mx + PW - 1) / PW;
    a.ncolw = mxp / PW;
    a.nc = 5;
    a.keep = az.keep();
    a.tw = dz.tw;
    a.sig = dz.sig;
    a.kk = dz.k;
    a.acc = reinterpret_cast<cx<CE>*>(S(ws));
    a.acc_fs = s_fs() / PW;
    a.acc_bs = mxp / PW;
    a.acc_ls = s_zs() / PW;
    a.stage = stage;
    a.dt = dt;
    a.dt_dev = dt_dev;
    a.colsig = xcol_sig;
    a.rowsig = dy.sig;
    a.cpb = -1;
    KScope ks("z_acc", st);
    launch_col<CE, CM_FWDINV_DIVACC>(nz, ay.nr, a, st);
    check_launch("k_col(z, spectral sum)");
  }
  // S = the masked spectrum of U (scaled as the RHS spectra): x and y
  // forward with the scale alone, then the z forward to S
  void spec_init(const T* U, void* ws, hipStream_t st, const int* gate = nullptr) const {
    {
      XFwdArgs<T> a = xfwd_args(U, 5, 0u, ws, gate);
      a.sig = dx.one;
      KScope ks("xfwd_copy", st);
      launch_xfwd<T, XF_COPY>(nx, ny, nz, a, st);
      check_launch("k_xfwd(state)");
    }
    yfwd(false, 5, 0, ws, st, true, gate);
    ColArgs<CE> a{};
    a.active = gate;
    a.act_all = 1;
    a.in = reinterpret_cast<cx<CE>*>(B(ws));
    a.in_fs = (b_fs()) / PW;
    a.in_bs = b_ys() / PW;
    a.in_ls = b_zs() / PW;
    a.out = reinterpret_cast<cx<CE>*>(S(ws));
    a.out_fs = s_fs() / PW;
    a.out_bs = mxp / PW;
    a.out_ls = s_zs() / PW;
    a.ncol = (mx + PW - 1) / PW;
    a.ncolw = mxp / PW;
    a.nc = 5;
    a.keep = az.keep();
    a.ikflag = 0;
    a.tw = dz.tw;
    a.sig = dz.one;
    a.kk = dz.k;
    KScope ks("z_fwd", st);
    launch_col<CE, CM_FWD_COPY>(nz, ay.nr, a, st);
    check_launch("k_col(z forward to S)");
  }
  // y inverse: B (compact) -> A (full)
  void yinv(int nc, void* ws, hipStream_t st, const int* gate = nullptr) const {
    ColArgs<CE> a{};
    a.active = gate;
    a.act_all = 1;
    a.in = reinterpret_cast<cx<CE>*>(B(ws));
    a.in_fs = (b_fs()) / PW;
    a.in_bs = b_zs() / PW;
    a.in_ls = b_ys() / PW;
    a.out = reinterpret_cast<cx<CE>*>(A(ws));
    a.out_fs = (a_fs()) / PW;
    a.out_bs = ((long long)ny * mxp) / PW;
    a.out_ls = (mxp) / PW;
    a.ncol = (mx + PW - 1) / PW;
    a.ncolw = mxp / PW;
    a.nc = nc;
    a.keep = ay.keep();
    a.ikflag = 0;
    a.tw = dy.tw;
    a.sig = dy.sig;
    a.kk = dy.k;
    a.cpb = PH_YINV_SPLIT ? -1 : 0;
    KScope ks("yinv", st);
    launch_col<CE, CM_INV>(ny, nz, a, st);
    check_launch("k_col(y inv)");
  }
  XInvArgs<T> xinv_args(void* ws, const int* gate = nullptr) const {
    XInvArgs<T> a{};
    a.active = gate;
    a.act_all = 1;
    a.in = A(ws);
    a.in_fs = a_fs();
    a.in_ss = (long long)ny * mxp;
    a.mx = mx;
    a.mxp = mxp;
    a.fs = ncell;
    a.ss = (long long)ny * nx;
    a.gamma = T(gamma);
    a.tw = dx.tw;
    return a;
  }

  // ---- public operations ----
  void spectral(const T* in, int nf, int axis, void* ws, hipStream_t st) const {
    // fields -> D (in A): transform, sigma*mask (and i*k along `axis` if >= 0)
    xfwd(XF_COPY, in, nf, axis == 0 ? 0x1Fu : 0u, ws, st);
    yfwd(false, nf, axis == 1, ws, st);
    zpass(false, nf, axis == 2, ws, st);
    yinv(nf, ws, st);
  }
  void deriv(int axis, int nf, const T* in, T* out, void* ws, hipStream_t st) const {
    spectral(in, nf, axis, ws, st);
    XInvArgs<T> a = xinv_args(ws);
    a.nf = nf;
    a.out = out;
    KScope ks("xinv_out", st);
    launch_xinv<T, 5, XI_OUT>(nx, ny, nz, a, st);
    check_launch("k_xinv");
  }
  void divergence(const T* U, void* ws, hipStream_t st) const {
    xfwd(XF_FLUX3, U, 5, 0u, ws, st);
    yfwd(true, 5, 0, ws, st);
    zpass(true, 5, 0, ws, st);
    yinv(5, ws, st);
  }
  void rhs(const T* U, T* out, void* ws, hipStream_t st) const {
    divergence(U, ws, st);
    XInvArgs<T> a = xinv_args(ws);
    a.nf = 5;
    a.out = out;
    launch_xinv<T, 5, XI_RHS>(nx, ny, nz, a, st);
    check_launch("k_xinv");
  }
  // RK4 (phrike/solver.py:1220-1230) then the post-step filter (1233-1234,
  // 1336).  `out` doubles as the stage accumulator; each stage boundary is
  // one fused k_xmid launch (x inverse + combine + next fluxes + x forward).
  // `bandlimited`: U has no content beyond the 2/3 cutoff (it is the output of
  // a previous step, whose RHS and filter are both masked).  FP32 plans then
  // apply the post-step filter in correction form, out = U' + F^-1((sx sy sz
  // - 1) F U') on the kept modes: the same operator on such states, but the
  // FP32 transforms carry only the small correction instead of the whole
  // state, whose round-trip rounding would otherwise accumulate step after
  // step.  FP64 plans always use the direct form (the reference's order).
  // `gate` (device int, or null): when gate[0] == 0 every kernel of the step
  // returns at once and out is left as it is (the device-resident run loop
  // keeps launching steps after its run has stopped)
  // cfl: also leave the new state's per-row wave-speed maxima in CflRows(ws)
  // (the last kernel takes them from the values it writes); max_rows() then
  // gives the max and dt without reading the state again
  // flux_in: the spectral buffer already holds the x-transformed fluxes of U
  // (left by the previous step's k_xend), so the first x pass is skipped;
  // flux_out: end with k_xend, which leaves those of out for the next step.
  //
  // With specacc the RK4 sum is kept in 3D spectral space instead: S starts
  // as the masked spectrum of U (spec_init; a chained step finds it in S,
  // left by the previous step's stage 4), each stage's z pass subtracts w D
  // from it (-D is the spectrum of k), and stage 4's z pass leaves
  // sigma_x sigma_y sigma_z (S - dt/6 D4) there and transforms it back: the
  // post-step filter of U + dt/6 (k1 + 2 k2 + 2 k3 + k4), since the k are
  // band-limited (F k = -D up to rounding).  The x kernels then form the
  // stage states U + c dt k with no accumulator, and the step has no
  // separate filter passes.
  // init_always: a step that computes its own fluxes (and spectrum) does so
  // even when gate stops it (a run loop's call then leaves them for its
  // state whether or not a step ran)
  void step(const T* U, T* out, double dt, const double* dt_dev, void* ws, hipStream_t st,
            bool bandlimited = false, const int* gate = nullptr, bool cfl = false, bool flux_in = false,
            bool flux_out = false, bool init_always = false) const {
    const int* igate = init_always ? nullptr : gate;
    if (specacc) {
      if (!flux_in) {
        spec_init(U, ws, st, igate);
        xfwd(XF_FLUX3, U, 5, 0u, ws, st, igate);
      }
      for (int s = 1; s <= 4; ++s) {
        yfwd(true, 5, 0, ws, st, false, gate);
        zpass_acc(s, dt, dt_dev, ws, st, gate);
        yinv(5, ws, st, gate);
        if (s == 4) break;
        XMidArgs<T> m{};
        m.inv = xinv_args(ws, gate);
        m.inv.nf = 5;
        m.inv.out = out;
        m.inv.un = U;
        m.inv.stage = s;
        m.inv.dt = dt;
        m.inv.dt_dev = dt_dev;
        m.inv.specacc = 1;
        m.fwd = xfwd_args(nullptr, 5, 0u, ws, gate);
        m.fwd_last = m.fwd;
        KScope ks("xmid_sa", st);
        launch_xmid<T, 5, XF_FLUX3>(nx, ny, nz, m, st);
        check_launch("k_xmid");
      }
      XInvArgs<T> a = xinv_args(ws, gate);
      a.nf = 5;
      a.out = out;
      a.cfl_rows = cfl ? CflRows(ws) : nullptr;
      if (flux_out) {
        XMidArgs<T> m{};
        m.inv = a;
        m.fwd = xfwd_args(nullptr, 5, 0u, ws, gate);
        KScope ks("xend_out", st);
        launch_xend<T, 5, XI_OUT>(nx, ny, nz, m, st);
        check_launch("k_xend");
      } else {
        KScope ks("xinv_out", st);
        launch_xinv<T, 5, XI_OUT>(nx, ny, nz, a, st);
        check_launch("k_xinv");
      }
      return;
    }
    const bool corr = bandlimited && sizeof(T) == 4;
    if (!flux_in) xfwd(XF_FLUX3, U, 5, 0u, ws, st, igate);
    for (int s = 1; s <= 4; ++s) {
      yfwd(true, 5, 0, ws, st, false, gate);
      zpass(true, 5, 0, ws, st, gate);
      yinv(5, ws, st, gate);
      XMidArgs<T> m{};
      m.inv = xinv_args(ws, gate);
      m.inv.nf = 5;
      m.inv.out = out;
      m.inv.un = U;
      m.inv.stage = s;
      m.inv.dt = dt;
      m.inv.dt_dev = dt_dev;
      m.inv.keep_state = corr ? 1 : 0;
      m.fwd = xfwd_args(nullptr, 5, 0u, ws, gate);
      m.fwd_last = xfwd_args(nullptr, 5, 0u, ws, gate);
      if (corr) m.fwd_last.sig = dx.one;
      KScope ks(s == 4 ? (corr ? "xmid_last_keep" : "xmid_last") : (s == 1 ? "xmid1" : "xmid23"), st);
      launch_xmid<T, 5, XF_FLUX3>(nx, ny, nz, m, st);
      check_launch("k_xmid");
    }
    // remaining passes of the post-step filter on the new state
    XInvArgs<T> a = xinv_args(ws, gate);
    a.nf = 5;
    a.out = out;
    a.cfl_rows = cfl ? CflRows(ws) : nullptr;
    if (corr) {
      yfwd(false, 5, 0, ws, st, true, gate);
      zcorr(5, ws, st, gate);
      yinv(5, ws, st, gate);
    } else {
      yfwd(false, 5, 0, ws, st, false, gate);
      zpass(false, 5, 0, ws, st, gate);
      yinv(5, ws, st, gate);
    }
    if (flux_out) {
      XMidArgs<T> m{};
      m.inv = a;
      m.fwd = xfwd_args(nullptr, 5, 0u, ws, gate);
      KScope ks(corr ? "xend_add" : "xend_out", st);
      if (corr) launch_xend<T, 5, XI_ADD>(nx, ny, nz, m, st);
      else launch_xend<T, 5, XI_OUT>(nx, ny, nz, m, st);
      check_launch("k_xend");
    } else {
      KScope ks(corr ? "xinv_add" : "xinv_out", st);
      if (corr) launch_xinv<T, 5, XI_ADD>(nx, ny, nz, a, st);
      else launch_xinv<T, 5, XI_OUT>(nx, ny, nz, a, st);
      check_launch("k_xinv");
    }
  }

  // Unnormalised forward transform of nf real fields to their half spectra
  // (nf, nz, ny, nx/2 + 1), interleaved complex; the inverse takes such half
  // spectra (Hermitian), scales by 1/(nx ny nz) and writes the real fields.
  // Only for a plan that keeps every mode and has no filter.
  void fft3d(int nf, const T* f, T* out, void* ws, hipStream_t st) const {
    if (!full || az.nr != nz || ay.nr != ny) throw Error(-2, "fft3d needs a plan without dealiasing or filter");
    XFwdArgs<T> a = xfwd_args(f, nf, 0u, ws);
    a.sig = unit_x;
    launch_xfwd<T, XF_COPY>(nx, ny, nz, a, st);
    check_launch("k_xfwd(fft3d)");
    yfwd(false, nf, 0, ws, st);
    zline(true, nf, ws, st);
    const long long rows = (long long)nf * nz * ny;
    hipLaunchKernelGGL((k_spec_out<T>), dim3((unsigned)((rows * mx + NT - 1) / NT)), dim3(NT), 0, st,
                       reinterpret_cast<const T*>(B(ws)), out, rows, mx, mxp);
    check_launch("k_spec_out");
  }
  void ifft3d(int nf, const T* in, T* f, void* ws, hipStream_t st) const {
    if (!full || az.nr != nz || ay.nr != ny) throw Error(-2, "ifft3d needs a plan without dealiasing or filter");
    const long long rows = (long long)nf * nz * ny;
    hipLaunchKernelGGL((k_spec_in<T>), dim3((unsigned)((rows * mxp + NT - 1) / NT)), dim3(NT), 0, st, in,
                       reinterpret_cast<T*>(B(ws)), rows, mx, mxp, T(1.0 / (double(nx) * double(ny) * double(nz))));
    check_launch("k_spec_in");
    zline(false, nf, ws, st);
    yinv(nf, ws, st);
    XInvArgs<T> a = xinv_args(ws);
    a.nf = nf;
    a.out = f;
    launch_xinv<T, 5, XI_OUT>(nx, ny, nz, a, st);
    check_launch("k_xinv(ifft3d)");
  }

  void max_speed(const T* U, double* out, double* dt_out, double cfl, void* ws, hipStream_t st,
                 const int* gate = nullptr) const {
    double* part = Part(ws);
    const int nb = (int)std::min<long long>(NPART, (ncell + NT - 1) / NT);
    hipLaunchKernelGGL((k_maxspeed_partial<T, 3>), dim3(nb, 1), dim3(NT), 0, st, U, ncell, 0ll, ncell, T(gamma), part,
                       gate);
    check_launch("k_maxspeed_partial");
    const double h = std::min(lx / nx, std::min(ly / ny, lz / nz));
    hipLaunchKernelGGL((k_max_final<3>), dim3(1), dim3(NT), 0, st, (const double*)part, nb, out, dt_out, cfl * h,
                       gate);
    check_launch("k_max_final");
  }
  // max and dt of the state written by the last step(..., cfl = true)
  void max_rows(double* out, double* dt_out, double cfl, void* ws, hipStream_t st, const int* gate = nullptr) const {
    double* part = Part(ws);
    const long long nrow = (long long)ny * nz;
    const int nb = (int)std::min<long long>(NPART, (nrow + NT - 1) / NT);
    hipLaunchKernelGGL(k_max_mid, dim3(nb, 1), dim3(NT), 0, st, (const double*)CflRows(ws), nrow, part);
    check_launch("k_max_mid");
    const double h = std::min(lx / nx, std::min(ly / ny, lz / nz));
    hipLaunchKernelGGL((k_max_final<3>), dim3(1), dim3(NT), 0, st, (const double*)part, nb, out, dt_out, cfl * h,
                       gate);
    check_launch("k_max_final(rows)");
  }
  void sums(const T* U, double* out, void* ws, hipStream_t st) const {
    double* part = Part(ws);
    const int nb = (int)std::min<long long>(NPART, (ncell + NT - 1) / NT);
    hipLaunchKernelGGL((k_sums_partial<T, 3>), dim3(nb, 1), dim3(NT), 0, st, U, ncell, 0ll, ncell, T(gamma), part);
    check_launch("k_sums_partial");
    hipLaunchKernelGGL((k_sums_final<5>), dim3(1), dim3(NT), 0, st, (const double*)part, nb, out);
    check_launch("k_sums_final");
  }
  // monitoring statistics (MON_* fields) of a state with NCOMP fields
  template <int NCOMP>
  void monitor(const T* U, double* out, void* ws, hipStream_t st) const {
    double* part = Part(ws);
    const int nb = (int)std::min<long long>(NPART, (ncell + NT - 1) / NT);
    hipLaunchKernelGGL((k_monitor_partial<T, NCOMP>), dim3(nb, 1), dim3(NT), 0, st, U, ncell, 0ll, ncell, T(gamma),
                       part);
    check_launch("k_monitor_partial");
    hipLaunchKernelGGL((k_monitor_final<T>), dim3(1), dim3(NT), 0, st, (const double*)part, nb, U, 0ll, out);
    check_launch("k_monitor_final");
  }
};

// ---------------------------------------------------------------------------
// batched 2D Fourier x Fourier plan: `batch` independent runs of one shape,
// state (batch, 4, ny, nx)
// ---------------------------------------------------------------------------
template <typename T>
struct Plan2D : PlanBase {
  int nx, ny, batch;
  double lx, ly, gamma;
  AxisTables ax, ay;
  using CE = typename ColElem<T>::type;  // column-pass element
  static constexpr int PW = ColElem<T>::PW;
  DevAxis<T> dx;
  DevAxis<CE> dy;
  int mx, mxp;
  long long ncell;
  CE* xcol_k = nullptr;  // kx per column of the y pass (zero past the kept modes)
  size_t offA, offUs, offPart, offCfl, wsBytes;
  static constexpr int NPART = 256;

  Plan2D(int nx_, int ny_, double lx_, double ly_, bool dealias, bool filt, int p, double alpha,
         double gamma_, int batch_)
      : nx(nx_), ny(ny_), batch(batch_), lx(lx_), ly(ly_), gamma(gamma_) {
    dtype = sizeof(T) == 8 ? 0 : 1;
    if (!supported_len(nx) || !supported_len(ny))
      throw Error(-2, "grid sizes must be powers of two in [8, 1024]");
    if (batch < 1 || batch > 65535) throw Error(-2, "batch must be in [1, 65535]");
    ax = AxisTables(nx, lx, dealias, filt, p, alpha, true);
    ay = AxisTables(ny, ly, dealias, filt, p, alpha, false);
    mx = ax.nr;
    const int align = sizeof(T) == 8 ? 8 : 16;  // 128-byte rows (cache lines)
    mxp = (mx + align - 1) / align * align;
    ncell = (long long)nx * ny;
    dx.upload(ax, 1.0 / (double(nx) * double(ny)));
    dy.upload(ay, 1.0);
    {
      const int ncol = (mx + PW - 1) / PW, ncolw = mxp / PW;
      std::vector<CE> hk(ncolw, CE(0.0));
      for (int c = 0; c < ncol; ++c) {
        if constexpr (PW == 2) {
          hk[c] = CE(float(ax.k[2 * c]), float(2 * c + 1 < mx ? ax.k[2 * c + 1] : 0.0));
        } else {
          hk[c] = CE(ax.k[c]);
        }
      }
      hip_check(hipMalloc(&xcol_k, sizeof(CE) * ncolw), "hipMalloc(column k)");
      hip_check(hipMemcpy(xcol_k, hk.data(), sizeof(CE) * ncolw, hipMemcpyHostToDevice), "copy column k");
    }
    const size_t cs = sizeof(cx<T>);
    auto al = [](size_t v) { return (v + 255) / 256 * 256; };
    offA = 0;
    offUs = al(8ull * batch * ny * mxp * cs);
    offPart = al(offUs + 4ull * batch * ncell * sizeof(T));
    offCfl = al(offPart + sizeof(double) * (size_t)NPART * MON_NRED * batch);
    wsBytes = al(offCfl + sizeof(double) * (size_t)batch * ny);
  }
  ~Plan2D() override {
    if (xcol_k) (void)hipFree(xcol_k);
    dx.release();
    dy.release();
  }
  cx<T>* A(void* ws) const { return reinterpret_cast<cx<T>*>((char*)ws + offA); }
  T* Us(void* ws) const { return reinterpret_cast<T*>((char*)ws + offUs); }
  double* Part(void* ws) const { return reinterpret_cast<double*>((char*)ws + offPart); }
  // per-row wave-speed maxima of the states the last step wrote ([run][y])
  double* CflRows(void* ws) const { return reinterpret_cast<double*>((char*)ws + offCfl); }

  void check_nb(int nb) const {
    if (nb < 1 || nb > batch) throw Error(-2, "nbatch must be in [1, plan batch]");
  }
  // workspace bytes step() touches for nb runs: only the spectral buffer A
  // (the fused step keeps stage states in registers; no reductions)
  size_t step_scratch_bytes(int nb) const { return 8ull * size_t(nb) * size_t(ny) * size_t(mxp) * sizeof(cx<T>); }
  XFwdArgs<T> xfwd_args(unsigned ikmask, void* ws) const {
    XFwdArgs<T> a{};
    a.in_fs = ncell;
    a.in_ss = 4 * ncell;
    a.out = A(ws);
    a.out_fs = (long long)ny * mxp;
    a.out_ss = 8ll * ny * mxp;
    a.nf = 4;
    a.mx = mx;
    a.mxp = mxp;
    a.ikmask = ikmask;
    a.tw = dx.tw;
    a.sig = dx.sig;
    a.kk = dx.k;
    a.gamma = T(gamma);
    return a;
  }
  void xfwd(int mode, const T* in, int nb, unsigned ikmask, void* ws, hipStream_t st,
            const int* active = nullptr) const {
    XFwdArgs<T> a{};
    a.active = active;
    a.in = in;
    a.in_fs = ncell;
    a.in_ss = 4 * ncell;
    a.out = A(ws);
    a.out_fs = (long long)ny * mxp;
    a.out_ss = 8ll * ny * mxp;
    a.nf = 4;
    a.mx = mx;
    a.mxp = mxp;
    a.ikmask = ikmask;
    a.tw = dx.tw;
    a.sig = dx.sig;
    a.kk = dx.k;
    a.gamma = T(gamma);
    if (mode == XF_FLUX2) launch_xfwd<T, XF_FLUX2>(nx, ny, nb, a, st);
    else launch_xfwd<T, XF_COPY>(nx, ny, nb, a, st);
    check_launch("k_xfwd(2d)");
  }
  void ypass(bool div, int ikflag, int nb, void* ws, hipStream_t st, const int* active = nullptr) const {
    ColArgs<CE> a{};
    a.active = active;
    a.in = reinterpret_cast<cx<CE>*>(A(ws));
    a.in_fs = ((long long)ny * mxp) / PW;
    a.in_bs = (8ll * ny * mxp) / PW;
    a.in_ls = (mxp) / PW;
    a.out = reinterpret_cast<cx<CE>*>(A(ws));
    a.out_fs = a.in_fs;
    a.out_bs = a.in_bs;
    a.out_ls = a.in_ls;
    a.ncol = (mx + PW - 1) / PW;
    a.ncolw = mxp / PW;
    a.nc = 4;
    a.keep = ay.keep();
    a.ikflag = ikflag;
    a.tw = dy.tw;
    a.sig = dy.sig;
    a.kk = dy.k;
    a.colk = xcol_k;
    // the symmetric 2D layout (7 fluxes, no i kx in the x pass)
    if (div && PH_2D_SYM) launch_col<CE, CM_FWDINV_DIV2S>(ny, nb, a, st);
    else if (div) launch_col<CE, CM_FWDINV_DIV>(ny, nb, a, st);
    else launch_col<CE, CM_FWDINV_COPY>(ny, nb, a, st);
    check_launch("k_col(2d y)");
  }
  XInvArgs<T> xinv_args(void* ws) const {
    XInvArgs<T> a{};
    a.in = A(ws);
    a.in_fs = (long long)ny * mxp;
    a.in_ss = 8ll * ny * mxp;
    a.mx = mx;
    a.mxp = mxp;
    a.fs = ncell;
    a.ss = 4 * ncell;
    a.gamma = T(gamma);
    a.rho_min = T(1e-8);
    a.p_min = T(1e-10);
    a.tw = dx.tw;
    return a;
  }
  void deriv(int axis, int nb, const T* in, T* out, const int* active, void* ws, hipStream_t st) const {
    xfwd(XF_COPY, in, nb, axis == 0 ? 0xFu : 0u, ws, st, active);
    ypass(false, axis == 1, nb, ws, st, active);
    XInvArgs<T> a = xinv_args(ws);
    a.nf = 4;
    a.out = out;
    a.active = active;
    launch_xinv<T, 4, XI_OUT>(nx, ny, nb, a, st);
    check_launch("k_xinv(2d)");
  }
  void rhs(int nb, const T* U, T* out, void* ws, hipStream_t st) const {
    xfwd(XF_FLUX2, U, nb, 0xFu, ws, st);
    ypass(true, 0, nb, ws, st);
    XInvArgs<T> a = xinv_args(ws);
    a.nf = 4;
    a.out = out;
    launch_xinv<T, 4, XI_RHS>(nx, ny, nb, a, st);
    check_launch("k_xinv(2d rhs)");
  }
  // RK4 with clamp (phrike/solver.py:502-513) and the filter when asked
  // (cadence decided by the caller, phrike/solver.py:645-647); stage
  // boundaries fused as in the 3D step
  // cfl: also leave the per-row wave-speed maxima of the new states in
  // CflRows(ws) for max_rows()
  // flux_in / flux_out as in Plan3D::step (chained steps of ph_run2d; the
  // filter must be on)
  void step(int nb, const T* U, T* out, double dt, const double* dt_dev, const int* active,
            bool filter, void* ws, hipStream_t st, bool cfl = false, bool flux_in = false,
            bool flux_out = false) const {
    if ((flux_in || flux_out) && !filter) throw Error(-9, "chained 2D steps need the post-step filter");
    if (!flux_in) xfwd(XF_FLUX2, U, nb, 0xFu, ws, st, active);
    for (int s = 1; s <= 4; ++s) {
      ypass(true, 0, nb, ws, st, active);
      XMidArgs<T> m{};
      m.inv = xinv_args(ws);
      m.inv.nf = 4;
      m.inv.out = out;
      m.inv.un = U;
      m.inv.stage = s;
      m.inv.dt = dt;
      m.inv.dt_dev = dt_dev;
      m.inv.dt_per_slab = 1;
      m.inv.active = active;
      if (s == 4 && !filter) {
        m.inv.cfl_rows = cfl ? CflRows(ws) : nullptr;
        launch_xinv<T, 4, XI_STAGE2>(nx, ny, nb, m.inv, st);
        check_launch("k_xinv(2d stage 4)");
        return;
      }
      m.fwd = xfwd_args(0xFu, ws);
      m.fwd_last = xfwd_args(0u, ws);
      launch_xmid<T, 4, XF_FLUX2>(nx, ny, nb, m, st);
      check_launch("k_xmid(2d)");
    }
    ypass(false, 0, nb, ws, st, active);
    XInvArgs<T> a = xinv_args(ws);
    a.nf = 4;
    a.out = out;
    a.active = active;
    a.cfl_rows = cfl ? CflRows(ws) : nullptr;
    if (flux_out) {
      XMidArgs<T> m{};
      m.inv = a;
      m.fwd = xfwd_args(0xFu, ws);
      launch_xend<T, 4, XI_OUT>(nx, ny, nb, m, st);
      check_launch("k_xend(2d)");
    } else {
      launch_xinv<T, 4, XI_OUT>(nx, ny, nb, a, st);
      check_launch("k_xinv(2d filter)");
    }
  }
  // max and dt per run of the states written by the last step(..., cfl = true)
  void max_rows(int nb, double* out, double* dt_out, double cfl, void* ws, hipStream_t st,
                const int* active = nullptr) const {
    const double h = std::min(lx / nx, ly / ny);
    hipLaunchKernelGGL((k_max_final<2>), dim3(nb), dim3(NT), 0, st, (const double*)CflRows(ws), ny, out, dt_out,
                       cfl * h, active);
    check_launch("k_max_final(2d rows)");
  }
  void max_speed(int nb, const T* U, double* out, double* dt_out, double cfl, void* ws,
                 hipStream_t st, const int* active = nullptr) const {
    double* part = Part(ws);
    const int npb = (int)std::min<long long>(NPART, (ncell + NT - 1) / NT);
    hipLaunchKernelGGL((k_maxspeed_partial<T, 2>), dim3(npb, nb), dim3(NT), 0, st, U, ncell, 4 * ncell, ncell, T(gamma), part,
                       active);
    check_launch("k_maxspeed_partial(2d)");
    const double h = std::min(lx / nx, ly / ny);
    hipLaunchKernelGGL((k_max_final<2>), dim3(nb), dim3(NT), 0, st, (const double*)part, npb, out, dt_out, cfl * h,
                       active);
    check_launch("k_max_final(2d)");
  }
  void sums(int nb, const T* U, double* out, void* ws, hipStream_t st) const {
    double* part = Part(ws);
    const int npb = (int)std::min<long long>(NPART, (ncell + NT - 1) / NT);
    hipLaunchKernelGGL((k_sums_partial<T, 2>), dim3(npb, nb), dim3(NT), 0, st, U, ncell, 4 * ncell, ncell, T(gamma), part);
    check_launch("k_sums_partial(2d)");
    hipLaunchKernelGGL((k_sums_final<4>), dim3(nb), dim3(NT), 0, st, (const double*)part, npb, out);
    check_launch("k_sums_final(2d)");
  }
  void monitor(int nb, const T* U, double* out, void* ws, hipStream_t st) const {
    double* part = Part(ws);
    const int npb = (int)std::min<long long>(NPART, (ncell + NT - 1) / NT);
    hipLaunchKernelGGL((k_monitor_partial<T, 4>), dim3(npb, nb), dim3(NT), 0, st, U, ncell, 4 * ncell, ncell,
                       T(gamma), part);
    check_launch("k_monitor_partial(2d)");
    hipLaunchKernelGGL((k_monitor_final<T>), dim3(nb), dim3(NT), 0, st, (const double*)part, npb, U, 4 * ncell, out);
    check_launch("k_monitor_final(2d)");
  }
};

}  // namespace ph
