// C ABI (include/phrike_hip.h): argument checks, error strings, dispatch on
// the plan's element type.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "../../include/phrike_hip.h"
#include "pipeline.hpp"

namespace ph {
extern template struct Plan3D<double>;
extern template struct Plan3D<float>;
extern template struct Plan2D<double>;
extern template struct Plan2D<float>;
}  // namespace ph

namespace {
thread_local std::string g_err;

template <typename F>
int guard(F&& f) {
  try {
    f();
    return PH_OK;
  } catch (const ph::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return PH_ERR_INTERNAL;
  } catch (...) {
    g_err = "unknown error";
    return PH_ERR_INTERNAL;
  }
}

void need(bool ok, const char* msg) {
  if (!ok) throw ph::Error(PH_ERR_ARG, msg);
}

ph::PlanBase* base3(ph_plan3d_t p) {
  need(p != nullptr, "null plan");
  return reinterpret_cast<ph::PlanBase*>(p);
}
ph::PlanBase* base2(ph_plan2d_t p) {
  need(p != nullptr, "null plan");
  return reinterpret_cast<ph::PlanBase*>(p);
}

// the plan's dynamic type decides (a plan of another kind is an argument error)
template <typename F>
void with3(ph_plan3d_t p, F&& f) {
  ph::PlanBase* b = base3(p);
  if (auto* d = dynamic_cast<ph::Plan3D<double>*>(b)) f(*d);
  else if (auto* s = dynamic_cast<ph::Plan3D<float>*>(b)) f(*s);
  else throw ph::Error(PH_ERR_ARG, "not a 3D plan");
}
template <typename F>
void with2(ph_plan2d_t p, F&& f) {
  ph::PlanBase* b = base2(p);
  if (auto* d = dynamic_cast<ph::Plan2D<double>*>(b)) f(*d);
  else if (auto* s = dynamic_cast<ph::Plan2D<float>*>(b)) f(*s);
  else throw ph::Error(PH_ERR_ARG, "not a 2D plan");
}

hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Scratch memory of ph_step2d_noalloc: a file-scope device array, so calls
// into that entry point allocate nothing.  Each GPU the library runs on holds
// its own copy (at its own address) until the process ends.  In host code the
// array's name is only a stand-in for the symbol: the device address comes
// from hipGetSymbolAddress, resolved once per device with that device current
// (at 2D plan creation, or at the entry point's first call on a device).
constexpr size_t kScratchBytes = size_t(32) << 20;  // 32 MiB
__device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

void* scratch_ptr() {
  int dev = -1;
  ph::hip_check(hipGetDevice(&dev), "hipGetDevice");
  need(dev >= 0 && dev < kMaxDevices, "device index out of range for the static scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    ph::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    need(p != nullptr, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

namespace ph {
// shared with the other translation units' entry points (ph_last_error)
void set_last_error(const char* msg) { g_err = msg; }
}  // namespace ph

extern "C" {

int ph_abi_version(void) { return PH_ABI_VERSION; }

int ph_timing_enable(int on) {
  return guard([&] {
    ph::KTimer& k = ph::ktimer();
    if (!on) k.collect();
    k.on = on != 0;
  });
}

int ph_timing_collect(char* buf, size_t len) {
  return guard([&] {
    need(buf != nullptr && len > 0, "null buffer");
    ph::KTimer& k = ph::ktimer();
    k.collect();
    std::string out;
    for (auto& kv : k.acc)
      out += kv.first + " " + std::to_string(kv.second.second) + " " + std::to_string(kv.second.first) + "\n";
    k.acc.clear();
    std::strncpy(buf, out.c_str(), len - 1);
    buf[len - 1] = 0;
  });
}

const char* ph_last_error(void) { return g_err.c_str(); }

int ph_plan3d_create(ph_plan3d_t* plan, int nx, int ny, int nz, double lx, double ly, double lz,
                     int dtype, int dealias, int filter_enabled, int filter_p, double filter_alpha,
                     double gamma) {
  return guard([&] {
    need(plan != nullptr, "null plan pointer");
    need(dtype == PH_F64 || dtype == PH_F32, "dtype must be PH_F64 or PH_F32");
    need(lx > 0 && ly > 0 && lz > 0, "domain lengths must be positive");
    need(gamma > 1.0, "gamma must exceed 1");
    ph::PlanBase* p;
    if (dtype == PH_F64)
      p = new ph::Plan3D<double>(nx, ny, nz, lx, ly, lz, dealias != 0, filter_enabled != 0, filter_p,
                                 filter_alpha, gamma);
    else
      p = new ph::Plan3D<float>(nx, ny, nz, lx, ly, lz, dealias != 0, filter_enabled != 0, filter_p,
                                filter_alpha, gamma);
    *plan = reinterpret_cast<ph_plan3d_t>(p);
  });
}

int ph_plan3d_destroy(ph_plan3d_t plan) {
  return guard([&] { delete base3(plan); });
}

int ph_plan3d_workspace_bytes(ph_plan3d_t plan, size_t* bytes) {
  return guard([&] {
    need(bytes != nullptr, "null output");
    with3(plan, [&](auto& p) { *bytes = p.wsBytes; });
  });
}

int ph_rhs3d(ph_plan3d_t plan, const void* U, void* rhs, void* ws, void* stream) {
  return guard([&] {
    need(U && rhs && ws, "null buffer");
    ph::chain_registry().forget(ws);
    with3(plan, [&](auto& p) {
      using R = std::conditional_t<std::is_same_v<std::decay_t<decltype(p)>, ph::Plan3D<double>>, double, float>;
      p.rhs(static_cast<const R*>(U), static_cast<R*>(rhs), ws, S(stream));
    });
  });
}

int ph_deriv3d(ph_plan3d_t plan, int axis, int nfields, const void* f, void* out, void* ws,
               void* stream) {
  return guard([&] {
    need(f && out && ws, "null buffer");
    ph::chain_registry().forget(ws);
    need(axis >= -1 && axis <= 2, "axis must be -1, 0, 1 or 2");
    need(nfields >= 1 && nfields <= 5, "nfields must be in [1, 5]");
    with3(plan, [&](auto& p) {
      using R = std::conditional_t<std::is_same_v<std::decay_t<decltype(p)>, ph::Plan3D<double>>, double, float>;
      p.deriv(axis, nfields, static_cast<const R*>(f), static_cast<R*>(out), ws, S(stream));
    });
  });
}

int ph_step3d_ex(ph_plan3d_t plan, const void* U, void* out, double dt, const double* dt_dev, int flags,
                 void* ws, void* stream) {
  return guard([&] {
    need(U && out && ws, "null buffer");
    need((flags & ~(PH_STEP_BANDLIMITED | PH_STEP_CHAIN | PH_STEP_CHAINED)) == 0, "unknown step flags");
    with3(plan, [&](auto& p) {
      using R = std::conditional_t<std::is_same_v<std::decay_t<decltype(p)>, ph::Plan3D<double>>, double, float>;
      const size_t n = 5ull * size_t(p.ncell) * sizeof(R);
      need(!ph::ranges_overlap(U, n, out, n), "U and out must not overlap");
      auto& reg = ph::chain_registry();
      const bool chain = (flags & PH_STEP_CHAIN) != 0;
      // the fluxes of U are in the workspace only if this plan's previous
      // chained step wrote U there and nothing has written the workspace since
      const bool flux_in = (flags & PH_STEP_CHAINED) != 0 && reg.holds(ws, &p, U);
      reg.forget(ws);
      p.step(static_cast<const R*>(U), static_cast<R*>(out), dt, dt_dev, ws, S(stream),
             (flags & PH_STEP_BANDLIMITED) != 0, nullptr, false, flux_in, chain);
      if (chain) reg.set(ws, &p, out);
    });
  });
}

int ph_step3d(ph_plan3d_t plan, const void* U, void* out, double dt, const double* dt_dev, void* ws,
              void* stream) {
  return ph_step3d_ex(plan, U, out, dt, dt_dev, 0, ws, stream);
}

int ph_run3d(ph_plan3d_t plan, void* Ua, void* Ub, int nsteps, double* rs_dev, int* rf_dev, double cfl, int flags,
             void* ws, void* stream) {
  return guard([&] {
    need(Ua && Ub && rs_dev && rf_dev && ws, "null buffer");
    need(nsteps >= 1, "nsteps must be >= 1");
    need((flags & ~(PH_STEP_BANDLIMITED | PH_STEP_CHAIN | PH_STEP_CHAINED)) == 0, "unknown step flags");
    with3(plan, [&](auto& p) {
      using R = std::conditional_t<std::is_same_v<std::decay_t<decltype(p)>, ph::Plan3D<double>>, double, float>;
      const size_t n = 5ull * size_t(p.ncell) * sizeof(R);
      need(!ph::ranges_overlap(Ua, n, Ub, n), "Ua and Ub must not overlap");
      hipStream_t st = S(stream);
      auto& reg = ph::chain_registry();
      const bool chain = (flags & PH_STEP_CHAIN) != 0;
      // PH_STEP_CHAINED: Ua is the state a previous chained call on this
      // workspace left, and nothing has written the workspace since
      const bool chained_in = (flags & PH_STEP_CHAINED) != 0 && reg.holds(ws, &p, Ua);
      reg.forget(ws);
      for (int k = 0; k < nsteps; ++k) {
        R* src = static_cast<R*>(k % 2 ? Ub : Ua);
        R* dst = static_cast<R*>(k % 2 ? Ua : Ub);
        hipLaunchKernelGGL(ph::k_run_clip, dim3(1), dim3(ph::NT), 0, st, 1, rs_dev, (const int*)rf_dev);
        ph::check_launch("k_run_clip");
        // steps after the first take a previous step's (band-limited) output
        // the step's last kernel leaves the new state's per-row wave-speed
        // maxima behind, so the next dt needs no second read of the state
        // steps chain within the call: each but the last leaves the fluxes of
        // its output for the next (a step of a stopped run does nothing, and
        // so do all after it)
        // with PH_STEP_CHAIN the last step also leaves them for the next
        // call, and a first step that computes its own does so even if the
        // run has stopped, so the workspace holds those of the call's final
        // state (Ua or Ub, by the number of steps taken) in every case
        p.step(src, dst, 0.0, rs_dev + ph::RS_DT, ws, st, k > 0 || (flags & PH_STEP_BANDLIMITED) != 0, rf_dev,
               true, k > 0 || chained_in, k + 1 < nsteps || chain, k == 0 && chain);
        p.max_rows(nullptr, rs_dev + ph::RS_DTNEXT, cfl, ws, st, rf_dev);
        hipLaunchKernelGGL(ph::k_run_advance, dim3(1), dim3(ph::NT), 0, st, 1, rs_dev, rf_dev);
        ph::check_launch("k_run_advance");
      }
      if (chain) reg.set(ws, &p, Ua, Ub);
    });
  });
}

int ph_max_wave_speed3d(ph_plan3d_t plan, const void* U, double* max_dev, double* dt_dev,
                        double cfl, void* ws, void* stream) {
  return guard([&] {
    need(U && ws && (max_dev || dt_dev), "null buffer");
    with3(plan, [&](auto& p) {
      using R = std::conditional_t<std::is_same_v<std::decay_t<decltype(p)>, ph::Plan3D<double>>, double, float>;
      p.max_speed(static_cast<const R*>(U), max_dev, dt_dev, cfl, ws, S(stream));
    });
  });
}

int ph_conserved_sums3d(ph_plan3d_t plan, const void* U, double* out5_dev, void* ws, void* stream) {
  return guard([&] {
    need(U && out5_dev && ws, "null buffer");
    with3(plan, [&](auto& p) {
      using R = std::conditional_t<std::is_same_v<std::decay_t<decltype(p)>, ph::Plan3D<double>>, double, float>;
      p.sums(static_cast<const R*>(U), out5_dev, ws, S(stream));
    });
  });
}

int ph_fft3d(ph_plan3d_t plan, int nfields, const void* f, void* spec, void* ws, void* stream) {
  return guard([&] {
    need(f && spec && ws, "null buffer");
    ph::chain_registry().forget(ws);
    need(nfields >= 1 && nfields <= 5, "nfields must be in [1, 5]");
    with3(plan, [&](auto& p) {
      using R = std::conditional_t<std::is_same_v<std::decay_t<decltype(p)>, ph::Plan3D<double>>, double, float>;
      p.fft3d(nfields, static_cast<const R*>(f), static_cast<R*>(spec), ws, S(stream));
    });
  });
}

int ph_ifft3d(ph_plan3d_t plan, int nfields, const void* spec, void* f, void* ws, void* stream) {
  return guard([&] {
    need(f && spec && ws, "null buffer");
    ph::chain_registry().forget(ws);
    need(nfields >= 1 && nfields <= 5, "nfields must be in [1, 5]");
    with3(plan, [&](auto& p) {
      using R = std::conditional_t<std::is_same_v<std::decay_t<decltype(p)>, ph::Plan3D<double>>, double, float>;
      p.ifft3d(nfields, static_cast<const R*>(spec), static_cast<R*>(f), ws, S(stream));
    });
  });
}

int ph_monitor3d(ph_plan3d_t plan, const void* U, double* out_dev, void* ws, void* stream) {
  return guard([&] {
    need(U && out_dev && ws, "null buffer");
    with3(plan, [&](auto& p) {
      using R = std::conditional_t<std::is_same_v<std::decay_t<decltype(p)>, ph::Plan3D<double>>, double, float>;
      p.template monitor<5>(static_cast<const R*>(U), out_dev, ws, S(stream));
    });
  });
}

int ph_plan2d_create(ph_plan2d_t* plan, int nx, int ny, double lx, double ly, int dtype, int dealias,
                     int filter_enabled, int filter_p, double filter_alpha, double gamma, int batch) {
  return guard([&] {
    need(plan != nullptr, "null plan pointer");
    need(dtype == PH_F64 || dtype == PH_F32, "dtype must be PH_F64 or PH_F32");
    need(lx > 0 && ly > 0, "domain lengths must be positive");
    need(gamma > 1.0, "gamma must exceed 1");
    ph::PlanBase* p;
    if (dtype == PH_F64)
      p = new ph::Plan2D<double>(nx, ny, lx, ly, dealias != 0, filter_enabled != 0, filter_p,
                                 filter_alpha, gamma, batch);
    else
      p = new ph::Plan2D<float>(nx, ny, lx, ly, dealias != 0, filter_enabled != 0, filter_p,
                                filter_alpha, gamma, batch);
    *plan = reinterpret_cast<ph_plan2d_t>(p);
    (void)scratch_ptr();  // set-up of the no-allocation step on this device
  });
}

int ph_plan2d_destroy(ph_plan2d_t plan) {
  return guard([&] { delete base2(plan); });
}

int ph_plan2d_workspace_bytes(ph_plan2d_t plan, size_t* bytes) {
  return guard([&] {
    need(bytes != nullptr, "null output");
    with2(plan, [&](auto& p) { *bytes = p.wsBytes; });
  });
}

#define PH_R2 \
  using R = std::conditional_t<std::is_same_v<std::decay_t<decltype(p)>, ph::Plan2D<double>>, double, float>

int ph_rhs2d(ph_plan2d_t plan, int nbatch, const void* U, void* rhs, void* ws, void* stream) {
  return guard([&] {
    need(U && rhs && ws, "null buffer");
    ph::chain_registry().forget(ws);
    with2(plan, [&](auto& p) {
      PH_R2;
      p.check_nb(nbatch);
      p.rhs(nbatch, static_cast<const R*>(U), static_cast<R*>(rhs), ws, S(stream));
    });
  });
}

int ph_deriv2d(ph_plan2d_t plan, int nbatch, int axis, const void* f, void* out, void* ws,
               void* stream) {
  return guard([&] {
    need(f && out && ws, "null buffer");
    ph::chain_registry().forget(ws);
    need(axis >= -1 && axis <= 1, "axis must be -1, 0 or 1");
    with2(plan, [&](auto& p) {
      PH_R2;
      p.check_nb(nbatch);
      p.deriv(axis, nbatch, static_cast<const R*>(f), static_cast<R*>(out), nullptr, ws, S(stream));
    });
  });
}

int ph_step2d(ph_plan2d_t plan, int nbatch, const void* U, void* out, double dt, const double* dt_dev,
              const int* active_dev, int apply_filter, void* ws, void* stream) {
  return guard([&] {
    need(U && out && ws, "null buffer");
    ph::chain_registry().forget(ws);
    with2(plan, [&](auto& p) {
      PH_R2;
      p.check_nb(nbatch);
      const size_t n = 4ull * size_t(nbatch) * size_t(p.ncell) * sizeof(R);
      need(!ph::ranges_overlap(U, n, out, n), "U and out must not overlap");
      p.step(nbatch, static_cast<const R*>(U), static_cast<R*>(out), dt, dt_dev, active_dev,
             apply_filter != 0, ws, S(stream));
    });
  });
}

int ph_step2d_noalloc(ph_plan2d_t plan, int nbatch, const void* U, void* out, double dt, const double* dt_dev,
                      const int* active_dev, int apply_filter, void* stream) {
  return guard([&] {
    need(U && out, "null buffer");
    with2(plan, [&](auto& p) {
      PH_R2;
      p.check_nb(nbatch);
      const size_t n = 4ull * size_t(nbatch) * size_t(p.ncell) * sizeof(R);
      need(!ph::ranges_overlap(U, n, out, n), "U and out must not overlap");
      const size_t want = p.step_scratch_bytes(nbatch);
      if (want > kScratchBytes)
        throw ph::Error(PH_ERR_ARG, "ph_step2d_noalloc: " + std::to_string(nbatch) + " run(s) of " +
                                        std::to_string(p.nx) + "x" + std::to_string(p.ny) + " need " +
                                        std::to_string(want) + " bytes of scratch; the static scratch holds " +
                                        std::to_string(kScratchBytes) + " (use ph_step2d with a workspace)");
      p.step(nbatch, static_cast<const R*>(U), static_cast<R*>(out), dt, dt_dev, active_dev, apply_filter != 0,
             scratch_ptr(), S(stream));
    });
  });
}

int ph_max_wave_speed2d(ph_plan2d_t plan, int nbatch, const void* U, double* max_dev, double* dt_dev,
                        double cfl, const int* active_dev, void* ws, void* stream) {
  return guard([&] {
    need(U && ws && (max_dev || dt_dev), "null buffer");
    with2(plan, [&](auto& p) {
      PH_R2;
      p.check_nb(nbatch);
      p.max_speed(nbatch, static_cast<const R*>(U), max_dev, dt_dev, cfl, ws, S(stream), active_dev);
    });
  });
}

int ph_run2d(ph_plan2d_t plan, int nbatch, void* Ua, void* Ub, int nsteps, double* rs_dev, int* rf_dev, double cfl,
             void* ws, void* stream) {
  return guard([&] {
    need(Ua && Ub && rs_dev && rf_dev && ws, "null buffer");
    ph::chain_registry().forget(ws);
    need(nsteps >= 1, "nsteps must be >= 1");
    with2(plan, [&](auto& p) {
      PH_R2;
      p.check_nb(nbatch);
      const size_t n = 4ull * size_t(nbatch) * size_t(p.ncell) * sizeof(R);
      need(!ph::ranges_overlap(Ua, n, Ub, n), "Ua and Ub must not overlap");
      hipStream_t st = S(stream);
      const int nblk = (nbatch + ph::NT - 1) / ph::NT;
      const int* act = rf_dev + ph::RF_ACTIVE * nbatch;
      for (int k = 0; k < nsteps; ++k) {
        const R* src = static_cast<const R*>(k % 2 ? Ub : Ua);
        R* dst = static_cast<R*>(k % 2 ? Ua : Ub);
        hipLaunchKernelGGL(ph::k_run_clip, dim3(nblk), dim3(ph::NT), 0, st, nbatch, rs_dev, (const int*)rf_dev);
        ph::check_launch("k_run_clip");
        // steps chain within the call (each but the last forms the next
        // step's fluxes; runs that stop stay stopped for the rest of the call)
        p.step(nbatch, src, dst, 0.0, rs_dev + ph::RS_DT * nbatch, act, true, ws, st, true, k > 0,
               k + 1 < nsteps);
        p.max_rows(nbatch, nullptr, rs_dev + ph::RS_DTNEXT * nbatch, cfl, ws, st, act);
        hipLaunchKernelGGL(ph::k_run_advance, dim3(nblk), dim3(ph::NT), 0, st, nbatch, rs_dev, rf_dev);
        ph::check_launch("k_run_advance");
      }
    });
  });
}

int ph_conserved_sums2d(ph_plan2d_t plan, int nbatch, const void* U, double* out4_dev, void* ws,
                        void* stream) {
  return guard([&] {
    need(U && out4_dev && ws, "null buffer");
    with2(plan, [&](auto& p) {
      PH_R2;
      p.check_nb(nbatch);
      p.sums(nbatch, static_cast<const R*>(U), out4_dev, ws, S(stream));
    });
  });
}

int ph_monitor2d(ph_plan2d_t plan, int nbatch, const void* U, double* out_dev, void* ws, void* stream) {
  return guard([&] {
    need(U && out_dev && ws, "null buffer");
    with2(plan, [&](auto& p) {
      PH_R2;
      p.check_nb(nbatch);
      p.monitor(nbatch, static_cast<const R*>(U), out_dev, ws, S(stream));
    });
  });
}

int ph_runs_clip(int nbatch, const double* t_dev, const double* t_end_dev, double* dt_dev,
                 const int* active_dev, void* stream) {
  return guard([&] {
    need(nbatch >= 1 && t_dev && t_end_dev && dt_dev && active_dev, "bad arguments");
    hipLaunchKernelGGL(ph::k_runs_clip, dim3((nbatch + ph::NT - 1) / ph::NT), dim3(ph::NT), 0, S(stream),
                       nbatch, t_dev, t_end_dev, dt_dev, active_dev);
    ph::check_launch("k_runs_clip");
  });
}

int ph_runs_advance(int nbatch, double* t_dev, const double* t_end_dev, const double* dt_dev,
                    int* active_dev, int* steps_dev, void* stream) {
  return guard([&] {
    need(nbatch >= 1 && t_dev && t_end_dev && dt_dev && active_dev && steps_dev, "bad arguments");
    hipLaunchKernelGGL(ph::k_runs_advance, dim3((nbatch + ph::NT - 1) / ph::NT), dim3(ph::NT), 0, S(stream),
                       nbatch, t_dev, t_end_dev, dt_dev, active_dev, steps_dev);
    ph::check_launch("k_runs_advance");
  });
}

int ph_gll_apply(int mode, int ny, int nx, int nbatch, const void* D, const void* Dhi, const void* Dlo,
                 const void* F, const void* add, void* out, void* stream) {
  return guard([&] {
    // every output row reads every row of its F line: no in-place apply
    if (F && out && ny > 0 && nx > 0 && nbatch > 0) {
      const size_t es = (mode & ~PH_GLL_OFFSET) == PH_GLL_SPLIT_F32 ? 4 : 8;
      const size_t n = size_t(nbatch) * size_t(ny) * size_t(nx) * es;
      const char* f0 = static_cast<const char*>(F);
      const char* o0 = static_cast<const char*>(out);
      if (f0 < o0 + n && o0 < f0 + n) throw ph::Error(PH_ERR_ARG, "F and out must not overlap");
    }
    ph::gll_apply(mode, ny, nx, nbatch, D, Dhi, Dlo, F, add, out, S(stream));
  });
}

}  // extern "C"
