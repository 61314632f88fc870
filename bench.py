"""Benchmark of the hip backend's 3D RK4 step (one GPU).

Workload (default): 3D Taylor-Green vortex, 512^3, FP64, Fourier basis on
[0, 2pi)^3, gamma = 1.4, 2/3 dealiasing, exponential filter alpha = 36,
order 16, RK4 with the post-step filter, dt from CFL 0.4 frozen at t = 0.
A step is one RK4 step of the whole grid (4 fused RHS evaluations, stage
combines, post-step filter).  Inputs are resident in HBM before timing.

Prints ONE JSON line:
  value      Mcell-updates/s = N^3 * steps / wall time
  roofline   the dominant kernel's algorithmic bytes per launch over its
             average launch duration (HIP events in the timed region)
             against 8 TB/s; plus the whole step against the pencil-pipeline
             byte model C(N) s N^3 (hbm_frac_step)
  cpu_baseline  the oracle (numpy/scipy restatement of the reference's
             NumPy path) timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "3D Euler RK4 Mcell-updates/s + %HBM peak; FP32 vs FP64 err; ensemble p99"


def byte_model(n: int):
    """Per-launch algorithmic bytes of each pipeline kernel in units of
    s*N^3 (s = element size), and the per-step total C(N)."""
    k = n // 3
    a = 2 * (k + 1) / n
    b = 2 * (k + 1) * (2 * k + 1) / n ** 2
    c = b * (2 * k + 1) / n  # one field's kept modes of all three axes
    # the x kernels write the 12 distinct Euler fluxes (the momentum flux
    # tensor is symmetric), which the y pass reads
    per = {
        "xfwd_flux": 5 + 12 * a, "yfwd_div": 12 * a + 10 * b, "z_div": 15 * b,
        "yinv": 5 * b + 5 * a, "xinv_stage1": 5 * a + 15, "xinv_stage23": 5 * a + 20,
        "xinv_stage4": 5 * a + 10, "xfwd_copy": 5 + 5 * a, "yfwd_copy": 5 * a + 5 * b,
        "z_copy": 10 * b, "xinv_out": 5 * a + 5,
        # fused stage boundaries: D in, U in, accumulator, next-stage spectra out
        "xmid1": 5 * a + 5 + 5 + 12 * a, "xmid23": 5 * a + 5 + 10 + 12 * a,
        # (stage 4 reads the accumulator, which carries U, and not U itself)
        "xmid_last": 5 * a + 5 + 5 * a,
        # FP32 filter in correction form: stage 4 also writes the new state,
        # unit-factor y pass, (sx sy sz - 1) z pass, x inverse added to it
        "xmid_last_keep": 5 * a + 5 + 5 * a + 5, "yfwd_unit": 5 * a + 5 * b,
        "z_corr": 10 * b, "xinv_add": 5 * a + 5 + 5,
        # chained steps: the filter's x inverse also forms the next step's
        # fluxes and their x transform (the state is not read again)
        "xend_out": 5 * a + 5 + 12 * a, "xend_add": 5 * a + 5 + 5 + 12 * a,
        # the RK4 sum in spectral space: stage boundaries without an
        # accumulator; z passes that also read and write the sum S (stage 4:
        # the filtered new state's spectrum); an unchained step's first pass
        # of S (after the x forward and unit-factor y pass of the state)
        "xmid_sa": 5 * a + 5 + 12 * a, "z_acc": 15 * b + 10 * c, "z_fwd": 5 * b + 5 * c,
    }
    # per-step model of the unfused pencil pipeline (stage states written and
    # re-read), the official figure for hbm_frac_step
    C = 4 * (5 + 40 * a + 30 * b) + 70 + 10 + 20 * a + 20 * b
    return per, C


def kernel_symbol(tag: str, n: int, prec: str) -> str:
    """The HIP kernel (as rocprofv3 names it) that a timing tag launches
    (x kernels instantiated for the 2/3 rule: kept-mode flag `true`)."""
    T = "double" if prec == "double" else "float"
    C = "double" if prec == "double" else "ph::f2"
    return {"xfwd_flux": f"k_xfwd<{T}, {n}, true, 0>", "xfwd_copy": f"k_xfwd<{T}, {n}, true, 2>",
            "yfwd_div": f"k_col<{C}, {n}, 0>", "yfwd_copy": f"k_col<{C}, {n}, 1>",
            "z_div": f"k_col<{C}, {n}, 2>", "z_copy": f"k_col<{C}, {n}, 3>",
            "yinv": f"k_col<{C}, {n}, 4>", "xinv_out": f"k_xinv<{T}, {n}, true, 5, 0>",
            # one kernel for every stage boundary (stage 4 branches at run time)
            "xmid1": f"k_xmid<{T}, {n}, true, 5, 0>", "xmid23": f"k_xmid<{T}, {n}, true, 5, 0>",
            "xmid_last": f"k_xmid<{T}, {n}, true, 5, 0>", "xmid_last_keep": f"k_xmid<{T}, {n}, true, 5, 0>",
            "yfwd_unit": f"k_col<{C}, {n}, 1>", "z_corr": f"k_col<{C}, {n}, 5>",
            "xinv_add": f"k_xinv<{T}, {n}, true, 5, 4>",
            "xend_out": f"k_xend<{T}, {n}, true, 5, 0>", "xend_add": f"k_xend<{T}, {n}, true, 5, 4>",
            "xmid_sa": f"k_xmid<{T}, {n}, true, 5, 0>", "z_acc": f"k_col<{C}, {n}, 7>",
            "z_fwd": f"k_col<{C}, {n}, 1>"}.get(tag, tag)


# launches of each timing tag in one RK4 step of the fused 3D pipeline, in
# a chain of steps (each step's last kernel forms the next step's first x
# pass, and leaves the filtered state's spectrum for the next step's RK4 sum)
STEP_LAUNCHES = {"yfwd_div": 4, "z_acc": 4, "yinv": 4, "xmid_sa": 3, "xend_out": 1}
# the library built with the RK4 accumulator in the x kernels (PH_SPECACC=0):
# FP64, and FP32 steps on a previous step's output (filter in correction form)
STEP_LAUNCHES_XACC = {"yfwd_div": 4, "z_div": 4, "yinv": 5, "xmid1": 1, "xmid23": 2,
                      "xmid_last": 1, "yfwd_copy": 1, "z_copy": 1, "xend_out": 1}
STEP_LAUNCHES_CORR = {"yfwd_div": 4, "z_div": 4, "yinv": 5, "xmid1": 1, "xmid23": 2,
                      "xmid_last_keep": 1, "yfwd_unit": 1, "z_corr": 1, "xend_add": 1}


def step_launches(prec: str, specacc: bool = True):
    if specacc:
        return STEP_LAUNCHES
    return STEP_LAUNCHES_CORR if prec == "single" else STEP_LAUNCHES_XACC


def fused_step_units(n: int, prec: str = "double", specacc: bool = True) -> float:
    """Bytes per step (units of s*N^3) that the fused pipeline moves."""
    per, _ = byte_model(n)
    return sum(per[t] * k for t, k in step_launches(prec, specacc).items())


def cpu_baseline(n: int, steps: int):
    """Time the oracle's step on the host: one untimed warm-up step, then
    `steps` timed steps of TGV n^3 FP64, and its compute_dt (vectorised
    numpy; the reference itself, without numba, takes the CFL maximum in a
    pure-Python loop over the cells)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import spectral_np as sn
    from phrike_amd.initial_conditions import taylor_green_3d

    cores = min(16, sn.default_workers())
    g = sn.Grid3DOracle(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
                        filter_params={"enabled": True, "p": 16, "alpha": 36.0}, fft_workers=cores)
    s = sn.Solver3DOracle(g, sn.Euler3DOracle(1.4), cfl=0.4)
    U = taylor_green_3d(n, n, n)
    t0 = time.perf_counter()
    dt = s.compute_dt(U)
    t_dt = time.perf_counter() - t0
    U = s.step(U, dt)  # warm-up step (FFT plans, allocations), not timed
    t0 = time.perf_counter()
    for _ in range(steps):
        U = s.step(U, dt)
    el = time.perf_counter() - t0
    return {"value": n ** 3 * steps / el / 1e6, "unit": "Mcell-updates/s", "cores": cores,
            "cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "compute_dt_s": round(t_dt, 3),
            "kind": "port",
            "sample": f"oracle (numpy/{sn.FFT_LIBRARY} restatement of the reference NumPy path) "
                      f"TGV {n}^3 FP64, {steps} timed RK4 step(s) incl. post-step filter after 1 warm-up step, "
                      f"{cores} FFT workers, {el:.1f} s"}


def precision_study(n: int, steps: int, dev):
    """Decaying turbulence n^3 (seeded random-phase solenoidal velocity,
    E(k) ~ k^4 exp(-2(k/4)^2), rms Mach 0.3, rho = 1 + 1% noise, uniform p):
    the same initial state stepped `steps` times in FP64 and in FP32 (FP64
    tables, sums and time), dt frozen from CFL 0.4 at t = 0.  Reports per-field
    relative L2 of FP32 against FP64, conservation drifts, and the speedup."""
    import torch
    from phrike_amd.equations import EulerEquations3D
    from phrike_amd.grid import Grid3D
    from phrike_amd.initial_conditions import decaying_turbulence_3d
    from phrike_amd.solver import SpectralSolver3D

    U0 = torch.as_tensor(decaying_turbulence_3d(n, seed=1234, density_seed=4321), device=dev)
    res, finals, sums = {}, {}, {}
    dt = None
    for prec in ("double", "single"):
        g = Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
                   filter_params={"enabled": True, "p": 16, "alpha": 36.0}, backend="hip", precision=prec)
        s = SpectralSolver3D(g, EulerEquations3D(1.4), scheme="rk4", cfl=0.4)
        U = U0.to(g.dtype).clone()  # never alias U0: the loop writes both buffers
        if dt is None:
            dt = s.compute_dt(U)
        s0 = s.conserved_sums(U)
        out = torch.empty_like(U)
        s.step_into(U, out, dt)  # warm-up step, not timed, not counted
        U = U0.to(g.dtype).clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.step_into(U, out, dt)
            U, out = out, U
        torch.cuda.synchronize()
        res[prec] = (time.perf_counter() - t0) / steps * 1e3
        s1 = s.conserved_sums(U)
        sums[prec] = (s0, s1)
        finals[prec] = U.double()
        del g, s, out
    a, b = finals["single"], finals["double"]
    per = [float(torch.linalg.vector_norm(a[f] - b[f]) / torch.linalg.vector_norm(b[f])) for f in range(5)]
    mom = max(float(torch.linalg.vector_norm(b[f])) for f in (1, 2, 3))
    grp = [per[0], *[float(torch.linalg.vector_norm(a[f] - b[f])) / mom for f in (1, 2, 3)], per[4]]
    s0, s1 = sums["double"]
    f0, f1 = sums["single"]
    return {"workload": f"decaying turbulence {n}^3, rms Mach 0.3, {steps} RK4 steps, dt frozen",
            "ms_per_step_f64": round(res["double"], 3), "ms_per_step_f32": round(res["single"], 3),
            "speedup_f32": round(res["double"] / res["single"], 3),
            "rel_l2_f32_vs_f64": [float(f"{e:.3e}") for e in per],
            "rel_l2_f32_vs_f64_momentum_scale": [float(f"{e:.3e}") for e in grp],
            "mass_drift_f64": float(f"{(s1[0] - s0[0]) / s0[0]:.3e}"),
            "energy_drift_f64": float(f"{(s1[4] - s0[4]) / s0[4]:.3e}"),
            "mass_drift_f32": float(f"{(f1[0] - f0[0]) / f0[0]:.3e}"),
            "energy_drift_f32": float(f"{(f1[4] - f0[4]) / f0[4]:.3e}")}


def torch_backend_step(n: int, dev):
    """The reference's own GPU path (its `torch` backend: phrike/grid.py:1424-1518,
    1705-1726; phrike/equations.py:582-607; phrike/solver.py:1177-1234), restated
    with torch.fft (the vendor FFT library), FP64, on [0, 2pi)^3 with n^3
    points, 2/3 mask and the exponential filter (alpha 36, p 16): full complex
    fftn of all 5 fields per derivative, full 3D mask and sigma tensors,
    ifftn(...).real, RK4 with state-sized temporaries, post-step filter.
    Returns step(U, dt)."""
    import torch

    gam = 1.4
    L = 2 * np.pi
    d = L / n
    m = np.fft.fftfreq(n, 1.0 / n)
    k = 2 * np.pi * (m * (1.0 / (n * d)))
    mask1 = (np.abs(m) <= n // 3).astype(np.float64)
    sig1 = np.exp(-36.0 * (np.abs(m) / max(n // 2, 1)) ** 16)
    t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev)
    mask = t(mask1[:, None, None] * mask1[None, :, None] * mask1[None, None, :])
    sigma = t((sig1[:, None, None] * sig1[None, :, None]) * sig1[None, None, :])
    kt = t(k).to(torch.complex128)
    ikx, iky, ikz = 1j * kt[None, None, :], 1j * kt[None, :, None], 1j * kt[:, None, None]
    dims = (-3, -2, -1)

    def masked(F):
        F *= mask
        F *= sigma
        return F

    def deriv(f, ik):
        return torch.fft.ifftn(masked(torch.fft.fftn(f, dim=dims)) * ik, dim=dims).real

    def rhs(U):
        rho = U[0]
        ux, uy, uz = U[1] / rho, U[2] / rho, U[3] / rho
        p = (gam - 1.0) * (U[4] - 0.5 * rho * (ux * ux + uy * uy + uz * uz))
        mx, my, mz = rho * ux, rho * uy, rho * uz
        h = U[4] + p
        Fx = torch.stack([mx, mx * ux + p, mx * uy, mx * uz, h * ux])
        Fy = torch.stack([my, my * ux, my * uy + p, my * uz, h * uy])
        Fz = torch.stack([mz, mz * ux, mz * uy, mz * uz + p, h * uz])
        return -(deriv(Fx, ikx) + deriv(Fy, iky) + deriv(Fz, ikz))

    def step(U, dt):
        k1 = rhs(U)
        k2 = rhs(U + 0.5 * dt * k1)
        k3 = rhs(U + 0.5 * dt * k2)
        k4 = rhs(U + dt * k3)
        Un = U + (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
        return torch.fft.ifftn(masked(torch.fft.fftn(Un, dim=dims)), dim=dims).real

    return step


def torch_backend_reference(n: int, steps: int, dev, dt: float, hip_state=None):
    """torch_backend_step on TGV n^3 as a second GPU reference point, timed
    over `steps` steps after one warm-up step; with `hip_state` (this
    backend's result after the same steps from the same start) the relative
    L2 difference is reported too."""
    import torch
    from phrike_amd.initial_conditions import taylor_green_3d_torch

    step = torch_backend_step(n, dev)
    U0 = taylor_green_3d_torch(n, dev, torch.float64)
    U = step(U0, dt)  # warm-up (FFT plans)
    torch.cuda.synchronize()
    U = U0
    t0 = time.perf_counter()
    for _ in range(steps):
        U = step(U, dt)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    out = {"workload": f"TGV {n}^3 FP64, the reference's torch backend arithmetic on torch.fft, "
                       f"{steps} RK4 step(s) after 1 warm-up",
           "ms_per_step": round(el * 1e3, 2), "mcell_updates_per_s": round(n ** 3 / el / 1e6, 2)}
    if hip_state is not None:
        out["rel_l2_vs_hip"] = [float(f"{float(torch.linalg.vector_norm(U[f] - hip_state[f]) / max(float(torch.linalg.vector_norm(hip_state[f])), 1e-300)):.3e}")
                                for f in range(5)]
    return out


def hybrid_study(nx: int, nys, steps: int, dev, ensembles=(16, 64), check=200):
    """2D Kelvin-Helmholtz on the hybrid grid (Fourier x on [0,1), LGL y on
    [-1/2, 1/2]): shear layers at y = +-1/4 of width 0.05, rho 1 -> 2, p = 2.5,
    v = 0.01 sin(4 pi x) + 1e-4 N(0,1) (seed 7), walls at y = +-1/2; RK4
    steps with dt frozen from CFL 0.3 at t = 0 (CFL 0.4 with this filter
    turns unstable at the walls after ~850 steps at 256 x 128), post-step
    x/y filters.  Three arithmetic forms of D.Fy: FP64 matrix cores on FP64
    state ("f64"), split precision (FP32 matrix cores on hi/lo parts of D
    and of the data) on FP64 state ("split_f64state") and on FP32 state
    ("split_f32state").

    * one run per size: time per step, D.U kernel rate (useful flops
      2 Ny^2 * 4 Nx per apply), relative L2 per field against FP64 after
      `check` and after `steps` steps, next to an FP64 run whose initial state
      is perturbed by 1e-15 relative (the KHI's own growth of a rounding-size
      difference);
    * ensembles of runs (batched, the D.U apply then takes the tiled
      matrix-core kernel): time per step of the batch and per run-step."""
    import torch
    from phrike_amd import _lib
    from phrike_amd.equations import EulerEquations2D
    from phrike_amd.hybrid import HybridGrid2D, HybridSolver2D
    from phrike_amd.initial_conditions import khi_2d_gll

    fp = {"enabled": True, "p": 16, "alpha": 36.0}
    forms = (("f64", "double", "f64"), ("split_f64state", "double", "split"), ("split_f32state", "single", "split"))
    out = {}

    def kernel_rate(sol, U0, dt, ny, nb=1):
        _lib.timing_collect()
        _lib.timing_enable(True)
        sol.advance(U0, 20, dt=dt)
        torch.cuda.synchronize()
        _lib.timing_enable(False)
        kt = _lib.timing_collect()
        if "h_gll_d" not in kt:
            return None
        cnt, tms = kt["h_gll_d"]
        flops = 2.0 * ny * ny * 4 * nx * nb
        return {"avg_us": round(tms / cnt * 1e3, 2), "useful_TFLOPs": round(flops / (tms / cnt / 1e3) / 1e12, 3)}

    for ny in nys:
        snaps, ms, dk = {}, {}, {}
        dt = None
        runs = list(forms) + [("f64_perturbed", "double", "f64")]
        for name, prec, gll in runs:
            g = HybridGrid2D(nx, ny, filter_params=fp, precision=prec, gll=gll, torch_device=str(dev))
            sol = HybridSolver2D(g, EulerEquations2D(1.4), cfl=0.3)
            U0 = torch.as_tensor(khi_2d_gll(nx, g.y_nodes, seed=7), device=dev)
            if name == "f64_perturbed":
                gen = torch.Generator(device="cpu").manual_seed(11)
                U0 = U0 * (1.0 + 1e-15 * torch.randn(U0.shape, generator=gen, dtype=torch.float64).to(dev))
            U0 = U0.to(g.dtype)
            if dt is None:
                dt = sol.compute_dt(U0)
            sol.advance(U0, 2, dt=dt)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            U1 = sol.advance(U0, check, dt=dt)
            U2 = sol.advance(U1, steps - check, dt=dt)
            torch.cuda.synchronize()
            ms[name] = (time.perf_counter() - t0) / steps * 1e3
            if name != "f64_perturbed":
                dk[name] = kernel_rate(sol, U0, dt, ny)
                # the same steps replayed from a captured HIP graph of 50 steps
                sol.advance(U0, 50, dt=dt, graph_steps=50)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                Ug = sol.advance(U0, steps, dt=dt, graph_steps=50)
                torch.cuda.synchronize()
                ms[name + "_graph"] = (time.perf_counter() - t0) / steps * 1e3
                if not torch.equal(Ug.to(torch.float64), U2.to(torch.float64)):
                    raise RuntimeError(f"hybrid {name}: graph-replayed steps differ from eager steps")
            snaps[name] = (U1.double(), U2.double())
            del g, sol
        ref = snaps["f64"]

        def errs(k, i):
            a, r = snaps[k][i], ref[i]
            return [float(f"{float(torch.linalg.vector_norm(a[f] - r[f]) / torch.linalg.vector_norm(r[f])):.3e}")
                    for f in range(4)]

        others = ("split_f64state", "split_f32state", "f64_perturbed")
        out[f"{nx}x{ny}"] = {
            "ms_per_step": {k: round(v, 4) for k, v in ms.items()},
            "finite": bool(torch.isfinite(ref[1]).all()),
            f"rel_l2_vs_f64_at_{check}": {k: errs(k, 0) for k in others},
            f"rel_l2_vs_f64_at_{steps}": {k: errs(k, 1) for k in others},
            "gll_d_kernel": dk}
    # ensembles of the first size
    ny = nys[0]
    ens = {}
    for nb in ensembles:
        row = {}
        for name, prec, gll in forms:
            g = HybridGrid2D(nx, ny, filter_params=fp, precision=prec, gll=gll, batch=nb, torch_device=str(dev))
            sol = HybridSolver2D(g, EulerEquations2D(1.4), cfl=0.3)
            U0 = torch.stack([torch.as_tensor(khi_2d_gll(nx, g.y_nodes, seed=7 + i), device=dev)
                              for i in range(nb)]).to(g.dtype)
            dt = sol.compute_dt(U0[0])
            sol.advance(U0, 2, dt=dt)
            torch.cuda.synchronize()
            k = 50
            t0 = time.perf_counter()
            sol.advance(U0, k, dt=dt)
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) / k * 1e3
            row[name] = {"ms_per_step": round(t, 4), "us_per_run_step": round(t * 1e3 / nb, 2),
                         "gll_d_kernel": kernel_rate(sol, U0, dt, ny, nb)}
            del g, sol, U0
            torch.cuda.empty_cache()
        ens[str(nb)] = row
    out[f"ensemble_{nx}x{ny}"] = ens
    return {"workload": f"KHI 2D hybrid Fourier-x ({nx}) x LGL-y with walls, {steps} RK4 steps, "
                        f"dt frozen, CFL 0.3; ensembles of {list(ensembles)} runs",
            **out}


def mhd_study(n: int, steps: int, dev):
    """3D ideal MHD (SpectralSolverMHD3D, reference defaults: projection after
    every substage, filter p=16 alpha=36, gamma 5/3, CFL 0.3) on a band-limited
    random state n^3 with a non-solenoidal B (seeded; the first projection
    removes its longitudinal part): time per step after one warm-up step in
    FP64 and FP32 (the solver's steps take the fused form, its input being
    solenoidal), max |div B| and relative mass / energy drift (FP64 sums)
    after the timed steps, then per-kernel times from two more steps; and the
    step with a projection after every substage (ph_step_mhd3d) timed on the
    same state, with its distance from the fused step."""
    import torch
    from phrike_amd import _lib
    from phrike_amd.grid import Grid3D
    from phrike_amd.initial_conditions import smooth_random_state_3d
    from phrike_amd.mhd import MHDEquations3D, SpectralSolverMHD3D

    H = smooth_random_state_3d(n, n, n, seed=31, kmax=6)
    Bn = smooth_random_state_3d(n, n, n, seed=32, kmax=5)
    B = np.stack([0.3 * Bn[1] / Bn[0], 0.3 * Bn[2] / Bn[0], 1.0 + 0.3 * Bn[3] / Bn[0]])
    U0 = np.concatenate([H, B])
    U0[4] = H[4] + 0.5 * (B ** 2).sum(axis=0)
    del H, Bn, B
    out = {"workload": f"ideal MHD {n}^3, band-limited random state (|m| <= 6), B0 = (0, 0, 1) + 30% "
                       f"random, {steps} RK4 steps; fused stages and one projection per step "
                       f"(PH_STEP_SOLENOIDAL) beside a projection after every substage"}
    for prec in ("double", "single"):
        g = Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
                   filter_params={"enabled": True, "p": 16, "alpha": 36.0}, backend="hip", precision=prec)
        g.mhd_config = {}
        s = SpectralSolverMHD3D(g, MHDEquations3D(5.0 / 3.0), cfl=0.3)
        U = s._project_B(torch.as_tensor(U0, device=dev).to(g.dtype))
        dt = s.compute_dt(U)
        c0 = s.conserved_quantities(U)
        buf = torch.empty_like(U)
        s.step_into(U, buf, dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.step_into(U, buf, dt)
            U, buf = buf, U
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / steps
        c1 = s.conserved_quantities(U)
        # per-kernel times from a separate pass (events around every launch)
        ksteps = 2
        _lib.timing_collect()
        _lib.timing_enable(True)
        for _ in range(ksteps):
            s.step_into(U, buf, dt)
            U, buf = buf, U
        torch.cuda.synchronize()
        _lib.timing_enable(False)
        kt = _lib.timing_collect()
        # the same steps with a projection after every substage
        ws, st = _lib.ptr(s._ws), s._st()
        V, vb = U.clone(), torch.empty_like(U)
        _lib.call("ph_step_mhd3d", s._plan, _lib.ptr(V), _lib.ptr(vb), dt, ws, st)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            _lib.call("ph_step_mhd3d", s._plan, _lib.ptr(V), _lib.ptr(vb), dt, ws, st)
            V, vb = vb, V
        torch.cuda.synchronize()
        el_sub = (time.perf_counter() - t0) / steps
        W = U.clone()
        s._mark_solenoidal(W)
        Wf = s.step(W, dt)
        Ws = torch.empty_like(W)
        _lib.call("ph_step_mhd3d", s._plan, _lib.ptr(W), _lib.ptr(Ws), dt, ws, st)
        fd = ((Wf - Ws).double().norm() / Ws.double().norm()).item()
        del V, vb, W, Wf, Ws
        out[prec] = {"ms_per_step": round(el * 1e3, 3), "mcell_updates_per_s": round(n ** 3 / el / 1e6, 1),
                     "ms_per_step_substage_projection": round(el_sub * 1e3, 3),
                     "fused_vs_substage_projection_rel_l2": float(f"{fd:.3e}"),
                     "max_div_b": float(f"{s.max_div_b(U):.3e}"),
                     "mass_drift": float(f"{(c1['mass'] - c0['mass']) / c0['mass']:.3e}"),
                     "energy_drift": float(f"{(c1['energy'] - c0['energy']) / c0['energy']:.3e}"),
                     "kernels_ms_per_step": {k: round(ms / ksteps, 3) for k, (cnt, ms) in sorted(kt.items())}}
        del g, s, U, buf
        torch.cuda.empty_cache()
    out["speedup_f32"] = round(out["double"]["ms_per_step"] / out["single"]["ms_per_step"], 3)
    return out


def gll_saturation(dev, runs: int = 64, nx: int = 256, nys=(128, 256), reps: int = 20):
    """The D.U kernel (ph_gll_apply) over a batch of `runs` hybrid-KHI-shaped
    states at once (4 fields x runs, Ny x nx each): the matrix cores at a size
    that fills the GPU, as an ensemble of hybrid runs would present it.
    Useful flops 2 Ny^2 nx per field and run; split modes issue 2 (FP32 data)
    or 3 (FP64 data) FP32 MFMA products per useful one.  HIP events around
    `reps` launches on the current stream."""
    import torch
    from phrike_amd.basis import LegendreLobattoBasis1D

    out = {}
    for ny in nys:
        b = LegendreLobattoBasis1D(ny, 1.0, torch_device=str(dev))
        F64 = torch.randn(4 * runs, ny, nx, dtype=torch.float64, device=dev)
        F32 = F64.float()
        flops = 2.0 * ny * ny * nx * 4 * runs
        row = {}
        for name, F, mode, peak in (("f64", F64, "f64", 78.6e12), ("split_f64io", F64, "split", 157.3e12),
                                    ("split_f32", F32, "split", 157.3e12)):
            o = torch.empty_like(F)
            b.apply(F, mode=mode, out=o)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                b.apply(F, mode=mode, out=o)
            e1.record()
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1) / reps / 1e3
            issued = flops * (1 if mode == "f64" else (3 if F.dtype == torch.float64 else 2))
            row[name] = {"avg_us": round(t * 1e6, 1), "useful_TFLOPs": round(flops / t / 1e12, 2),
                         "mfma_TFLOPs": round(issued / t / 1e12, 2),
                         "mfma_frac": round(issued / t / peak, 3)}
        out[f"ny{ny}"] = row
    return {"workload": f"D.U over {runs} runs x 4 fields, Ny x {nx}, {reps} launches; peak FP64 "
                        "matrix 78.6, FP32 matrix 157.3 TFLOP/s", **out}


def hbm_reference(dev, gib: int = 2, reps: int = 5):
    """What plain torch kernels reach on this GPU over `gib` GiB buffers: a
    device copy (bytes read + written), a read-only sum and a write-only fill;
    HIP events around `reps` launches each.  The achievable rate a mixed
    read/write pencil kernel can be compared with, beside the 8 TB/s peak."""
    import torch

    a = torch.ones(gib << 27, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    nbytes = a.numel() * 8
    out = {}
    for name, fn, moved in (("copy", lambda: b.copy_(a), 2 * nbytes), ("read", lambda: a.sum(), nbytes),
                            ("write", lambda: b.fill_(2.0), nbytes)):
        fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[f"{name}_GBs"] = round(moved * reps / (e0.elapsed_time(e1) / 1e3) / 1e9, 1)
    del a, b
    torch.cuda.empty_cache()
    return {"workload": f"torch copy_/sum/fill_ over {gib} GiB FP64 buffers, {reps} launches each", **out}


def _ensemble_summary(res, wall, sizes):
    from phrike_amd.serving import latency_summary
    summ = latency_summary(res, wall)
    per = {}
    for n in sizes:
        lat = [r.latency for r in res if r.n == n]
        if lat:
            per[str(n)] = {"count": len(lat), "p50_s": round(float(np.percentile(lat, 50)), 3),
                           "p99_s": round(float(np.percentile(lat, 99)), 3)}
    out = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in summ.items()}
    out["per_size"] = per
    # the 2D pencil-pipeline byte model (khi2d_words, FP32) over the wall
    # time, against 8 TB/s
    moved = 0.0
    for r in res:
        moved += r.steps * khi2d_words(r.n) * 4 * r.n ** 2
    out["model_GBs"] = round(moved / wall / 1e9, 1)
    out["hbm_frac_model"] = round(moved / wall / HBM_PEAK, 4)
    return out


def khi2d_words(n: int) -> float:
    """Words moved per cell and run-step by the chained batched 2D step with
    the filter every step, C2(N) = 40 + 104 a (a = kept x columns as a
    fraction of N).  Real fields: stage 1 reads U and writes the RK4
    accumulator (8), stages 2-3 read U and the accumulator and write it (12
    each), stage 4 reads the accumulator (4), the last x pass writes the new
    state (4).  Kept-column spectra: per stage the y pass reads the 7 fluxes
    and writes 4 fields (11 a), stages 1-3 read 4 and write 7 (11 a each),
    stage 4 reads 4 and writes the state's 4 for the filter (8 a), the filter's
    y pass reads and writes 4 (8 a), the last x pass reads 4 and writes the
    next step's 7 fluxes (11 a)."""
    a = 2 * (n // 3 + 1) / n
    return 40 + 104 * a


def khi2d_batch_study(dev, cases=((1024, 32), (512, 128), (256, 256), (128, 512)), precisions=("single", "double"),
                      steps: int = 32, launches: int = 3):
    """The batched 2D step at fixed batch sizes, as the ensemble server runs
    it: periodic KHI runs (seeds 0..B-1, filter every step, CFL 0.4), one
    ph_run2d call advancing every run by `steps` chained steps with its dt
    from the CFL maxima the previous step left (no host round trip), timed
    over `launches` calls after one warm-up call.  Reports ms per batched
    step, Mcell-updates/s and the byte model C2(N) against 8 TB/s."""
    from phrike_amd import _lib
    from phrike_amd.equations import EulerEquations2D
    from phrike_amd.grid import Grid2D
    from phrike_amd.initial_conditions import khi_2d_periodic
    from phrike_amd.solver import SpectralSolver2D
    import torch
    fp = {"enabled": True, "p": 16, "alpha": 36.0, "interval": 1}
    out = {"workload": f"KHI 2D periodic, B runs of N^2 per batch, filter every step, CFL 0.4, {steps} chained "
                       f"steps per ph_run2d call, {launches} timed calls after 1 warm-up"}
    for prec in precisions:
        es = 4 if prec == "single" else 8
        for n, nb in cases:
            print(f"[bench] khi2d batch {n}^2 x {nb} {prec}", file=sys.stderr, flush=True)
            g = Grid2D(n, n, 1.0, 1.0, dealias=True, filter_params=fp, backend="hip", precision=prec, batch=nb,
                       torch_device=str(dev))
            s = SpectralSolver2D(g, EulerEquations2D(1.4), cfl=0.4)
            U0 = np.stack([khi_2d_periodic(n, n, seed=i) for i in range(min(nb, 4))])
            U = [torch.as_tensor(U0, dtype=g.dtype, device=dev).repeat((nb + 3) // 4, 1, 1, 1)[:nb].contiguous(),
                 torch.empty((nb, 4, n, n), dtype=g.dtype, device=dev)]
            rs = torch.zeros(_lib.PH_RS_N * nb, dtype=torch.float64, device=dev)
            rf = torch.zeros(_lib.PH_RF_N * nb, dtype=torch.int32, device=dev)
            rs[_lib.PH_RS_TEND * nb:(_lib.PH_RS_TEND + 1) * nb] = 1e30
            rs[_lib.PH_RS_TSTOP * nb:(_lib.PH_RS_TSTOP + 1) * nb] = float("inf")
            rf[_lib.PH_RF_ACTIVE * nb:(_lib.PH_RF_ACTIVE + 1) * nb] = 1
            plan, ws, st = s._ctx()
            _lib.call("ph_max_wave_speed2d", plan, nb, _lib.ptr(U[0]), None,
                      _lib.ptr(rs[_lib.PH_RS_DT * nb:]), 0.4, None, ws, st)

            def launch():
                _lib.call("ph_run2d", plan, nb, _lib.ptr(U[0]), _lib.ptr(U[1]), steps, _lib.ptr(rs), _lib.ptr(rf),
                          0.4, ws, st)

            launch()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(launches):
                launch()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / (launches * steps)
            cells = nb * n * n
            gbs = khi2d_words(n) * es * cells / (ms * 1e-3) / 1e9
            done = int(rf[_lib.PH_RF_STEPS * nb:(_lib.PH_RF_STEPS + 1) * nb].min())
            finite = bool(torch.isfinite(U[0]).all())
            out[f"{n}x{nb}_{prec}"] = {"ms_per_step": round(ms, 4), "mcell_updates_per_s": round(cells / ms / 1e3, 1),
                                       "model_GBs": round(gbs, 1), "hbm_frac_model": round(gbs * 1e9 / HBM_PEAK, 4),
                                       "steps_done_min": done, "finite": finite}
            del g, s, U, rs, rf
            torch.cuda.empty_cache()
    return out


def ensemble_study(rate: float, window: float, budget_gb: float, seed: int = 0, sustain_windows=(10.0, 20.0)):
    """Poisson streams of 2D Fourier KHI runs, sizes {128, 256, 512, 1024}^2
    with weights 0.4/0.3/0.2/0.1, own seed each, t_end ~ U[0.5, 2.0], FP32
    (FP64 time, dt and sums), served with continuous batching from one HBM
    pool (buckets grow and shrink between replays within the budget).

    * the stated stream (`rate` req/s over `window` s): p50/p99 and throughput;
    * the service rate: a closed-loop saturation run (requests queued at
      once) gives completed requests per second of wall time;
    * the sustainable rate: an open-loop stream at 70 % of that rate over two
      windows (`sustain_windows`); p99 that does not grow with the window says
      the queue is stable there."""
    from phrike_amd.serving import EnsembleServer, poisson_requests
    sizes = {128: 0.4, 256: 0.3, 512: 0.2, 1024: 0.1}
    budget = int(budget_gb * 2 ** 30)

    def run(reqs, realtime, label):
        print(f"[bench] ensemble {label}: {len(reqs)} requests", file=sys.stderr, flush=True)
        srv = EnsembleServer(hbm_budget=budget, precision="single", steps_per_replay=32, sizes=tuple(sizes))
        t0 = time.perf_counter()
        res = srv.serve(reqs, realtime=realtime)
        wall = time.perf_counter() - t0
        s = _ensemble_summary(res, wall, sizes)
        s["peak_hbm_gb"] = round(srv.peak_bytes / 2 ** 30, 3)
        s["bucket_resizes"] = {str(n): b.resizes for n, b in srv.buckets.items()}
        return res, wall, s

    out = {"workload": "KHI 2D sizes 128..1024 (0.4/0.3/0.2/0.1), t_end U[0.5,2], FP32, HBM pool "
                       f"{budget_gb} GB", "hbm_budget_gb": budget_gb}
    reqs = poisson_requests(rate, window, sizes, seed=seed)
    res, wall, s = run(reqs, True, "stated stream")
    out["stated"] = {"offered_req_per_s": rate, "window_s": window, "requests": len(reqs),
                     "completed_req_per_s": round(len(res) / wall, 3), **s}
    # closed loop: service rate of the size mix
    sat = poisson_requests(1000.0, 0.06, sizes, seed=seed + 1)
    res, wall, s = run(sat, False, "saturation")
    mu = len(res) / wall
    out["saturation"] = {"requests": len(sat), "service_req_per_s": round(mu, 3), **s}
    lam = round(0.7 * mu, 3)
    sus = {"offered_req_per_s": lam}
    for w in sustain_windows:
        reqs = poisson_requests(lam, w, sizes, seed=seed + 2)
        res, wall, s = run(reqs, True, f"{lam} req/s over {w} s")
        sus[f"window_{int(w)}s"] = {"requests": len(reqs), "completed_req_per_s": round(len(res) / wall, 3), **s}
    p99s = [sus[f"window_{int(w)}s"]["p99_s"] for w in sustain_windows]
    sus["p99_growth"] = round(p99s[-1] / p99s[0], 3) if p99s[0] else None
    out["sustainable"] = sus
    out["sustainable_req_per_s"] = lam if sus["p99_growth"] is not None and sus["p99_growth"] < 1.5 else None
    # the headline figures of the stated stream
    out.update({k: out["stated"][k] for k in ("p50_s", "p99_s", "mcell_updates_per_s", "completed_req_per_s")})
    return out


def load_traffic(kernel: str, n: int, prec: str):
    """Measured HBM bytes per launch of `kernel` (rocprofv3 FETCH_SIZE and
    WRITE_SIZE, scaled by a plain copy's readings), from profiles/."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic.json")
    try:
        d = json.load(open(path))
        return d["kernels"].get(f"{kernel}@{n}^3")
    except Exception:
        return None


def _pmc_per_dispatch(path: str, counter: str):
    """{dispatch id: (kernel symbol, value)} from a rocprofv3 --pmc CSV."""
    import csv
    import re
    out = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            m = re.search(r"(k_\w+<[^>]*>)", r["Kernel_Name"])
            name = m.group(1) if m else r["Kernel_Name"]
            did = int(r["Dispatch_Id"])
            prev = out.get(did, (name, 0.0))[1]
            out[did] = (name, prev + float(r["Counter_Value"]))
    return out


def traffic_table(fetch_csv, write_csv, cal_fetch_csv, cal_write_csv, n, prec, command):
    """HBM bytes per launch of each pipeline kernel from two rocprofv3 passes
    (FETCH_SIZE, WRITE_SIZE: KiB per dispatch), each scaled by what the same
    counter reads for a plain device copy of known size (--calibrate-copy)."""
    cal_bytes = (1 << 27) * 8

    def factor(path, counter):
        # the five 1 GiB copies are the largest dispatches of the calibration run
        vals = sorted((v for _, v in _pmc_per_dispatch(path, counter).values()), reverse=True)[:5]
        med = vals[len(vals) // 2]
        return cal_bytes / (med * 1024.0)

    ff = factor(cal_fetch_csv, "FETCH_SIZE")
    fw = factor(cal_write_csv, "WRITE_SIZE")
    fe = _pmc_per_dispatch(fetch_csv, "FETCH_SIZE")
    wr = _pmc_per_dispatch(write_csv, "WRITE_SIZE")
    es = 8 if prec == "double" else 4
    per, _ = byte_model(n)
    alg = {}
    launches = step_launches(prec)
    for tag, units in per.items():
        if launches.get(tag, 0) > 0:  # tags of this precision's step only
            alg.setdefault(kernel_symbol(tag, n, prec), []).extend([units * es * n ** 3] * launches[tag])
    acc = {}
    for did, (name, v) in fe.items():
        acc.setdefault(name, [0, 0.0, 0.0])
        acc[name][0] += 1
        acc[name][1] += v * 1024.0 * ff
    for did, (name, v) in wr.items():
        if name in acc:
            acc[name][2] += v * 1024.0 * fw
    kern = {}
    for name, (cnt, rd, wt) in acc.items():
        if not name.startswith("k_"):
            continue
        a = alg.get(name)
        kern[f"{name}@{n}^3"] = {"launches": cnt, "read_bytes": int(rd / cnt), "write_bytes": int(wt / cnt),
                                 "bytes_per_launch": int((rd + wt) / cnt),
                                 "alg_bytes_per_launch": int(sum(a) / len(a)) if a else None}
    return {"precision": prec, "size": n, "fetch_factor": round(ff, 4), "write_factor": round(fw, 4),
            "command": command, "kernels": kern}


def run_loop_timing(s, U, steps):
    """The device run loop (ph_run3d) over `steps` steps with dt from the
    CFL condition after every step, as the reference's run loop does
    (phrike/solver.py:1395-1513): ms per step, against the fixed-dt step and
    the standalone max wave-speed reduction (one read of the state) that the
    fused per-row maxima of the step's last kernel replace."""
    import torch
    from phrike_amd import _lib
    plan, ws, st = s._ctx()
    dev = U.device
    rs = torch.zeros(_lib.PH_RS_N, dtype=torch.float64, device=dev)
    rf = torch.zeros(_lib.PH_RF_N, dtype=torch.int32, device=dev)
    rs[_lib.PH_RS_TEND] = 1e30
    rs[_lib.PH_RS_TSTOP] = float("inf")
    a, b = U.clone(), torch.empty_like(U)
    _lib.call("ph_max_wave_speed3d", plan, _lib.ptr(a), None, _lib.ptr(rs[_lib.PH_RS_DT:]), s.cfl, ws, st)

    def run(k):
        rf[_lib.PH_RF_ACTIVE] = 1
        _lib.call("ph_run3d", plan, _lib.ptr(a), _lib.ptr(b), k, _lib.ptr(rs), _lib.ptr(rf), s.cfl,
                  _lib.PH_STEP_BANDLIMITED, ws, st)

    run(2)  # warm-up (even: the state is back in a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    run(steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    nsteps = int(rf[_lib.PH_RF_STEPS].item())
    dts = float(rs[_lib.PH_RS_DTUSED].item())
    m = torch.zeros(1, dtype=torch.float64, device=dev)
    e0.record()
    for _ in range(steps):
        _lib.call("ph_max_wave_speed3d", plan, _lib.ptr(a), _lib.ptr(m), None, s.cfl, ws, st)
    e1.record()
    torch.cuda.synchronize()
    sep = e0.elapsed_time(e1) / steps
    del a, b
    return {"ms_per_step": round(ms, 3), "steps": nsteps, "last_dt": dts, "cfl_every_step": True,
            "fused_cfl_max": True, "separate_max_wave_speed_ms": round(sep, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--precision", choices=["double", "single"], default="double")
    ap.add_argument("--cpu-size", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--no-run-loop", action="store_true",
                    help="skip the device run-loop timing (profiling passes of the step alone)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the FP32-vs-FP64 turbulence study")
    ap.add_argument("--precision-size", type=int, default=512)
    ap.add_argument("--precision-size2", type=int, default=256,
                    help="second size of the FP32-vs-FP64 study (0: none)")
    ap.add_argument("--precision-steps", type=int, default=200)
    ap.add_argument("--hybrid-steps", type=int, default=1000)
    ap.add_argument("--mhd-size", type=int, default=256, help="ideal-MHD study size (0: none)")
    ap.add_argument("--mhd-steps", type=int, default=10)
    ap.add_argument("--torch-steps", type=int, default=2,
                    help="steps of the reference's torch backend arithmetic timed beside (0: none)")
    ap.add_argument("--ensemble-window", type=float, default=3.0)
    ap.add_argument("--ensemble-rate", type=float, default=20.0)
    ap.add_argument("--ensemble-budget-gb", type=float, default=8.0)
    ap.add_argument("--hybrid-only", action="store_true",
                    help="only run the hybrid KHI study (split-precision D.U, ensembles)")
    ap.add_argument("--mhd-only", action="store_true", help="only run the ideal-MHD study")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only run the ensemble serving study (stated stream, saturation, sustainable rate)")
    ap.add_argument("--khi2d-only", action="store_true",
                    help="only run the batched 2D step study (fixed batches, ph_run2d)")
    ap.add_argument("--gll-only", action="store_true",
                    help="only run the batched D.U matrix-core study (for profiling the MFMA kernels)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 1 GiB five times (HBM counter calibration)")
    ap.add_argument("--traffic", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "CAL_FETCH_CSV", "CAL_WRITE_CSV"),
                    help="only build profiles/pmc_traffic.json from rocprofv3 --pmc CSVs of this "
                         "benchmark (FETCH_SIZE, WRITE_SIZE) and of --calibrate-copy")
    args = ap.parse_args()
    if args.traffic:
        tab = traffic_table(*args.traffic, args.size, args.precision,
                            command="rocprofv3 --pmc FETCH_SIZE|WRITE_SIZE -f csv -- python3 bench.py "
                                    f"--size {args.size} --precision {args.precision} --steps 2 --warmup 0 "
                                    "--no-cpu --no-extras --no-kernel-timing --no-run-loop; and -- python3 "
                                    "bench.py --calibrate-copy")
        out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic.json")
        old = {}
        if os.path.exists(out):
            with open(out) as f:
                old = json.load(f)
        merged = {"kernels": {**old.get("kernels", {}), **tab.pop("kernels")},
                  "runs": {**old.get("runs", {}), f"{args.size}^3/{args.precision}": tab}}
        with open(out, "w") as f:
            json.dump(merged, f, indent=1)
        print(json.dumps(tab))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the 3D pseudo-spectral step is not sharded "
              "across GPUs in this backend (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.calibrate_copy:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read_per_launch": a.numel() * 8,
                          "bytes_written_per_launch": b.numel() * 8, "launches": 5}))
        return
    if args.gll_only:
        print(json.dumps({"gll_batched": gll_saturation(torch.device("cuda:0"))}))
        return
    if args.hybrid_only:
        print(json.dumps({"hybrid_khi": hybrid_study(256, (128, 256), args.hybrid_steps, torch.device("cuda:0"))}))
        return
    if args.mhd_only:
        print(json.dumps({"mhd3d": mhd_study(args.mhd_size, args.mhd_steps, torch.device("cuda:0"))}))
        return
    if args.khi2d_only:
        print(json.dumps({"khi2d_batch": khi2d_batch_study(torch.device("cuda:0"))}))
        return
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_study(args.ensemble_rate, args.ensemble_window,
                                                     args.ensemble_budget_gb)}))
        return
    from phrike_amd import _lib
    from phrike_amd.equations import EulerEquations3D
    from phrike_amd.grid import Grid3D
    from phrike_amd.initial_conditions import taylor_green_3d_torch
    from phrike_amd.solver import SpectralSolver3D

    n = args.size
    dev = torch.device("cuda:0")
    g = Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
               filter_params={"enabled": True, "p": 16, "alpha": 36.0}, backend="hip",
               precision=args.precision)
    s = SpectralSolver3D(g, EulerEquations3D(1.4), scheme="rk4", cfl=0.4)
    U = taylor_green_3d_torch(n, dev, g.dtype)
    out = torch.empty_like(U)
    dt = s.compute_dt(U)
    print(f"[bench] TGV {n}^3 {args.precision} dt={dt:.6e} warmup={args.warmup} steps={args.steps}",
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        s.step_into(U, out, dt)
        U, out = out, U
    torch.cuda.synchronize()
    if not args.no_kernel_timing:
        _lib.timing_collect()
        _lib.timing_enable(True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        s.step_into(U, out, dt)
        U, out = out, U
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = e0.elapsed_time(e1)
    kt = {}
    if not args.no_kernel_timing:
        _lib.timing_enable(False)
        kt = _lib.timing_collect()
    ms_per_step = ev_ms / args.steps
    value = n ** 3 * args.steps / (ev_ms / 1e3) / 1e6
    # the step keeps its RK4 sum in spectral space unless the library was
    # built without it (its timing tags say which)
    specacc = not kt or "xmid_sa" in kt
    es = 8 if args.precision == "double" else 4
    per, C = byte_model(n)
    step_bytes = C * es * n ** 3
    kernels = {}
    dominant = None
    for tag, (cnt, ms) in kt.items():
        if tag not in per:
            continue
        avg_ms = ms / cnt
        alg = per.get(tag, 0.0) * es * n ** 3
        kernels[tag] = {"launches": cnt, "avg_ms": round(avg_ms, 4), "total_ms": round(ms, 3),
                        "alg_GBs": round(alg / (avg_ms / 1e3) / 1e9, 1) if alg else None}
        if dominant is None or ms > kt[dominant][1]:
            dominant = tag
    # group the tags by HIP kernel symbol (xmid1 and xmid23 are one kernel),
    # so the average launch duration is the one rocprofv3 reports for it
    groups = {}
    for tag, (cnt, ms) in kt.items():
        if tag not in per:
            continue
        g_ = groups.setdefault(kernel_symbol(tag, n, args.precision), [0, 0.0, 0.0])
        g_[0] += cnt
        g_[1] += ms
        g_[2] += cnt * per[tag] * es * n ** 3
    roof = None
    if groups:
        sym = max(groups, key=lambda k: groups[k][1])
        cnt, ms, alg_tot = groups[sym]
        alg = alg_tot / cnt
        ach = alg / (ms / cnt / 1e3)
        roof = {"bound": "hbm", "achieved": round(ach / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK, 4), "traffic": load_traffic(sym, n, args.precision),
                "kernel": sym, "launches": cnt, "avg_ms": round(ms / cnt, 4),
                "alg_bytes_per_launch": int(alg)}
    sums = s.conserved_sums(U)
    del out
    extras = {}
    if not args.no_run_loop:
        print("[bench] device run loop (dt from CFL every step)", file=sys.stderr, flush=True)
        extras["run_loop"] = run_loop_timing(s, U, args.steps)
        extras["run_loop"]["overhead_vs_fixed_dt_ms"] = round(extras["run_loop"]["ms_per_step"] - ms_per_step, 3)
    if not args.no_extras:
        extras["hbm_reference"] = hbm_reference(dev)
        if roof is not None:
            roof["frac_of_copy"] = round(roof["achieved"] / extras["hbm_reference"]["copy_GBs"], 4)
    if not args.no_extras and args.torch_steps > 0:
        print(f"[bench] the reference's torch backend on torch.fft: TGV {n}^3 FP64 x {args.torch_steps}",
              file=sys.stderr, flush=True)
        del U, s, g
        torch.cuda.empty_cache()
        g2 = Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
                    filter_params={"enabled": True, "p": 16, "alpha": 36.0}, backend="hip", precision="double")
        s2 = SpectralSolver3D(g2, EulerEquations3D(1.4), scheme="rk4", cfl=0.4)
        V = taylor_green_3d_torch(n, dev, torch.float64)
        dt2 = s2.compute_dt(V)
        for _ in range(args.torch_steps):
            V = s2.step(V, dt2)
        del g2, s2
        torch.cuda.empty_cache()
        extras["torch_backend"] = torch_backend_reference(n, args.torch_steps, dev, dt2, V)
        if args.precision == "double":
            extras["torch_backend"]["hip_speedup"] = round(extras["torch_backend"]["ms_per_step"] / ms_per_step, 2)
        del V
        torch.cuda.empty_cache()
    if not args.no_extras:
        print(f"[bench] precision study: turbulence {args.precision_size}^3 x {args.precision_steps}",
              file=sys.stderr, flush=True)
        U = s = g = None
        torch.cuda.empty_cache()
        extras["fp32_vs_fp64"] = precision_study(args.precision_size, args.precision_steps, dev)
        torch.cuda.empty_cache()
        if args.precision_size2 > 0:
            extras[f"fp32_vs_fp64_{args.precision_size2}"] = precision_study(args.precision_size2,
                                                                             args.precision_steps, dev)
            torch.cuda.empty_cache()
        if args.hybrid_steps > 0:
            print(f"[bench] hybrid KHI: {args.hybrid_steps} steps", file=sys.stderr, flush=True)
            extras["hybrid_khi"] = hybrid_study(256, (128, 256), args.hybrid_steps, dev)
            extras["gll_batched"] = gll_saturation(dev)
        if args.mhd_size > 0:
            print(f"[bench] ideal MHD {args.mhd_size}^3", file=sys.stderr, flush=True)
            extras["mhd3d"] = mhd_study(args.mhd_size, args.mhd_steps, dev)
            torch.cuda.empty_cache()
        extras["khi2d_batch"] = khi2d_batch_study(dev)
        torch.cuda.empty_cache()
        if args.ensemble_window > 0:
            print(f"[bench] ensemble: {args.ensemble_rate} req/s over {args.ensemble_window} s",
                  file=sys.stderr, flush=True)
            extras["ensemble"] = ensemble_study(args.ensemble_rate, args.ensemble_window,
                                                args.ensemble_budget_gb)
    cpu = None
    if not args.no_cpu:
        print(f"[bench] cpu baseline: oracle TGV {args.cpu_size}^3 x {args.cpu_steps}", file=sys.stderr,
              flush=True)
        cpu = cpu_baseline(args.cpu_size, args.cpu_steps)
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "double" else "f32", "data": "synthetic",
        "config": {"workload": f"TGV {n}^3 Fourier pseudo-spectral Euler, RK4 + post-step filter "
                               f"(2/3 dealias, exp filter alpha=36 p=16), dt frozen from CFL 0.4",
                   "grid": [n, n, n], "parallelism": "single GPU"},
        "hbm_frac_step": round(step_bytes / (ms_per_step / 1e3) / HBM_PEAK, 4),
        "step_model_bytes": int(step_bytes),
        "step_fused_bytes": int(fused_step_units(n, args.precision, specacc) * es * n ** 3),
        "hbm_frac_step_fused": round(fused_step_units(n, args.precision, specacc) * es * n ** 3
                                     / (ms_per_step / 1e3) / HBM_PEAK, 4),
        # the dense (unpruned) pencil model, C = 420, for transparency
        "hbm_frac_step_dense": round(420 * es * n ** 3 / (ms_per_step / 1e3) / HBM_PEAK, 4),
        "wall_s": round(wall, 3),
        "roofline": roof, "cpu_baseline": cpu, "kernels": kernels,
        "mass_drift": float(sums[0] / n ** 3 - 1.0),
        **extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
