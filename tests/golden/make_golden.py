"""Write golden input/output vectors from the reference, in this container only.

Usage:  python tests/golden/make_golden.py [--skip-long]

The reference package at /root/reference is imported by path (its package
``__init__`` is bypassed so nothing beyond the numerics modules is loaded, and
no bytecode is written into the reference tree).  Only functions whose exact
composition the closed-form checks cannot pin are recorded:

* F1  3D RK4 step + post-step filter (SpectralSolver3D.step) on the benchmark
      TGV at 16^3 and 32^3, dt frozen from compute_dt(U0) at cfl 0.4, after 1
      and 10 steps (phrike/solver.py:1220-1337);
* F2  the same at 64^3 for 200 steps: per-field norms/sums, a stride-4
      subsample and the conserved sums;
* F3  one _compute_rhs_3d on a band-limited random state at 16^3 and on a
      non-cubic 16x8x32 grid (phrike/solver.py:1177-1184);
* F5  LegendreLobattoBasis1D nodes, weights, D_x and nodal filter for
      N = 8, 16, 32, 128 and checksums for N = 256 (phrike/basis/legendre.py);
* F8  TGV 256^3, 200 steps, dt frozen at t=0: stride-8 subsample, the
      Fourier modes |m| <= 8, norms and conserved sums after 1, 3, 200 steps;
* F6  2D KHI (periodic profile) 10 steps of SpectralSolver2D.step with the
      per-step compute_dt, clamp and filter, at 64^2 and 32x16, plus a
      filter-cadence case (phrike/solver.py:428-648).
* F9  3D ideal MHD (SpectralSolverMHD3D with the reference defaults,
      phrike/solver.py:840-1050, 1173; MHDEquations3D, phrike/equations.py:
      911-1100; Grid3D.project_solenoidal / divergence, phrike/grid.py:
      1635-1684): the circularly polarised Alfven wave on 32x8x8 (RHS, 1 and 5
      steps) and a band-limited random state with a non-solenoidal B on 16^3
      (fluxes, RHS, projection, CFL speed, 1 and 3 steps, conserved sums).
* F10 1D Euler on the LGL grid with the wall rule (Grid1D basis="legendre",
      bc="wall"; phrike/grid.py:144-738, phrike/solver.py:33-131): one RHS
      and 1 / 5 RK4 steps (clamp, filter, wall rule) with the filter on and
      off: the x-uniform restriction of the hybrid Fourier-x / LGL-y path.
* F11 snapshots written by the reference's save_solution_snapshot
      (phrike/io.py:103-236) for 3D Euler, 2D Euler and 3D MHD states (meta
      stored as a pickled dict), for the restart loader's tests.
* F12 BaseProblem's monitoring statistics (phrike/problems/base.py:490-822)
      on torch tensors: conservation errors, velocity, density and MHD
      statistics of 3D Euler, 2D Euler and 3D MHD states.

The outputs are plain ``.npz`` data (no pickles), except F11, whose ``meta``
entry is the reference's own pickled dict (the case the loader has to
handle; nothing here unpickles it).  numba, pyfftw and finufft
are absent here, so the reference runs its plain numpy/scipy branches; the
versions used are recorded in every file.
"""
from __future__ import annotations

import argparse
import os
import sys
import time
import types

import numpy as np

sys.dont_write_bytecode = True
REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from phrike_amd import initial_conditions as ics  # noqa: E402


def _load_reference():
    pkg = types.ModuleType("phrike")
    pkg.__path__ = [os.path.join(REF, "phrike")]
    sys.modules["phrike"] = pkg
    import importlib

    grid = importlib.import_module("phrike.grid")
    eqs = importlib.import_module("phrike.equations")
    solver = importlib.import_module("phrike.solver")
    leg = importlib.import_module("phrike.basis.legendre")
    return grid, eqs, solver, leg


def _meta():
    import scipy

    return dict(numpy=np.__version__, scipy=scipy.__version__, numba="absent",
                pyfftw="absent", created=time.strftime("%Y-%m-%d"))


FILT = {"enabled": True, "p": 16, "alpha": 36.0}


def f1_f2(grid_m, eq_m, sol_m, out, skip_long):
    for n, nsteps_list in ((16, (1, 10)), (32, (1, 10))):
        g = grid_m.Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
                          filter_params=dict(FILT), fft_workers=1)
        e = eq_m.EulerEquations3D(gamma=1.4)
        s = sol_m.SpectralSolver3D(g, e, scheme="rk4", cfl=0.4)
        U0 = ics.taylor_green_3d(n, n, n)
        dt = s.compute_dt(U0)
        res = {"U0": U0, "dt": np.float64(dt)}
        U = U0
        for k in range(1, max(nsteps_list) + 1):
            U = s.step(U, dt)
            if k in nsteps_list:
                res[f"U_{k}"] = U
        np.savez_compressed(os.path.join(out, f"tgv3d_{n}.npz"), **res, **_meta())
        print("F1", n, dt, flush=True)
    if skip_long:
        return
    n, nsteps = 64, 200
    g = grid_m.Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
                      filter_params=dict(FILT), fft_workers=os.cpu_count())
    e = eq_m.EulerEquations3D(gamma=1.4)
    s = sol_m.SpectralSolver3D(g, e, scheme="rk4", cfl=0.4)
    U = ics.taylor_green_3d(n, n, n)
    dt = s.compute_dt(U)
    t0 = time.time()
    for k in range(nsteps):
        U = s.step(U, dt)
    el = time.time() - t0
    cq = e.conserved_quantities(U)
    np.savez_compressed(
        os.path.join(out, "tgv3d_64_200.npz"),
        dt=np.float64(dt), steps=np.int64(nsteps),
        sub4=U[:, ::4, ::4, ::4].copy(),
        l2=np.sqrt((U ** 2).reshape(5, -1).sum(axis=1)),
        amax=np.abs(U).reshape(5, -1).max(axis=1),
        sums=U.reshape(5, -1).sum(axis=1),
        conserved=np.array([cq["mass"], cq["momentum_x"], cq["momentum_y"],
                            cq["momentum_z"], cq["energy"]]),
        ref_seconds=np.float64(el), **_meta())
    print("F2", dt, el, flush=True)


def lowmodes(U, m=8):
    """Fourier coefficients |mx|,|my|,|mz| <= m of each field (rfftn order)."""
    F = np.fft.rfftn(U, axes=(-3, -2, -1))
    idx = np.r_[0:m + 1, -m:0]
    return F[:, idx][:, :, idx][:, :, :, : m + 1]


def f8(grid_m, eq_m, sol_m, out, n=256, nsteps=200):
    """TGV n^3, dt frozen at t=0, nsteps reference steps: stride-8 subsample,
    low-mode spectrum, norms and conserved sums after 1, 3 and nsteps steps."""
    g = grid_m.Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
                      filter_params=dict(FILT), fft_workers=int(os.environ.get("GOLDEN_WORKERS", "6")))
    e = eq_m.EulerEquations3D(gamma=1.4)
    s = sol_m.SpectralSolver3D(g, e, scheme="rk4", cfl=0.4)
    U = ics.taylor_green_3d(n, n, n)
    t0 = time.time()
    dt = s.compute_dt(U)
    print("F8 dt", dt, time.time() - t0, flush=True)
    res = {"dt": np.float64(dt), "steps": np.int64(nsteps)}
    t0 = time.time()
    for k in range(1, nsteps + 1):
        U = s.step(U, dt)
        if k in (1, 3, nsteps):
            cq = e.conserved_quantities(U)
            res[f"sub8_{k}"] = U[:, ::8, ::8, ::8].copy()
            res[f"low8_{k}"] = lowmodes(U)
            res[f"l2_{k}"] = np.sqrt((U ** 2).reshape(5, -1).sum(axis=1))
            res[f"sums_{k}"] = U.reshape(5, -1).sum(axis=1)
            res[f"conserved_{k}"] = np.array([cq["mass"], cq["momentum_x"], cq["momentum_y"],
                                              cq["momentum_z"], cq["energy"]])
            np.savez_compressed(os.path.join(out, f"tgv3d_{n}_{nsteps}.npz"), **res,
                                ref_seconds=np.float64(time.time() - t0), **_meta())
        if k % 10 == 0:
            print("F8 step", k, time.time() - t0, flush=True)


def f3(grid_m, eq_m, sol_m, out):
    res = {}
    for tag, (nx, ny, nz) in (("cube16", (16, 16, 16)), ("box16x8x32", (16, 8, 32))):
        g = grid_m.Grid3D(nx, ny, nz, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
                          filter_params=dict(FILT), fft_workers=1)
        e = eq_m.EulerEquations3D(gamma=1.4)
        U = ics.smooth_random_state_3d(nx, ny, nz, seed=11, kmax=4)
        res[f"{tag}_U"] = U
        res[f"{tag}_rhs"] = sol_m._compute_rhs_3d(g, e, U)
        res[f"{tag}_dx_rho"] = g.dx1(U[0])
        res[f"{tag}_dy_rho"] = g.dy1(U[0])
        res[f"{tag}_dz_rho"] = g.dz1(U[0])
        res[f"{tag}_filter"] = g.apply_spectral_filter(U)
        if tag == "cube16":
            Fx, Fy, Fz = e.flux(U)
            res[f"{tag}_Fx"], res[f"{tag}_Fy"], res[f"{tag}_Fz"] = Fx, Fy, Fz
        res[f"{tag}_maxc"] = np.float64(e.max_wave_speed(U))
        s = sol_m.SpectralSolver3D(g, e, scheme="rk4", cfl=0.4)
        dt = s.compute_dt(U)
        res[f"{tag}_dt"] = np.float64(dt)
        res[f"{tag}_step"] = s.step(U, dt)
    # filter disabled, dealias on: the post-step filter still applies the mask
    g = grid_m.Grid3D(16, 16, 16, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
                      filter_params={"enabled": False}, fft_workers=1)
    e = eq_m.EulerEquations3D(gamma=1.4)
    s = sol_m.SpectralSolver3D(g, e, scheme="rk4", cfl=0.4)
    U = res["cube16_U"]
    res["nofilter_step"] = s.step(U, 1e-3)
    np.savez_compressed(os.path.join(out, "rhs3d.npz"), **res, **_meta())
    print("F3", flush=True)


def f5(leg_m, out):
    res = {}
    for n in (8, 16, 32, 128, 256):
        L = 1.0
        b = leg_m.LegendreLobattoBasis1D(n, L)
        D = np.asarray(b._D_x)
        if n <= 128:
            res[f"N{n}_x"] = np.asarray(b.nodes())
            res[f"N{n}_w"] = np.asarray(b.quadrature_weights())
            res[f"N{n}_D"] = D
            res[f"N{n}_filter_p16_a36"] = np.asarray(b._build_nodal_filter_np(16, 36.0))
        res[f"N{n}_rowsum_absmax"] = np.float64(np.abs(D.sum(axis=1)).max())
        res[f"N{n}_fro"] = np.float64(np.linalg.norm(D))
        res[f"N{n}_x_sum"] = np.float64(np.asarray(b.nodes()).sum())
        res[f"N{n}_D_sum_abs"] = np.float64(np.abs(D).sum())
        res[f"N{n}_D_diag"] = np.diag(D).copy()
    np.savez_compressed(os.path.join(out, "lgl.npz"), **res, **_meta())
    print("F5", flush=True)


def f6(grid_m, eq_m, sol_m, out):
    res = {}
    cases = (("sq64", 64, 64, 1), ("rect32x16", 32, 16, 1), ("cad2_32", 32, 32, 2))
    for tag, nx, ny, interval in cases:
        fp = dict(FILT)
        fp["interval"] = interval
        g = grid_m.Grid2D(nx, ny, 1.0, 1.0, dealias=True, filter_params=fp, fft_workers=1,
                          basis_x="fourier", basis_y="fourier")
        e = eq_m.EulerEquations2D(gamma=1.4)
        s = sol_m.SpectralSolver2D(g, e, scheme="rk4", cfl=0.4)
        U = ics.khi_2d_periodic(nx, ny, seed=5)
        res[f"{tag}_U0"] = U
        dts = []
        for k in range(10):
            dt = s.compute_dt(U)
            dts.append(dt)
            U = s.step(U, dt)
            if k == 0:
                res[f"{tag}_U1"] = U
        res[f"{tag}_U10"] = U
        res[f"{tag}_dts"] = np.array(dts)
        res[f"{tag}_rhs0"] = sol_m._compute_rhs_2d(g, e, res[f"{tag}_U0"])
    np.savez_compressed(os.path.join(out, "khi2d.npz"), **res, **_meta())
    print("F6", flush=True)


def _alfven_ic(e, X, rho0=1.0, p0=0.1, B0=1.0, b1=0.1, m=1, Lx=1.0):
    # phrike/problems/problem_list/alfven3d.py:24-47 (direction +1, t = 0)
    k = 2.0 * np.pi * m / Lx
    By, Bz = b1 * np.cos(k * X), b1 * np.sin(k * X)
    inv = 1.0 / np.sqrt(rho0)
    return e.conservative(np.full_like(X, rho0), np.zeros_like(X), -By * inv, -Bz * inv,
                          np.full_like(X, p0), np.full_like(X, B0), By, Bz)


def f9(grid_m, eq_m, sol_m, out):
    res = {}
    mhd = {"resistivity": 0.0, "viscosity": 0.0, "divergence_clean": "projection",
           "project_each_substage": True}
    fp8 = {"enabled": True, "p": 8, "alpha": 36.0}
    g = grid_m.Grid3D(32, 8, 8, 1.0, 1.0, 1.0, dealias=True, filter_params=dict(fp8), fft_workers=1)
    g.mhd_config = dict(mhd)
    e = eq_m.MHDEquations3D(gamma=5.0 / 3.0)
    s = sol_m.SpectralSolverMHD3D(g, e, scheme="rk4", cfl=0.3)
    X, _, _ = g.xyz_mesh()
    U0 = _alfven_ic(e, X)
    dt = s.compute_dt(U0)
    res["alf_U0"], res["alf_dt"] = U0, np.float64(dt)
    res["alf_rhs"] = s._compute_rhs(U0)
    U = U0
    for k in range(5):
        U = s.step(U, dt)
        if k == 0:
            res["alf_U1"] = U
    res["alf_U5"] = U
    # band-limited random state, 16^3 on [0, 2 pi)^3, B with a longitudinal part
    g = grid_m.Grid3D(16, 16, 16, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
                      filter_params=dict(FILT), fft_workers=1)
    g.mhd_config = dict(mhd)
    s = sol_m.SpectralSolverMHD3D(g, e, scheme="rk4", cfl=0.3)
    H = ics.smooth_random_state_3d(16, 16, 16, seed=21, kmax=4)
    Bn = ics.smooth_random_state_3d(16, 16, 16, seed=22, kmax=3)
    B = np.stack([0.3 * (Bn[1] / Bn[0]), 0.3 * (Bn[2] / Bn[0]), 1.0 + 0.3 * (Bn[3] / Bn[0])])
    U = np.concatenate([H, B])
    U[4] = H[4] + 0.5 * (B[0] ** 2 + B[1] ** 2 + B[2] ** 2)
    res["rnd_U"] = U
    Fx, Fy, Fz = e.flux(U)
    res["rnd_Fx"], res["rnd_Fy"], res["rnd_Fz"] = Fx, Fy, Fz
    res["rnd_rhs"] = s._compute_rhs(U)
    res["rnd_proj"] = np.array(g.project_solenoidal(U[5], U[6], U[7]))
    res["rnd_div"] = g.divergence(U[5], U[6], U[7])
    res["rnd_maxc"] = np.float64(e.max_wave_speed(U))
    dt = s.compute_dt(U)
    res["rnd_dt"] = np.float64(dt)
    cq = e.conserved_quantities(U)
    res["rnd_cons_keys"] = np.array(list(cq.keys()))
    res["rnd_cons"] = np.array(list(cq.values()))
    V = U
    for k in range(3):
        V = s.step(V, dt)
        if k == 0:
            res["rnd_U1"] = V
    res["rnd_U3"] = V
    np.savez_compressed(os.path.join(out, "mhd3d.npz"), **res, **_meta())
    print("F9", flush=True)


def _wall1d_state(x, L):
    # smooth, y-varying, with non-zero momentum at both walls (so the wall
    # rule changes the boundary nodes of every RHS input)
    s = x / L
    rho = 1.0 + 0.2 * np.sin(2.0 * np.pi * s) + 0.05 * np.cos(5.0 * np.pi * s)
    u = 0.3 * np.sin(np.pi * s + 0.4) + 0.05
    p = 1.0 + 0.1 * np.cos(3.0 * np.pi * s)
    return rho, u, p


def f10(grid_m, eq_m, sol_m, out):
    """1D Euler on the Legendre-Gauss-Lobatto grid with the wall boundary rule
    (Grid1D basis="legendre", bc="wall"; phrike/grid.py:144-738,
    phrike/solver.py:33-131): one _compute_rhs and 1 / 5 calls of _rk4_step
    (RK4, clamp, filter, wall rule) with the filter on and off, at N = 16, 33
    and 64.  An x-uniform hybrid Fourier-x / LGL-y state with zero x momentum
    is this 1D problem along y."""
    res = {}
    for n in (16, 33, 64):
        L = 1.0
        for ftag, fp in (("f", dict(FILT)), ("nof", None)):
            g = grid_m.Grid1D(n, L, basis="legendre", bc="wall", dealias=False, filter_params=fp)
            e = eq_m.EulerEquations1D(gamma=1.4)
            x = np.asarray(g.x)
            rho, u, p = _wall1d_state(x, L)
            U0 = e.conservative(rho, u, p)
            key = f"N{n}_{ftag}"
            res[f"{key}_x"] = x
            res[f"{key}_U0"] = U0.copy()
            res[f"{key}_rhs0"] = sol_m._compute_rhs(g, e, U0.copy())
            dt = 0.2 * float(g.dx) / float(np.max(np.abs(u) + np.sqrt(1.4 * p / rho)))
            res[f"{key}_dt"] = np.float64(dt)
            U = U0.copy()
            for k in range(5):
                U = sol_m._rk4_step(g, e, U, dt)
                if k == 0:
                    res[f"{key}_U1"] = U.copy()
            res[f"{key}_U5"] = U
    np.savez_compressed(os.path.join(out, "wall1d.npz"), **res, **_meta())
    print("F10", flush=True)


def f11(grid_m, eq_m, sol_m, out):
    """Snapshots written by the reference's own save_solution_snapshot
    (phrike/io.py:103-236) for a 3D, a 2D and a 3D MHD state: their meta is a
    pickled dict, which the hip backend's loader must not unpickle."""
    import shutil
    import tempfile
    io_m = __import__("importlib").import_module("phrike.io")
    d = os.path.join(out, "ref_snapshots")
    os.makedirs(d, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        g3 = grid_m.Grid3D(8, 8, 16, 2 * np.pi, np.pi, 4 * np.pi, dealias=True, fft_workers=1)
        e3 = eq_m.EulerEquations3D(gamma=1.4)
        U3 = ics.smooth_random_state_3d(8, 8, 16, seed=3, kmax=2)
        f = io_m.save_solution_snapshot(tmp, 0.125, U3, g3, e3)
        shutil.copy(f, os.path.join(d, "euler3d_" + os.path.basename(f)))
        g2 = grid_m.Grid2D(16, 8, 1.0, 0.5, dealias=True, fft_workers=1, basis_x="fourier", basis_y="fourier")
        e2 = eq_m.EulerEquations2D(gamma=5.0 / 3.0)
        U2 = ics.khi_2d_periodic(16, 8, seed=2)
        f = io_m.save_solution_snapshot(tmp, 0.5, U2, g2, e2, tag="khi")
        shutil.copy(f, os.path.join(d, "euler2d_" + os.path.basename(f)))
        gm = grid_m.Grid3D(8, 8, 8, 1.0, 1.0, 1.0, dealias=True, fft_workers=1)
        em = eq_m.MHDEquations3D(gamma=5.0 / 3.0)
        X, _, _ = gm.xyz_mesh()
        Um = _alfven_ic(em, X)
        f = io_m.save_solution_snapshot(tmp, 0.25, Um, gm, em)
        shutil.copy(f, os.path.join(d, "mhd3d_" + os.path.basename(f)))
    print("F11", sorted(os.listdir(d)), flush=True)


def f12(grid_m, eq_m, sol_m, out):
    """The monitoring statistics of BaseProblem (phrike/problems/base.py:
    490-822) on torch CPU tensors (the branch a tensor state takes):
    conservation errors, velocity, density and MHD statistics of a 3D Euler,
    a 2D Euler and a 3D MHD state."""
    import importlib
    import types as _t

    import torch
    base = importlib.import_module("phrike.problems.base")
    BP = base.BaseProblem
    res = {}

    def stats(tag, solver, U0, U1):
        me = _t.SimpleNamespace(monitoring_initial_values=None)
        me.monitoring_initial_values = solver.equations.conserved_quantities(torch.as_tensor(U0))
        T1 = torch.as_tensor(U1)
        out = {}
        out.update(BP.compute_conservation_errors(me, solver, T1))
        out.update(BP.compute_velocity_stats(me, solver, T1))
        out.update(BP.compute_density_stats(me, solver, T1))
        out.update(BP.compute_mhd_stats(me, solver, T1))
        res[f"{tag}_U0"], res[f"{tag}_U1"] = U0, U1
        res[f"{tag}_keys"] = np.array(list(out.keys()))
        res[f"{tag}_vals"] = np.array([float(v) for v in out.values()])

    rng = np.random.default_rng(12)
    e3 = eq_m.EulerEquations3D(gamma=1.4)
    U0 = ics.smooth_random_state_3d(16, 8, 32, seed=31, kmax=3)
    U1 = U0 * (1.0 + 0.01 * rng.standard_normal(U0.shape))
    stats("e3", _t.SimpleNamespace(equations=e3, grid=None), U0, U1)
    e2 = eq_m.EulerEquations2D(gamma=1.4)
    V0 = ics.khi_2d_periodic(32, 16, seed=8)
    V1 = V0 * (1.0 + 0.01 * rng.standard_normal(V0.shape))
    stats("e2", _t.SimpleNamespace(equations=e2, grid=_t.SimpleNamespace()), V0, V1)
    em = eq_m.MHDEquations3D(gamma=5.0 / 3.0)
    H = ics.smooth_random_state_3d(8, 8, 8, seed=41, kmax=2)
    B = np.stack([1.0 + 0.2 * H[1] / H[0], 0.2 * H[2] / H[0], 0.2 * H[3] / H[0]])
    M0 = np.concatenate([H, B])
    M0[4] = H[4] + 0.5 * (B ** 2).sum(axis=0)
    M1 = M0 * (1.0 + 0.01 * rng.standard_normal(M0.shape))
    # max|divB| comes from the solver; a stand-in returning a fixed value
    stats("mhd", _t.SimpleNamespace(equations=em, grid=None, max_div_b=lambda U: 0.125), M0, M1)
    np.savez_compressed(os.path.join(out, "monitor.npz"), **res, **_meta())
    print("F12", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-long", action="store_true")
    ap.add_argument("--only", default="")
    a = ap.parse_args()
    grid_m, eq_m, sol_m, leg_m = _load_reference()
    only = set(a.only.split(",")) if a.only else {"f1", "f3", "f5", "f6", "f9", "f10", "f11", "f12"}
    if "f12" in only:
        f12(grid_m, eq_m, sol_m, HERE)
    if "f11" in only:
        f11(grid_m, eq_m, sol_m, HERE)
    if "f10" in only:
        f10(grid_m, eq_m, sol_m, HERE)
    if "f3" in only:
        f3(grid_m, eq_m, sol_m, HERE)
    if "f5" in only:
        f5(leg_m, HERE)
    if "f6" in only:
        f6(grid_m, eq_m, sol_m, HERE)
    if "f1" in only:
        f1_f2(grid_m, eq_m, sol_m, HERE, a.skip_long)
    if "f8" in only:
        f8(grid_m, eq_m, sol_m, HERE)
    if "f9" in only:
        f9(grid_m, eq_m, sol_m, HERE)


if __name__ == "__main__":
    main()
