"""Solvers of the ``hip`` backend.

SpectralSolver3D / SpectralSolver2D keep the reference's constructor and
step/compute_dt/run contract (phrike/solver.py:527-812, 1237-1513).  A step is
one C-ABI call (ph_step3d / ph_step2d): four fused RHS evaluations, the RK4
stage combine, the 2D positivity clamp and the post-step filter, all on the
GPU stream.  The CFL reduction and the conservation sums are HIP kernels too.
"""
from __future__ import annotations

import weakref
from typing import Any, Callable, Dict, List, Optional

import numpy as np
import torch

from . import _lib
from .equations import EulerEquations2D, EulerEquations3D
from .grid import Grid2D, Grid3D


def _check_config(scheme: str, adaptive_config, gravity_config, av_config, dim: int):
    if str(scheme).lower() != "rk4":
        raise NotImplementedError(f"the hip backend implements the RK4 path only (scheme={scheme!r})")
    if adaptive_config and adaptive_config.get("enabled", False):
        fallback = adaptive_config.get("scheme", "rk45")
        if fallback in ("rk23", "rk45", "rk78"):
            raise NotImplementedError("adaptive embedded RK is not part of the hip backend")
    if gravity_config and gravity_config.get("enabled", False):
        raise NotImplementedError("gravity source terms are not part of the hip backend")
    if dim == 2 and av_config and av_config.get("enabled", False):
        raise NotImplementedError("artificial viscosity is not part of the hip backend")


class SpectralSolver3D:
    """Fixed-step RK4 + post-step filter every step (phrike/solver.py:1237-1337)."""

    def __init__(self, grid: Grid3D, equations: EulerEquations3D, U0=None, scheme: str = "rk4",
                 cfl: float = 0.25, adaptive_config: Optional[Dict] = None,
                 artificial_viscosity_config: Optional[Dict] = None,
                 gravity_config: Optional[Dict] = None):
        # the 3D reference ignores artificial viscosity (phrike/solver.py:1266)
        _check_config(scheme, adaptive_config, gravity_config, None, 3)
        self.grid = grid
        self.equations = equations
        self.scheme = "rk4"
        self.cfl = float(cfl)
        self.t = 0.0
        self.U = U0
        self.adaptive_enabled = False
        self.gravity_config = gravity_config
        self._plan = grid._plan(equations.gamma)
        self._scalars = torch.zeros(8, dtype=torch.float64, device=grid.device)
        self._last_out = None  # weak reference to the last step's output
        self._last_version = -1

    def _ctx(self):
        return self._plan, _lib.ptr(self.grid.workspace(self.equations.gamma)), \
            _lib.stream_handle(torch, self.grid.device)

    def _state(self, U):
        g = self.grid
        if not isinstance(U, torch.Tensor) or U.device.type != "cuda":
            raise TypeError("U must be a GPU tensor (backend='hip')")
        if tuple(U.shape) != (5, g.Nz, g.Ny, g.Nx) or U.dtype != g.dtype:
            raise ValueError(f"U must be (5, {g.Nz}, {g.Ny}, {g.Nx}) {g.dtype}, got "
                             f"{tuple(U.shape)} {U.dtype}")
        return U.contiguous()

    def max_wave_speed(self, U) -> float:
        plan, ws, st = self._ctx()
        U = self._state(U)
        _lib.call("ph_max_wave_speed3d", plan, _lib.ptr(U), _lib.ptr(self._scalars), None,
                  self.cfl, ws, st)
        return float(self._scalars[0].item())

    def compute_dt(self, U) -> float:
        """cfl*min(dx,dy,dz)/max_speed, 1e-6 if max_speed <= 0
        (phrike/solver.py:1307-1311)."""
        s = self.max_wave_speed(U)
        if s <= 0.0:
            return 1e-6
        return self.cfl * min(self.grid.dx, self.grid.dy, self.grid.dz) / s

    def step(self, U, dt: float):
        U = self._state(U)
        out = torch.empty_like(U)
        self.step_into(U, out, dt)
        return out

    _fixed_step = step

    def step_into(self, U, out, dt: float, dt_dev=None) -> None:
        """out = filter(RK4(U, dt)) with caller-owned output (no allocation).

        When U is this solver's own previous output, unmodified since (same
        tensor object, same version counter), it has no content beyond the 2/3
        cutoff and the step says so (PH_STEP_BANDLIMITED): FP32 plans then
        filter in correction form.  Any other input takes the direct form."""
        plan, ws, st = self._ctx()
        last = self._last_out() if self._last_out is not None else None
        flags = _lib.PH_STEP_BANDLIMITED if (last is U and U._version == self._last_version) else 0
        _lib.call("ph_step3d_ex", plan, _lib.ptr(U), _lib.ptr(out), float(dt), _lib.ptr(dt_dev), flags, ws, st)
        self._last_out = weakref.ref(out)
        self._last_version = out._version

    def rhs(self, U):
        U = self._state(U)
        out = torch.empty_like(U)
        plan, ws, st = self._ctx()
        _lib.call("ph_rhs3d", plan, _lib.ptr(U), _lib.ptr(out), ws, st)
        return out

    def conserved_sums(self, U) -> np.ndarray:
        """[mass, momentum_x, momentum_y, momentum_z, energy] in FP64."""
        plan, ws, st = self._ctx()
        U = self._state(U)
        _lib.call("ph_conserved_sums3d", plan, _lib.ptr(U), _lib.ptr(self._scalars), ws, st)
        return self._scalars[:5].cpu().numpy().copy()

    def run(self, U0, t0: float, t_end: float, output_interval: float = 0.2,
            checkpoint_interval: float = 0.0, outdir: Optional[str] = None,
            on_output: Optional[Callable] = None, on_step: Optional[Callable] = None,
            tracers: Optional[Any] = None) -> Dict[str, List[float]]:
        """Run loop of phrike/solver.py:1395-1513: dt from the CFL condition,
        recomputed after every step, last step clipped to t_end; conserved
        sums recorded at t0, every output_interval and at the end."""
        if tracers is not None:
            raise NotImplementedError("tracers are not part of the hip backend")
        self.U = self._state(U0).clone()
        self.t = float(t0)
        next_output = self.t + output_interval
        next_ckpt = self.t + checkpoint_interval if checkpoint_interval and checkpoint_interval > 0 else np.inf
        keys = ["time", "mass", "momentum_x", "momentum_y", "momentum_z", "energy"]
        history: Dict[str, List[float]] = {k: [] for k in keys}

        def record():
            s = self.conserved_sums(self.U)
            history["time"].append(self.t)
            for k, v in zip(keys[1:], s):
                history[k].append(float(v))

        record()
        if on_output is not None:
            on_output(self.t, self.U)
        dt = self.compute_dt(self.U)
        step_count = 0
        buf = torch.empty_like(self.U)
        while self.t < t_end - 1e-12:
            dt = min(dt, t_end - self.t)
            self.step_into(self.U, buf, dt)
            self.U, buf = buf, self.U
            self.t += dt
            if on_step is not None:
                on_step(step_count, dt, self.U)
            dt = self.compute_dt(self.U)
            step_count += 1
            if self.t + 1e-12 >= next_output:
                record()
                if on_output is not None:
                    on_output(self.t, self.U)
                next_output += output_interval
            if outdir and self.t + 1e-12 >= next_ckpt:
                from .io import save_solution_snapshot
                save_solution_snapshot(outdir, self.t, U=self.U, grid=self.grid,
                                       equations=self.equations)
                next_ckpt += checkpoint_interval
        record()
        if on_output is not None:
            on_output(self.t, self.U)
        return history


class SpectralSolver2D:
    """Fixed-step RK4 with positivity clamp; filter at cadence
    ``filter_params['interval']`` (phrike/solver.py:527-812).

    Accepts one run (4, Ny, Nx) or a batch (B, 4, Ny, Nx) with B <= grid.batch."""

    def __init__(self, grid: Grid2D, equations: EulerEquations2D, U0=None, scheme: str = "rk4",
                 cfl: float = 0.3, adaptive_config: Optional[Dict] = None,
                 artificial_viscosity_config: Optional[Dict] = None,
                 gravity_config: Optional[Dict] = None):
        _check_config(scheme, adaptive_config, gravity_config, artificial_viscosity_config, 2)
        self.grid = grid
        self.equations = equations
        self.scheme = "rk4"
        self.cfl = float(cfl)
        self.t = 0.0
        self.U = U0
        self.adaptive_enabled = False
        self.artificial_viscosity = None
        self.gravity_config = gravity_config
        self.filter_interval = 1
        fp = grid.filter_params
        if isinstance(fp, dict):
            try:
                self.filter_interval = max(1, int(fp.get("interval", 1)))
            except Exception:
                self.filter_interval = 1
        self._step_counter = 0
        self._plan = grid._plan(equations.gamma)
        self._scalars = torch.zeros(8 * max(1, grid.batch), dtype=torch.float64, device=grid.device)

    def _ctx(self):
        return self._plan, _lib.ptr(self.grid.workspace(self.equations.gamma)), \
            _lib.stream_handle(torch, self.grid.device)

    def _runs(self, U):
        g = self.grid
        if not isinstance(U, torch.Tensor) or U.device.type != "cuda":
            raise TypeError("U must be a GPU tensor (backend='hip')")
        if U.dtype != g.dtype or tuple(U.shape[-3:]) != (4, g.Ny, g.Nx) or U.dim() not in (3, 4):
            raise ValueError(f"U must be ([B,] 4, {g.Ny}, {g.Nx}) {g.dtype}, got "
                             f"{tuple(U.shape)} {U.dtype}")
        Ub = U.reshape(-1, 4, g.Ny, g.Nx).contiguous()
        if Ub.shape[0] > g.batch:
            raise ValueError(f"batch {Ub.shape[0]} exceeds the grid's plan batch {g.batch}")
        return Ub

    def max_wave_speed_runs(self, U) -> torch.Tensor:
        plan, ws, st = self._ctx()
        Ub = self._runs(U)
        out = torch.empty(Ub.shape[0], dtype=torch.float64, device=self.grid.device)
        _lib.call("ph_max_wave_speed2d", plan, Ub.shape[0], _lib.ptr(Ub), _lib.ptr(out), None,
                  self.cfl, None, ws, st)
        return out

    def dt_runs(self, U, out=None) -> torch.Tensor:
        """Per-run CFL dt on the device (phrike/solver.py:613-620)."""
        plan, ws, st = self._ctx()
        Ub = self._runs(U)
        if out is None:
            out = torch.empty(Ub.shape[0], dtype=torch.float64, device=self.grid.device)
        _lib.call("ph_max_wave_speed2d", plan, Ub.shape[0], _lib.ptr(Ub), None, _lib.ptr(out),
                  self.cfl, None, ws, st)
        return out

    def compute_dt(self, U) -> float:
        s = float(self.max_wave_speed_runs(U).max().item())
        if s <= 0.0:
            return 1e-6
        dxe = self.grid.dx_min or self.grid.dx
        dye = self.grid.dy_min or self.grid.dy
        return self.cfl * min(dxe, dye) / s

    def step(self, U, dt: float):
        Ub = self._runs(U)
        out = torch.empty_like(Ub)
        apply = (self._step_counter % self.filter_interval) == 0
        self.step_into(Ub, out, dt, apply_filter=apply)
        self._step_counter += 1
        return out.reshape(U.shape)

    _fixed_step = step

    def step_into(self, Ub, out, dt: float, dt_dev=None, active_dev=None, apply_filter: bool = True,
                  nbatch: Optional[int] = None) -> None:
        plan, ws, st = self._ctx()
        nb = Ub.shape[0] if nbatch is None else nbatch
        _lib.call("ph_step2d", plan, nb, _lib.ptr(Ub), _lib.ptr(out), float(dt), _lib.ptr(dt_dev),
                  _lib.ptr(active_dev), int(apply_filter), ws, st)

    def rhs(self, U):
        Ub = self._runs(U)
        out = torch.empty_like(Ub)
        plan, ws, st = self._ctx()
        _lib.call("ph_rhs2d", plan, Ub.shape[0], _lib.ptr(Ub), _lib.ptr(out), ws, st)
        return out.reshape(U.shape)

    def conserved_sums(self, U) -> np.ndarray:
        """Per run [mass, momentum_x, momentum_y, energy] in FP64, shape (B, 4)."""
        plan, ws, st = self._ctx()
        Ub = self._runs(U)
        out = torch.empty(Ub.shape[0] * 4, dtype=torch.float64, device=self.grid.device)
        _lib.call("ph_conserved_sums2d", plan, Ub.shape[0], _lib.ptr(Ub), _lib.ptr(out), ws, st)
        return out.reshape(-1, 4).cpu().numpy()

    def run(self, U0, t0: float, t_end: float, output_interval: float = 0.1,
            checkpoint_interval: float = 0.0, outdir: Optional[str] = None,
            on_output: Optional[Callable] = None, on_step: Optional[Callable] = None,
            tracers: Optional[Any] = None) -> Dict[str, List[float]]:
        """Run loop of phrike/solver.py:706-812 for one run (4, Ny, Nx)."""
        if tracers is not None:
            raise NotImplementedError("tracers are not part of the hip backend")
        self.U = self._runs(U0)[0].clone()
        self.t = float(t0)
        next_output = self.t + output_interval
        next_ckpt = self.t + checkpoint_interval if checkpoint_interval and checkpoint_interval > 0 else np.inf
        keys = ["time", "mass", "momentum_x", "momentum_y", "energy"]
        history: Dict[str, List[float]] = {k: [] for k in keys}

        def record():
            s = self.conserved_sums(self.U)[0]
            history["time"].append(self.t)
            for k, v in zip(keys[1:], s):
                history[k].append(float(v))

        record()
        if on_output is not None:
            on_output(self.t, self.U)
        dt = self.compute_dt(self.U)
        step_count = 0
        while self.t < t_end - 1e-12:
            dt = min(dt, t_end - self.t)
            self.U = self.step(self.U, dt)
            self.t += dt
            if on_step is not None:
                on_step(step_count, dt, self.U)
            dt = self.compute_dt(self.U)
            step_count += 1
            if self.t + 1e-12 >= next_output:
                record()
                if on_output is not None:
                    on_output(self.t, self.U)
                next_output += output_interval
            if outdir and self.t + 1e-12 >= next_ckpt:
                from .io import save_solution_snapshot
                save_solution_snapshot(outdir, self.t, U=self.U, grid=self.grid,
                                       equations=self.equations)
                next_ckpt += checkpoint_interval
        record()
        if on_output is not None:
            on_output(self.t, self.U)
        return history
