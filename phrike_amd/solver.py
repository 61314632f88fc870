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


def _check_pair(U, out, name: str) -> None:
    """out must match U (shape, dtype, device), be contiguous and share no
    memory with it: the library writes it while it still reads U."""
    if not isinstance(out, torch.Tensor):
        raise TypeError(f"{name} must be a GPU tensor")
    if out.shape != U.shape or out.dtype != U.dtype or out.device != U.device:
        raise ValueError(f"{name} must be {tuple(U.shape)} {U.dtype} on {U.device}, got "
                         f"{tuple(out.shape)} {out.dtype} on {out.device}")
    if not out.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    a, b = U.data_ptr(), out.data_ptr()
    n = U.numel() * U.element_size()
    if a < b + n and b < a + n:
        raise ValueError(f"U and {name} must not overlap")


def _check_scalars(t, n: int, dtype, device, name: str) -> None:
    """A per-run device array (dt_dev, active_dev) of at least n entries."""
    if t is None:
        return
    if not isinstance(t, torch.Tensor) or t.dtype != dtype or t.device.type != "cuda" or t.device != device:
        raise ValueError(f"{name} must be a {dtype} tensor on {device}")
    if not t.is_contiguous() or t.numel() < n:
        raise ValueError(f"{name} must be contiguous with at least {n} entries")


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

    def _state(self, U, copy: bool = True):
        """U checked against the grid: (5, Nz, Ny, Nx), the grid's dtype, on the
        grid's GPU; a contiguous copy when it is not contiguous (copy=True)
        or an error (copy=False, for buffers the library writes or reads in
        place)."""
        g = self.grid
        if not isinstance(U, torch.Tensor) or U.device.type != "cuda":
            raise TypeError("U must be a GPU tensor (backend='hip')")
        if tuple(U.shape) != (5, g.Nz, g.Ny, g.Nx) or U.dtype != g.dtype:
            raise ValueError(f"U must be (5, {g.Nz}, {g.Ny}, {g.Nx}) {g.dtype}, got "
                             f"{tuple(U.shape)} {U.dtype}")
        if U.device != g.device and not (g.device.index is None and U.device.index == torch.cuda.current_device()):
            raise ValueError(f"U is on {U.device}, the grid on {g.device}")
        if not copy and not U.is_contiguous():
            raise ValueError("U must be contiguous")
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
        filter in correction form.  Any other input takes the direct form.
        Every step leaves the x-transformed fluxes of its output in the
        workspace (PH_STEP_CHAIN); the next step on that unmodified output
        says so (PH_STEP_CHAINED) and, if nothing has used the workspace in
        between, skips its first x pass."""
        _check_pair(self._state(U, copy=False), out, "out")
        _check_scalars(dt_dev, 1, torch.float64, U.device, "dt_dev")
        plan, ws, st = self._ctx()
        last = self._last_out() if self._last_out is not None else None
        own = last is U and U._version == self._last_version
        flags = _lib.PH_STEP_CHAIN | ((_lib.PH_STEP_BANDLIMITED | _lib.PH_STEP_CHAINED) if own else 0)
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
            tracers: Optional[Any] = None, max_steps_per_launch: int = 64) -> Dict[str, List[float]]:
        """Run loop of phrike/solver.py:1395-1513: dt from the CFL condition,
        recomputed after every step, last step clipped to t_end; conserved
        sums recorded at t0, every output_interval and at the end.  The loop
        itself runs on the device (ph_run3d; see _device_run_loop)."""
        if tracers is not None:
            raise NotImplementedError("tracers are not part of the hip backend")
        keys = ["time", "mass", "momentum_x", "momentum_y", "momentum_z", "energy"]
        plan, ws, st = self._ctx()

        def dt0(U, rs):
            _lib.call("ph_max_wave_speed3d", plan, _lib.ptr(U), None, _lib.ptr(rs[_lib.PH_RS_DT:]),
                      self.cfl, ws, st)

        calls = [0]

        def launch(a, b, n, rs, rf, flags):
            # every call chains its last step to the next call, which starts
            # from the buffer holding this call's final state, so the run's
            # result does not depend on how its steps fall into calls
            flags |= _lib.PH_STEP_CHAIN | (_lib.PH_STEP_CHAINED if calls[0] else 0)
            calls[0] += 1
            _lib.call("ph_run3d", plan, _lib.ptr(a), _lib.ptr(b), n, _lib.ptr(rs), _lib.ptr(rf), self.cfl,
                      flags, ws, st)

        return _device_run_loop(self, self._state(U0).clone(), t0, t_end, output_interval, checkpoint_interval,
                                outdir, on_output, on_step, keys, lambda U: self.conserved_sums(U), dt0, launch,
                                max_steps_per_launch)


def _device_run_loop(solver, U, t0, t_end, output_interval, checkpoint_interval, outdir, on_output, on_step,
                     keys, sums, dt0, launch, max_steps_per_launch):
    """The reference run loop (phrike/solver.py:1395-1513, 706-812) with its
    per-step arithmetic on the device.  t, dt, the clip to t_end and the stop
    test live in device memory (ph_run3d / ph_run2d), so the host launches
    steps in chunks and reads t back only when a chunk ends: at the next
    output or checkpoint time, at t_end, or after max_steps_per_launch steps
    (chunks are sized from the last dt).  With on_step every step ends a
    chunk, since the callback takes each step's dt and state.  Callbacks get
    a copy of the state, as the reference's functional step hands each step a
    fresh array."""
    dev = U.device
    bufs = [U, torch.empty_like(U)]
    cur = 0
    solver.U = bufs[0]
    solver.t = float(t0)
    t_end = float(t_end)
    next_output = solver.t + output_interval
    next_ckpt = solver.t + checkpoint_interval if checkpoint_interval and checkpoint_interval > 0 else np.inf
    history: Dict[str, List[float]] = {k: [] for k in keys}

    def record():
        v = sums(solver.U)
        history["time"].append(solver.t)
        for k, x in zip(keys[1:], v):
            history[k].append(float(x))

    record()
    if on_output is not None:
        on_output(solver.t, solver.U.clone())
    rs = torch.zeros(_lib.PH_RS_N, dtype=torch.float64, device=dev)
    rf = torch.zeros(_lib.PH_RF_N, dtype=torch.int32, device=dev)
    rs[_lib.PH_RS_T] = solver.t
    rs[_lib.PH_RS_TEND] = t_end
    dt0(solver.U, rs)  # dt of the initial state (phrike/solver.py:1472)
    flags = 0  # the initial state may carry modes beyond the 2/3 cutoff
    step_count = 0
    dt_est = None
    while solver.t < t_end - 1e-12:
        t_stop = min(next_output, next_ckpt)
        n = 1
        if on_step is None and dt_est is not None and dt_est > 0 and np.isfinite(dt_est):
            n = int(min(max_steps_per_launch, max(1.0, np.ceil((min(t_stop, t_end) - solver.t) / dt_est) + 1)))
        rs[_lib.PH_RS_TSTOP] = t_stop
        rf[_lib.PH_RF_ACTIVE] = 1
        launch(bufs[cur], bufs[1 - cur], n, rs, rf, flags)
        hs = rs.cpu().numpy()
        steps = int(rf[_lib.PH_RF_STEPS].item())
        done = steps - step_count
        step_count = steps
        cur = (cur + done) % 2
        solver.U = bufs[cur]
        solver.t = float(hs[_lib.PH_RS_T])
        dt_est = float(hs[_lib.PH_RS_DT])
        if done > 0:
            flags = _lib.PH_STEP_BANDLIMITED
        if on_step is not None and done:
            on_step(step_count - 1, float(hs[_lib.PH_RS_DTUSED]), solver.U.clone())
        if solver.t + 1e-12 >= next_output:
            record()
            if on_output is not None:
                on_output(solver.t, solver.U.clone())
            next_output += output_interval
        if outdir and solver.t + 1e-12 >= next_ckpt:
            from .io import save_solution_snapshot
            save_solution_snapshot(outdir, solver.t, U=solver.U, grid=solver.grid, equations=solver.equations)
            next_ckpt += checkpoint_interval
        if done == 0:  # only a non-finite time stops a launched run without a step
            break
    record()
    if on_output is not None:
        on_output(solver.t, solver.U.clone())
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
        """out = RK4 step (+ filter) of the first nbatch runs of Ub (B, 4, Ny, Nx);
        per-run dt from dt_dev and runs with active_dev == 0 left untouched
        when given."""
        g = self.grid
        if not isinstance(Ub, torch.Tensor) or Ub.dim() != 4 or tuple(Ub.shape[1:]) != (4, g.Ny, g.Nx) \
                or Ub.dtype != g.dtype or Ub.device.type != "cuda" or not Ub.is_contiguous():
            raise ValueError(f"Ub must be a contiguous (B, 4, {g.Ny}, {g.Nx}) {g.dtype} GPU tensor")
        _check_pair(Ub, out, "out")
        nb = Ub.shape[0] if nbatch is None else int(nbatch)
        if not 1 <= nb <= min(Ub.shape[0], g.batch):
            raise ValueError(f"nbatch must be in [1, {min(Ub.shape[0], g.batch)}]")
        _check_scalars(dt_dev, nb, torch.float64, Ub.device, "dt_dev")
        _check_scalars(active_dev, nb, torch.int32, Ub.device, "active_dev")
        plan, ws, st = self._ctx()
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
            tracers: Optional[Any] = None, max_steps_per_launch: int = 64) -> Dict[str, List[float]]:
        """Run loop of phrike/solver.py:706-812 for one run (4, Ny, Nx).  With
        the filter every step (interval 1, the default) the loop runs on the
        device (ph_run2d; see _device_run_loop); another filter interval keeps
        the host loop, whose cadence counter is the solver's."""
        if tracers is not None:
            raise NotImplementedError("tracers are not part of the hip backend")
        keys = ["time", "mass", "momentum_x", "momentum_y", "energy"]
        if self.filter_interval == 1:
            plan, ws, st = self._ctx()

            def dt0(U, rs):
                _lib.call("ph_max_wave_speed2d", plan, 1, _lib.ptr(U), None, _lib.ptr(rs[_lib.PH_RS_DT:]),
                          self.cfl, None, ws, st)

            def launch(a, b, n, rs, rf, flags):
                _lib.call("ph_run2d", plan, 1, _lib.ptr(a), _lib.ptr(b), n, _lib.ptr(rs), _lib.ptr(rf), self.cfl,
                          ws, st)

            return _device_run_loop(self, self._runs(U0)[0].clone(), t0, t_end, output_interval,
                                    checkpoint_interval, outdir, on_output, on_step, keys,
                                    lambda U: self.conserved_sums(U)[0], dt0, launch, max_steps_per_launch)
        self.U = self._runs(U0)[0].clone()
        self.t = float(t0)
        next_output = self.t + output_interval
        next_ckpt = self.t + checkpoint_interval if checkpoint_interval and checkpoint_interval > 0 else np.inf
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
