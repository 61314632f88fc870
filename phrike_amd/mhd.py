"""3D ideal MHD on the ``hip`` backend.

MHDEquations3D and SpectralSolverMHD3D keep the reference's names, constructor
arguments and step / compute_dt / run / max_div_b contract
(phrike/equations.py:911-1100, phrike/solver.py:840-1163, 1173).  The state is
(8, Nz, Ny, Nx) on the GPU: rho, rho ux, rho uy, rho uz, E, Bx, By, Bz.  A step
is one C-ABI call (ph_step_mhd3d_ex): four RHS evaluations on the pencil
pipeline with 8 fields, the clamp and the Helmholtz projection of B after every
substage, then the spectral filter and the projection.  When B of the input is
known to be solenoidal (the solver's own previous output, or its projected
initial state) the step says so (PH_STEP_SOLENOIDAL): the substage projections
are then the identity up to round-off and the library runs the fused form.  The runtime options
are read from ``grid.mhd_config`` as in the reference; only its defaults
(no resistivity or viscosity, projection after every substage) are built.
"""
from __future__ import annotations

import weakref
from typing import Any, Callable, Dict, List, Optional

import numpy as np
import torch

from . import _lib
from .grid import Grid3D


class MHDEquations3D:
    """Ideal MHD, conservative form (phrike/equations.py:911-1100).  The
    pointwise helpers work on numpy arrays or torch tensors."""

    def __init__(self, gamma: float = 5.0 / 3.0):
        self.gamma = float(gamma)

    def primitive(self, U):
        rho = U[0]
        ux, uy, uz = U[1] / rho, U[2] / rho, U[3] / rho
        Bx, By, Bz = U[5], U[6], U[7]
        kinetic = 0.5 * rho * (ux * ux + uy * uy + uz * uz)
        magnetic = 0.5 * (Bx * Bx + By * By + Bz * Bz)
        p = (self.gamma - 1.0) * (U[4] - kinetic - magnetic)
        return rho, ux, uy, uz, p, Bx, By, Bz

    def conservative(self, rho, ux, uy, uz, p, Bx, By, Bz):
        E = (p / (self.gamma - 1.0) + 0.5 * rho * (ux * ux + uy * uy + uz * uz)
             + 0.5 * (Bx * Bx + By * By + Bz * Bz))
        parts = [rho, rho * ux, rho * uy, rho * uz, E, Bx, By, Bz]
        if isinstance(rho, torch.Tensor):
            return torch.stack(parts, dim=0)
        return np.array(parts)


def _check_mhd_config(grid, scheme: str, adaptive_config, gravity_config):
    if str(scheme).lower() != "rk4":
        raise NotImplementedError(f"the hip backend implements the RK4 path only (scheme={scheme!r})")
    if adaptive_config and adaptive_config.get("enabled", False):
        raise NotImplementedError("adaptive embedded RK is not part of the hip backend")
    if gravity_config and gravity_config.get("enabled", False):
        raise NotImplementedError("gravity source terms are not part of the hip backend")
    cfg = getattr(grid, "mhd_config", None) or {}
    if float(cfg.get("resistivity", 0.0)) != 0.0 or float(cfg.get("viscosity", 0.0)) != 0.0:
        raise NotImplementedError("resistive / viscous MHD is not part of the hip backend")
    if str(cfg.get("divergence_clean", "projection")) != "projection" or \
            not bool(cfg.get("project_each_substage", True)):
        raise NotImplementedError("the hip backend projects B after every substage "
                                  "(divergence_clean='projection', project_each_substage=True)")
    if float(cfg.get("rho_min", 1e-10)) != 1e-10 or float(cfg.get("p_min", 1e-12)) != 1e-12:
        raise NotImplementedError("the hip backend clamps at rho_min=1e-10, p_min=1e-12")
    fp = grid.filter_params or {}
    if int(fp.get("interval", 1)) != 1:
        raise NotImplementedError("the hip backend filters every step (interval 1)")


class SpectralSolverMHD3D:
    """phrike/solver.py:840-1163 (3D), reference defaults."""

    def __init__(self, grid: Grid3D, equations: MHDEquations3D, U0=None, scheme: str = "rk4",
                 cfl: float = 0.3, adaptive_config: Optional[Dict] = None,
                 artificial_viscosity_config: Optional[Dict] = None,
                 gravity_config: Optional[Dict] = None):
        _check_mhd_config(grid, scheme, adaptive_config, gravity_config)
        self.grid = grid
        self.equations = equations
        self.scheme = "rk4"
        self.cfl = float(cfl)
        self.t = 0.0
        self.U = U0
        self.dim = 3
        self.adaptive_enabled = False
        h = _lib.c_void_p()
        fon, fp, fa = grid._filter
        with torch.cuda.device(grid.device):
            _lib.call("ph_planmhd3d_create", _lib.ctypes.byref(h), grid.Nx, grid.Ny, grid.Nz,
                      grid.Lx, grid.Ly, grid.Lz, grid._dtype_code, int(grid.dealias), fon, fp, fa,
                      float(equations.gamma))
        self._plan = h
        n = _lib.c_size_t()
        _lib.call("ph_planmhd3d_workspace_bytes", h, _lib.ctypes.byref(n))
        self._ws = torch.empty(n.value, dtype=torch.uint8, device=grid.device)
        self._scalars = torch.zeros(8, dtype=torch.float64, device=grid.device)
        self._sol_ref = None  # weak reference to the last tensor known to hold a solenoidal B
        self._sol_version = -1

    def _mark_solenoidal(self, T) -> None:
        self._sol_ref = weakref.ref(T)
        self._sol_version = T._version

    def _is_solenoidal(self, U) -> bool:
        t = self._sol_ref() if self._sol_ref is not None else None
        return t is U and U._version == self._sol_version

    def __del__(self):
        try:
            _lib.load().ph_planmhd3d_destroy(self._plan)
        except Exception:
            pass

    def _st(self):
        return _lib.stream_handle(torch, self.grid.device)

    def _state(self, U):
        g = self.grid
        if not isinstance(U, torch.Tensor) or U.device.type != "cuda":
            raise TypeError("U must be a GPU tensor (backend='hip')")
        if tuple(U.shape) != (8, g.Nz, g.Ny, g.Nx) or U.dtype != g.dtype:
            raise ValueError(f"U must be (8, {g.Nz}, {g.Ny}, {g.Nx}) {g.dtype}, got "
                             f"{tuple(U.shape)} {U.dtype}")
        return U.contiguous()

    # --- physics -------------------------------------------------------------
    def _compute_rhs(self, U):
        U = self._state(U)
        out = torch.empty_like(U)
        _lib.call("ph_rhs_mhd3d", self._plan, _lib.ptr(U), _lib.ptr(out), _lib.ptr(self._ws), self._st())
        return out

    def _project_B(self, U):
        """B projected (Nyquist planes zeroed), E at fixed thermal pressure;
        returns a new tensor (phrike/solver.py:966-986)."""
        V = self._state(U).clone()
        _lib.call("ph_project_mhd3d", self._plan, _lib.ptr(V), _lib.ptr(self._ws), self._st())
        self._mark_solenoidal(V)
        return V

    def max_div_b(self, U) -> float:
        """max |dx Bx + dy By + dz Bz| with the masked derivatives
        (phrike/solver.py:1008-1015, phrike/grid.py:1635-1646)."""
        U = self._state(U)
        g = self.grid
        d = g.dx1(U[5]) + g.dy1(U[6]) + g.dz1(U[7])
        return float(d.abs().max().item())

    def max_wave_speed(self, U) -> float:
        U = self._state(U)
        _lib.call("ph_max_wave_speed_mhd3d", self._plan, _lib.ptr(U), _lib.ptr(self._scalars), None,
                  self.cfl, _lib.ptr(self._ws), self._st())
        return float(self._scalars[0].item())

    def compute_dt(self, U) -> float:
        """cfl*min(dx,dy,dz)/max_speed (phrike/solver.py:1018-1035, no diffusion)."""
        s = self.max_wave_speed(U)
        if s <= 0.0:
            return 1e-6
        return self.cfl * min(self.grid.dx, self.grid.dy, self.grid.dz) / s

    def conserved_quantities(self, U) -> Dict[str, float]:
        U = self._state(U)
        _lib.call("ph_conserved_sums_mhd3d", self._plan, _lib.ptr(U), _lib.ptr(self._scalars),
                  _lib.ptr(self._ws), self._st())
        v = self._scalars[:7].cpu().numpy()
        keys = ("mass", "momentum_x", "momentum_y", "momentum_z", "energy", "magnetic_energy",
                "cross_helicity")
        return {k: float(x) for k, x in zip(keys, v)}

    # --- time stepping ---------------------------------------------------------
    def step(self, U, dt: float):
        U = self._state(U)
        out = torch.empty_like(U)
        self.step_into(U, out, dt)
        return out

    _fixed_step = step

    def step_into(self, U, out, dt: float) -> None:
        """out = one step of U with caller-owned output (checked: same shape,
        dtype and device as U, contiguous, no shared memory).  When U is this
        solver's own previous output or projected state, unmodified since
        (same tensor object, same version counter), the step takes the fused
        form (PH_STEP_SOLENOIDAL)."""
        from .solver import _check_pair
        U = self._state(U)
        _check_pair(U, out, "out")
        flags = _lib.PH_STEP_SOLENOIDAL if self._is_solenoidal(U) else 0
        _lib.call("ph_step_mhd3d_ex", self._plan, _lib.ptr(U), _lib.ptr(out), float(dt), flags,
                  _lib.ptr(self._ws), self._st())
        self._mark_solenoidal(out)

    def run(self, U0, t0: float, t_end: float, output_interval: float = 0.1,
            checkpoint_interval: float = 0.0, outdir: Optional[str] = None,
            on_output: Optional[Callable] = None, on_step: Optional[Callable] = None,
            tracers: Optional[Any] = None) -> Dict[str, List[float]]:
        """phrike/solver.py:1095-1163: the initial state is projected, dt from
        the CFL condition after every step, the last step clipped to t_end,
        conserved quantities and max |div B| recorded at t0, every
        output_interval and at the end.  Callbacks get a copy of the state
        (the loop reuses two buffers)."""
        if tracers is not None:
            raise NotImplementedError("tracers are not part of the hip backend")
        self.U = self._project_B(U0)
        self.t = float(t0)
        next_output = self.t + output_interval
        next_ckpt = self.t + checkpoint_interval if checkpoint_interval and checkpoint_interval > 0 else np.inf
        keys = ["time", "mass", "momentum_x", "momentum_y", "momentum_z", "energy", "magnetic_energy",
                "cross_helicity", "div_b_max"]
        history: Dict[str, List[float]] = {k: [] for k in keys}

        def record():
            cons = self.conserved_quantities(self.U)
            history["time"].append(self.t)
            for k in keys[1:8]:
                history[k].append(cons[k])
            history["div_b_max"].append(self.max_div_b(self.U))

        record()
        if on_output is not None:
            on_output(self.t, self.U.clone())
        dt = self.compute_dt(self.U)
        step_count = 0
        buf = torch.empty_like(self.U)
        while self.t < t_end - 1e-12:
            dt = min(dt, t_end - self.t)
            self.step_into(self.U, buf, dt)
            self.U, buf = buf, self.U
            self.t += dt
            if on_step is not None:
                on_step(step_count, dt, self.U.clone())
            dt = self.compute_dt(self.U)
            step_count += 1
            if self.t + 1e-12 >= next_output:
                record()
                if on_output is not None:
                    on_output(self.t, self.U.clone())
                next_output += output_interval
            if outdir and self.t + 1e-12 >= next_ckpt:
                from .io import save_solution_snapshot
                save_solution_snapshot(outdir, self.t, U=self.U, grid=self.grid, equations=self.equations)
                next_ckpt += checkpoint_interval
        record()
        if on_output is not None:
            on_output(self.t, self.U.clone())
        return history
