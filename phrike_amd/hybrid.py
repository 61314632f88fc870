"""2D Euler on a hybrid grid: Fourier along x, Legendre-Gauss-Lobatto along y.

The reference's Grid2D(basis_y="legendre") cannot serve here (Gauss nodes, an
ad-hoc derivative matrix and FFTs across the non-periodic axis:
phrike/grid.py:738-772, 852-891, 1231-1256), so this grid composes the
reference's working 1D pieces:

* x: Grid1D's Fourier derivative and filter per row (phrike/grid.py:250-266,
  512-538): FFT, times dealias mask and sigma (and i k), inverse FFT, real part;
* y: LegendreLobattoBasis1D on [y0, y0 + Ly] (phrike/basis/legendre.py:681-729):
  collocation derivative D applied along y on matrix cores, and the nodal
  filter (phrike/basis/legendre.py:817-872) after each step;
* rhs = -(dx Fx + D Fy); RK4 with positivity clamp (phrike/solver.py:467-513);
  dt = cfl * min(dx, dy_min) / max(|ux| + a) (phrike/solver.py:613-620);
* y walls (``y_bc="wall"``, the default): the 1D Legendre grid's wall rule
  (phrike/grid.py:621-738) along y -- my = 0 on the two boundary rows, rho and
  E copied from the adjacent row -- applied to every RHS input (the step's
  base state included, as the reference's in-place BC at
  phrike/solver.py:64-68) and after the filter (phrike/solver.py:120-130).
  ``y_bc="none"`` is the reference hybrid grid's no-op
  (phrike/grid.py:1082-1084); collocation without walls is unstable for KHI
  (the boundary rows blow up within ~100 steps at 256 x 128).

``gll`` selects the D application: "f64" (FP64 MFMA) or "split" (FP32 MFMA on
hi + lo parts of D, and of FP64 data).  precision="single" keeps the state in
FP32 and always uses the split form.
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np
import torch

from . import _lib
from .basis import LegendreLobattoBasis1D
from .equations import EulerEquations2D


def _filter_cfg(fp):
    if fp and bool(fp.get("enabled", False)):
        return 1, int(fp.get("p", 8)), float(fp.get("alpha", 36.0))
    return 0, 8, 36.0


class HybridGrid2D:
    """Arrays (..., Ny, Nx): x periodic on [0, Lx), y on LGL nodes in [y0, y0+Ly]."""

    def __init__(self, Nx: int, Ny: int, Lx: float = 1.0, Ly: float = 1.0, y0: float = -0.5,
                 dealias: bool = True, filter_params: Optional[Dict] = None,
                 precision: str = "double", gll: str = "f64", torch_device: Optional[str] = None,
                 batch: int = 1, backend: str = "hip", y_bc: str = "wall"):
        if backend != "hip":
            raise ValueError("HybridGrid2D is the hip backend's grid")
        self.Nx, self.Ny = int(Nx), int(Ny)
        self.Lx, self.Ly, self.y0 = float(Lx), float(Ly), float(y0)
        self.dealias = bool(dealias)
        self.filter_params = filter_params
        self.precision = "single" if precision == "single" else "double"
        self.dtype = torch.float32 if self.precision == "single" else torch.float64
        if gll not in ("f64", "split"):
            raise ValueError("gll must be 'f64' or 'split'")
        if self.precision == "single":
            gll = "split"
        self.gll = gll
        if y_bc not in ("wall", "none"):
            raise ValueError("y_bc must be 'wall' or 'none'")
        self.y_bc = y_bc
        self.batch = int(batch)
        self.torch_device = torch_device or "cuda"
        self.device = torch.device(self.torch_device)
        self.basis = LegendreLobattoBasis1D(self.Ny, self.Ly, torch_device=self.torch_device)
        yn = self.basis.nodes() + self.y0
        self.dx = self.Lx / self.Nx
        self.dy_min = float(np.min(np.diff(self.basis.nodes())))
        self.x = torch.tensor(np.linspace(0.0, self.Lx, self.Nx, endpoint=False), dtype=self.dtype,
                              device=self.device)
        self.y = torch.tensor(yn, dtype=self.dtype, device=self.device)
        self.y_nodes = yn
        self._filter = _filter_cfg(filter_params)
        fon, fp, fa = self._filter
        self.D = np.ascontiguousarray(self.basis.D_x)
        self.Fn = np.ascontiguousarray(self.basis.nodal_filter(fp, fa) if fon else np.eye(self.Ny))
        self._plans = {}
        self._ws = None

    def _mode(self) -> int:
        if self.precision == "single":
            return _lib.PH_GLL_SPLIT_F32
        return _lib.PH_GLL_F64 if self.gll == "f64" else _lib.PH_GLL_SPLIT_F64IO

    def _plan(self, gamma: float = 1.4):
        g = float(gamma)
        if g not in self._plans:
            h = _lib.c_void_p()
            fon, fp, fa = self._filter
            with torch.cuda.device(self.device):
                _lib.call("ph_planh_create", _lib.ctypes.byref(h), self.Nx, self.Ny, self.Lx, self.dy_min,
                          0 if self.dtype == torch.float64 else 1, int(self.dealias), fon, fp, fa, g,
                          self.batch, self.D.ctypes.data, self.Fn.ctypes.data, self._mode(),
                          int(self.y_bc == "wall"))
            self._plans[g] = h
        return self._plans[g]

    def workspace(self, gamma: float = 1.4):
        plan = self._plan(gamma)
        n = _lib.c_size_t()
        _lib.call("ph_planh_workspace_bytes", plan, _lib.ctypes.byref(n))
        if self._ws is None or self._ws.numel() < n.value:
            self._ws = torch.empty(n.value, dtype=torch.uint8, device=self.device)
        return self._ws

    def xy_mesh(self):
        Y, X = torch.meshgrid(self.y, self.x, indexing="ij")
        return X, Y

    def apply_boundary_conditions(self, U, t=None):
        return U

    def __del__(self):
        try:
            lib = _lib.load()
            for h in getattr(self, "_plans", {}).values():
                lib.ph_planh_destroy(h)
        except Exception:
            pass


class HybridSolver2D:
    """Fixed-step RK4 + clamp + post-step filter on HybridGrid2D; one run
    (4, Ny, Nx) or a batch (B, 4, Ny, Nx) with B <= grid.batch."""

    def __init__(self, grid: HybridGrid2D, equations: EulerEquations2D, cfl: float = 0.4):
        self.grid = grid
        self.equations = equations
        self.cfl = float(cfl)
        self.t = 0.0
        self._plan = grid._plan(equations.gamma)

    def _ctx(self):
        return self._plan, _lib.ptr(self.grid.workspace(self.equations.gamma)), \
            _lib.stream_handle(torch, self.grid.device)

    def _runs(self, U):
        g = self.grid
        if not isinstance(U, torch.Tensor) or U.device.type != "cuda":
            raise TypeError("U must be a GPU tensor (backend='hip')")
        if U.dtype != g.dtype or tuple(U.shape[-3:]) != (4, g.Ny, g.Nx) or U.dim() not in (3, 4):
            raise ValueError(f"U must be ([B,] 4, {g.Ny}, {g.Nx}) {g.dtype}, got {tuple(U.shape)} {U.dtype}")
        Ub = U.reshape(-1, 4, g.Ny, g.Nx).contiguous()
        if Ub.shape[0] > g.batch:
            raise ValueError(f"batch {Ub.shape[0]} exceeds the grid's plan batch {g.batch}")
        return Ub

    def rhs(self, U):
        Ub = self._runs(U)
        out = torch.empty_like(Ub)
        plan, ws, st = self._ctx()
        _lib.call("ph_rhs2dh", plan, Ub.shape[0], _lib.ptr(Ub), _lib.ptr(out), ws, st)
        return out.reshape(U.shape)

    def max_wave_speed_runs(self, U) -> torch.Tensor:
        plan, ws, st = self._ctx()
        Ub = self._runs(U)
        out = torch.empty(Ub.shape[0], dtype=torch.float64, device=self.grid.device)
        _lib.call("ph_max_wave_speed2dh", plan, Ub.shape[0], _lib.ptr(Ub), _lib.ptr(out), None,
                  self.cfl, ws, st)
        return out

    def dt_runs(self, U, out=None) -> torch.Tensor:
        plan, ws, st = self._ctx()
        Ub = self._runs(U)
        if out is None:
            out = torch.empty(Ub.shape[0], dtype=torch.float64, device=self.grid.device)
        _lib.call("ph_max_wave_speed2dh", plan, Ub.shape[0], _lib.ptr(Ub), None, _lib.ptr(out),
                  self.cfl, ws, st)
        return out

    def compute_dt(self, U) -> float:
        s = float(self.max_wave_speed_runs(U).max().item())
        if s <= 0.0:
            return 1e-6
        return self.cfl * min(self.grid.dx, self.grid.dy_min) / s

    def step_into(self, Ub, out, dt: float, dt_dev=None, active_dev=None, apply_filter: bool = True) -> None:
        """out = one step of every run of Ub (B, 4, Ny, Nx), caller-owned; runs
        with active_dev == 0 are left untouched in out."""
        from .solver import _check_pair, _check_scalars
        g = self.grid
        if not isinstance(Ub, torch.Tensor) or Ub.dim() != 4 or tuple(Ub.shape[1:]) != (4, g.Ny, g.Nx) \
                or Ub.dtype != g.dtype or Ub.device.type != "cuda" or not Ub.is_contiguous() \
                or not 1 <= Ub.shape[0] <= g.batch:
            raise ValueError(f"Ub must be a contiguous (B <= {g.batch}, 4, {g.Ny}, {g.Nx}) {g.dtype} GPU tensor")
        _check_pair(Ub, out, "out")
        _check_scalars(dt_dev, Ub.shape[0], torch.float64, Ub.device, "dt_dev")
        _check_scalars(active_dev, Ub.shape[0], torch.int32, Ub.device, "active_dev")
        plan, ws, st = self._ctx()
        _lib.call("ph_step2dh", plan, Ub.shape[0], _lib.ptr(Ub), _lib.ptr(out), float(dt), _lib.ptr(dt_dev),
                  _lib.ptr(active_dev), int(apply_filter), ws, st)

    def step(self, U, dt: float, apply_filter: bool = True):
        Ub = self._runs(U)
        out = torch.empty_like(Ub)
        self.step_into(Ub, out, dt, apply_filter=apply_filter)
        return out.reshape(U.shape)

    def advance(self, U, nsteps: int, dt: Optional[float] = None, graph_steps: int = 0):
        """nsteps RK4 steps; dt fixed when given, else the CFL dt recomputed on
        the device before every step (no host synchronisation).  With
        graph_steps > 0 the steps are launched as replays of a captured HIP
        graph of graph_steps (rounded up to even) steps, the remainder eagerly:
        the same kernels and arguments, without the per-step host launch cost
        (one step at the C4 sizes is ~10 latency-bound launches)."""
        Ub = self._runs(U)
        a, b = Ub.clone(), torch.empty_like(Ub)
        dtv = None if dt is not None else torch.empty(Ub.shape[0], dtype=torch.float64, device=self.grid.device)

        def steps(k, x, y):
            for _ in range(k):
                if dtv is not None:
                    self.dt_runs(x, out=dtv)
                self.step_into(x, y, 0.0 if dt is None else dt, dt_dev=dtv)
                x, y = y, x
            return x, y

        G = graph_steps + (graph_steps % 2)
        if G > 0 and nsteps >= G:
            dev = self.grid.device
            torch.cuda.synchronize(dev)
            g = torch.cuda.CUDAGraph()
            cs = torch.cuda.Stream(dev)
            with torch.cuda.graph(g, stream=cs):
                steps(G, a, b)  # an even count: the state ends in a again
            torch.cuda.synchronize(dev)
            for _ in range(nsteps // G):
                g.replay()
            torch.cuda.current_stream(dev).wait_stream(cs)
            nsteps %= G
        a, b = steps(nsteps, a, b)
        return a.reshape(U.shape)

    def conserved_sums(self, U) -> np.ndarray:
        """Per run, in FP64: [mass, x-momentum, y-momentum, energy] as
        sum_x dx * sum_y w_y U (LGL quadrature along y, rectangle rule along
        the periodic x)."""
        Ub = self._runs(U).double()
        w = torch.as_tensor(self.grid.basis.quadrature_weights(), dtype=torch.float64, device=Ub.device)
        return (torch.einsum("bcyx,y->bc", Ub, w) * (self.grid.Lx / self.grid.Nx)).cpu().numpy()
