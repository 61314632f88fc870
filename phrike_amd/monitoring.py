"""Run monitoring of the ``hip`` backend: the statistics of
BaseProblem.compute_conservation_errors / compute_velocity_stats /
compute_density_stats / compute_mhd_stats (phrike/problems/base.py:490-822),
formed from one device reduction per call (ph_monitor3d / ph_monitor2d /
ph_monitor_mhd3d) and the conserved sums.

The reference computes these with torch when the state is a tensor; the
results here follow that branch: ``rho_std`` is the unbiased (N - 1)
standard deviation of torch.std, and density statistics carry no pressure
entries.  ``Monitor.on_step`` is a run-loop callback that records them every
``interval`` steps, as the problem classes do with their monitoring interval.
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _lib


class Monitor:
    """Monitoring statistics of one solver's states (3D Euler, 2D Euler for a
    single run, or 3D ideal MHD)."""

    def __init__(self, solver, interval: int = 1):
        self.solver = solver
        self.interval = max(1, int(interval))
        self.monitoring_initial_values: Optional[Dict[str, float]] = None
        self.records: List[Dict[str, float]] = []
        self._out = torch.zeros(_lib.PH_MON_N, dtype=torch.float64, device=solver.grid.device)
        self._kind = type(solver).__name__

    # -- device reduction --------------------------------------------------
    def _reduce(self, U) -> np.ndarray:
        s = self.solver
        st = _lib.stream_handle(torch, s.grid.device)
        if self._kind == "SpectralSolverMHD3D":
            U = s._state(U)
            _lib.call("ph_monitor_mhd3d", s._plan, _lib.ptr(U), _lib.ptr(self._out), _lib.ptr(s._ws), st)
        elif self._kind == "SpectralSolver3D":
            U = s._state(U)
            plan, ws, _ = s._ctx()
            _lib.call("ph_monitor3d", plan, _lib.ptr(U), _lib.ptr(self._out), ws, st)
        elif self._kind == "SpectralSolver2D":
            Ub = s._runs(U)
            if Ub.shape[0] != 1:
                raise ValueError("Monitor takes one 2D run (4, Ny, Nx)")
            plan, ws, _ = s._ctx()
            _lib.call("ph_monitor2d", plan, 1, _lib.ptr(Ub), _lib.ptr(self._out), ws, st)
        else:
            raise TypeError(f"no monitoring reduction for {self._kind}")
        return self._out.cpu().numpy()

    def _ncells(self) -> int:
        g = self.solver.grid
        return int(np.prod([getattr(g, k) for k in ("Nx", "Ny", "Nz") if hasattr(g, k)]))

    # -- the reference's statistics -----------------------------------------
    def conserved(self, U) -> Dict[str, float]:
        s = self.solver
        if self._kind == "SpectralSolverMHD3D":
            return s.conserved_quantities(U)
        if self._kind == "SpectralSolver3D":
            v = s.conserved_sums(U)
            return dict(zip(("mass", "momentum_x", "momentum_y", "momentum_z", "energy"), map(float, v)))
        # 2D: EulerEquations2D.conserved_quantities, as the reference calls it
        return s.equations.conserved_quantities(s._runs(U)[0])

    def set_initial_values(self, U) -> None:
        self.monitoring_initial_values = self.conserved(U)

    def compute_conservation_errors(self, U) -> Dict[str, float]:
        """(current - initial) / initial (or the difference when the initial
        value is 0); signed for mass and energy, absolute for momenta
        (phrike/problems/base.py:490-664)."""
        if self.monitoring_initial_values is None:
            return {}
        cur = self.conserved(U)
        errors = {}
        for key, init in self.monitoring_initial_values.items():
            if key in cur:
                delta = (cur[key] - init) / init if init != 0 else cur[key] - init
                errors[f"{key}_error"] = abs(delta) if key.startswith("momentum") else delta
        return errors

    def compute_velocity_stats(self, U) -> Dict[str, float]:
        """Density-weighted rms speed, max and min speed
        (phrike/problems/base.py:666-723)."""
        m = self._reduce(U)
        mass = m[_lib.PH_MON_MASS]
        if mass > 0:
            return {"v_rms": float(math.sqrt(m[_lib.PH_MON_RV2] / mass)), "v_max": float(m[_lib.PH_MON_VMAX]),
                    "v_min": float(m[_lib.PH_MON_VMIN])}
        return {"v_rms": 0.0, "v_max": 0.0, "v_min": 0.0}

    def compute_density_stats(self, U) -> Dict[str, float]:
        """Mean, max, min and (unbiased) standard deviation of rho
        (phrike/problems/base.py:725-766, torch branch)."""
        m = self._reduce(U)
        n = self._ncells()
        s1, s2, c = m[_lib.PH_MON_RS1], m[_lib.PH_MON_RS2], m[_lib.PH_MON_SHIFT]
        var = (s2 - s1 * s1 / n) / (n - 1) if n > 1 else float("nan")
        return {"rho_mean": float(c + s1 / n), "rho_max": float(m[_lib.PH_MON_RMAX]),
                "rho_min": float(m[_lib.PH_MON_RMIN]), "rho_std": float(math.sqrt(max(var, 0.0)))}

    def compute_mhd_stats(self, U) -> Dict[str, float]:
        """max |div B|, rms |B|, magnetic energy, plasma beta <p>/<|B|^2/2> and
        cross helicity (phrike/problems/base.py:768-822); empty for Euler."""
        if self._kind != "SpectralSolverMHD3D":
            return {}
        m = self._reduce(U)
        n = self._ncells()
        cons = self.solver.conserved_quantities(U)
        b2 = m[_lib.PH_MON_B2]
        bmag_mean = 0.5 * b2 / n
        p_mean = m[_lib.PH_MON_PSUM] / n
        try:
            div_b = float(self.solver.max_div_b(U))
        except Exception:
            div_b = float("nan")
        return {"max|divB|": div_b, "b_rms": float(math.sqrt(b2 / n)), "magnetic_energy": float(0.5 * b2),
                "plasma_beta": float(p_mean / bmag_mean) if bmag_mean > 0 else float("inf"),
                "cross_helicity": float(cons["cross_helicity"])}

    def stats(self, U) -> Dict[str, float]:
        out = {}
        out.update(self.compute_conservation_errors(U))
        out.update(self.compute_velocity_stats(U))
        out.update(self.compute_density_stats(U))
        out.update(self.compute_mhd_stats(U))
        return out

    def on_step(self, step_count: int, dt: float, U) -> None:
        """Run-loop callback: the statistics every `interval` steps."""
        if self.monitoring_initial_values is None:
            self.set_initial_values(U)
        if step_count % self.interval == 0:
            rec = {"step": int(step_count), "dt": float(dt), "t": float(self.solver.t)}
            rec.update(self.stats(U))
            self.records.append(rec)
