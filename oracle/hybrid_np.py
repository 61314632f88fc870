"""Oracle (test infrastructure, never the product): NumPy FP64 restatement of
the hybrid 2D Euler path, Fourier along x and Legendre-Gauss-Lobatto
collocation along y.

The reference has no working composition of this grid (its hybrid Grid2D
uses Gauss nodes, an ad-hoc derivative matrix and FFTs across the
non-periodic axis: phrike/grid.py:738-772, 852-891, 1231-1256), so the path
is defined from the reference's working pieces, each restated here:

  dx f  = Re IFFT( i k * mask * sigma * FFT(f) ) per row   Grid1D.dx1,
          phrike/grid.py:512-519, tables phrike/grid.py:250-266 (a full
          complex transform; the Nyquist term of i k F is imaginary and
          drops out with .real)
  D f   = sum_j D[i, j] f[j, x]                            LegendreLobattoBasis1D
          D_x along y (phrike/basis/legendre.py:714-717, 1020-1028)
  rhs   = -(dx Fx + D Fy)                                   fluxes
          phrike/equations.py:104-134
  RK4 with the positivity clamp after each stage and at the end
          (phrike/solver.py:467-513), then the filter: x mask*sigma
          (Grid1D.apply_spectral_filter, phrike/grid.py:533-538), then the
          LGL nodal filter along y (phrike/basis/legendre.py:817-872)
  dt    = cfl * min(dx, dy_min) / max(|ux| + a)          phrike/solver.py:613-620
  walls (y_bc="wall"): the 1D Legendre grid's wall rule along y
          (phrike/grid.py:621-738): my = 0 on the boundary rows, rho and E
          copied from the adjacent row; applied to every RHS input, in place
          as phrike/solver.py:64-68 does (so the step's base state carries
          it), and after the filter (phrike/solver.py:120-130).  y_bc="none"
          is the reference hybrid grid's no-op (phrike/grid.py:1082-1084).

Pinning (tests/test_oracle_hybrid.py): dx against closed-form derivatives of
trigonometric polynomials and against the Fourier Grid2D oracle (pinned by
the reference's goldens) on rows; D against exact polynomial derivatives; the
nodal filter against its modal definition; the wall rule entry by entry; the
RK4/clamp/filter composition against the Fourier 2D solver oracle on states
uniform along y.  The D Fy leg on y-varying states: the RHS of an entropy
wave (rho = 1 + eps P(y) + delta sin(kx) at constant u, v, p) and of
x-uniform states whose fluxes are polynomials of degree < Ny (with and
without boundary rows for the wall rule to overwrite) against exact
derivatives, and one RK4 step of the entropy wave against its exact RK4
polynomial.  The wall-rule composition (rule applied to every RHS input, to
the step's base state in place, and after the filter) against the
reference's own 1D Legendre solver (Grid1D basis="legendre", bc="wall",
_compute_rhs and _rk4_step; golden tests/golden/wall1d.npz), of which an
x-uniform hybrid state with zero x momentum is the restriction.
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np

from .gll_np import LGLBasisOracle
from .spectral_np import (Euler2DOracle, _filter_settings, clamp2d, dealias_factor, filter_factor,
                          wavenumbers)


class HybridGridOracle:
    """x: Fourier on [0, lx) with nx points; y: LGL nodes on [y0, y0 + ly]."""

    def __init__(self, nx: int, ny: int, lx: float = 1.0, ly: float = 1.0, y0: float = -0.5,
                 dealias: bool = True, filter_params: Optional[Dict] = None, y_bc: str = "wall"):
        self.y_bc = y_bc
        self.nx, self.ny, self.lx, self.ly = nx, ny, float(lx), float(ly)
        self.x = np.linspace(0.0, lx, nx, endpoint=False)
        self.basis = LGLBasisOracle(ny, ly)
        self.y = self.basis.x + y0
        self.dx = self.lx / nx
        self.dy_min = float(np.min(np.diff(self.basis.x)))
        self.k = wavenumbers(nx, lx)
        self.mask = dealias_factor(nx, dealias)
        on, p, alpha = _filter_settings(filter_params)
        self.filter_on = on
        self.sigma = filter_factor(nx, p, alpha) if on else np.ones(nx)
        self.D = self.basis.D_x
        self.Fn = self.basis.nodal_filter(p, alpha) if on else np.eye(ny)

    def dx1(self, f):
        F = np.fft.fft(f, axis=-1)
        F = F * self.mask
        F = F * self.sigma
        return np.fft.ifft(1j * self.k * F, axis=-1).real

    def dy1(self, f):
        return np.einsum("ij,...jx->...ix", self.D, f)

    def walls(self, U):
        """phrike/grid.py:621-738 wall rule along y (returns a new array)."""
        if self.y_bc != "wall":
            return U
        U = np.array(U, copy=True)
        U[0, 0] = U[0, 1]
        U[0, -1] = U[0, -2]
        U[2, 0] = 0.0
        U[2, -1] = 0.0
        U[3, 0] = U[3, 1]
        U[3, -1] = U[3, -2]
        return U

    def filter(self, U):
        F = np.fft.fft(U, axis=-1) * self.mask * self.sigma
        Ux = np.fft.ifft(F, axis=-1).real
        return np.einsum("ij,...jx->...ix", self.Fn, Ux)


def rhs_hybrid(grid: HybridGridOracle, eqs: Euler2DOracle, U):
    Fx, Fy = eqs.flux(grid.walls(U))
    return -(grid.dx1(Fx) + grid.dy1(Fy))


def rk4_hybrid(grid: HybridGridOracle, eqs: Euler2DOracle, U, dt: float):
    U = grid.walls(U)  # the reference applies the BC to the step's state in place
    k1 = rhs_hybrid(grid, eqs, U)
    U1 = clamp2d(U + 0.5 * dt * k1, eqs)
    k2 = rhs_hybrid(grid, eqs, U1)
    U2 = clamp2d(U + 0.5 * dt * k2, eqs)
    k3 = rhs_hybrid(grid, eqs, U2)
    U3 = clamp2d(U + dt * k3, eqs)
    k4 = rhs_hybrid(grid, eqs, U3)
    return clamp2d(U + (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4), eqs)


class HybridSolverOracle:
    def __init__(self, grid: HybridGridOracle, eqs: Euler2DOracle, cfl: float = 0.4):
        self.grid, self.eqs, self.cfl = grid, eqs, cfl

    def compute_dt(self, U) -> float:
        h = min(self.grid.dx, self.grid.dy_min)
        return self.cfl * h / self.eqs.max_wave_speed(U)

    def step(self, U, dt: float, apply_filter: bool = True):
        Un = rk4_hybrid(self.grid, self.eqs, U, dt)
        if apply_filter and self.grid.filter_on:
            Un = self.grid.filter(Un)
        return self.grid.walls(Un)
