"""numpy/scipy restatement of the reference's Fourier pseudo-spectral Euler path.

TEST INFRASTRUCTURE ONLY (the oracle).  The shipped ``phrike_amd`` package never
imports this module.

Every operator keeps the reference's *mathematical* operation order, so that a
GPU result can be compared with it at round-off level:

* full complex ``fftn`` per derivative (not ``rfftn``), then ``*= mask``,
  ``*= sigma``, ``* ik`` and ``ifftn(...).real``
  (phrike/grid.py:1497-1518, 1705-1726);
* tensor-product 2/3 mask and exponential filter
  (phrike/grid.py:87-141);
* fluxes from recomputed momenta (phrike/equations.py:104-196);
* ``rhs = -((dx Fx + dy Fy) + dz Fz)`` (phrike/solver.py:1177-1184);
* RK4 coefficient order and post-step filter (phrike/solver.py:1220-1234,
  1330-1337), 2D positivity clamp after each stage and at the end
  (phrike/solver.py:467-513), 2D filter cadence (phrike/solver.py:639-648);
* CFL dt (phrike/solver.py:1307-1311, 613-620; phrike/equations.py:137-147,
  199-211).

Parity pinning (tests/test_oracle_*.py):
* axis tables, derivatives, filter: closed forms (single Fourier modes, the
  N//3 cutoff, sigma(0)=1, sigma(N/2)=exp(-alpha)) and the reference golden
  vectors in tests/golden/;
* flux / primitive: algebraic identities plus golden vectors;
* RHS, RK4 step, 2D clamp and cadence, compute_dt: golden vectors written from
  the reference (tests/golden/make_golden.py), the linear-acoustics RK4
  amplification factor, and conservation of the k=0 mode.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Dict, Optional, Tuple

import numpy as np

try:  # the reference's default FFT (phrike/grid.py:16-25)
    import scipy.fft as _sfft

    FFT_LIBRARY = "scipy.fft"
except Exception:  # pragma: no cover - scipy is present in this image
    _sfft = None
    FFT_LIBRARY = "numpy.fft"


def default_workers() -> int:
    """Reference default when ``fft_workers`` is unset: all host cores
    (phrike/problems/base.py:193-202)."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except Exception:  # pragma: no cover
        return max(1, os.cpu_count() or 1)


# ----------------------------------------------------------------------------
# 1D factor tables (phrike/grid.py:87-141, 1381-1392)
# ----------------------------------------------------------------------------

def mode_numbers(n: int) -> np.ndarray:
    """Integer mode numbers in FFT order, as float: 0,1,..,n/2-1,-n/2,..,-1."""
    return np.fft.fftfreq(n) * n


def wavenumbers(n: int, length: float) -> np.ndarray:
    """k = 2*pi*fftfreq(n, d=length/n) (phrike/grid.py:1387-1392)."""
    return 2.0 * np.pi * np.fft.fftfreq(n, d=length / n)


def dealias_factor(n: int, dealias: bool) -> np.ndarray:
    """Per-axis 2/3 rule: keep |m| <= n//3 (phrike/grid.py:114-123)."""
    if not dealias:
        return np.ones(n)
    return (np.abs(mode_numbers(n)) <= (n // 3)).astype(float)


def filter_factor(n: int, p: int, alpha: float) -> np.ndarray:
    """Per-axis exponential filter exp(-alpha*(|m|/max(max|m|,1))**p)
    (phrike/grid.py:129-141)."""
    m = np.abs(mode_numbers(n))
    mmax = np.max(m) if n > 0 else 1.0
    eta = m / max(mmax, 1.0)
    return np.exp(-alpha * eta ** p)


def _filter_settings(filter_params: Optional[Dict]) -> Tuple[bool, int, float]:
    if filter_params and bool(filter_params.get("enabled", False)):
        return True, int(filter_params.get("p", 8)), float(filter_params.get("alpha", 36.0))
    return False, 8, 36.0


# ----------------------------------------------------------------------------
# 3D grid (phrike/grid.py:1357-1744)
# ----------------------------------------------------------------------------

@dataclass
class Grid3DOracle:
    Nx: int
    Ny: int
    Nz: int
    Lx: float
    Ly: float
    Lz: float
    dealias: bool = True
    filter_params: Optional[Dict] = None
    fft_workers: Optional[int] = None

    def __post_init__(self) -> None:
        self.dx = self.Lx / self.Nx
        self.dy = self.Ly / self.Ny
        self.dz = self.Lz / self.Nz
        self.x = np.linspace(0.0, self.Lx, self.Nx, endpoint=False)
        self.y = np.linspace(0.0, self.Ly, self.Ny, endpoint=False)
        self.z = np.linspace(0.0, self.Lz, self.Nz, endpoint=False)
        self.kx = wavenumbers(self.Nx, self.Lx)
        self.ky = wavenumbers(self.Ny, self.Ly)
        self.kz = wavenumbers(self.Nz, self.Lz)
        self.ikx = (1j * self.kx)[None, None, :]
        self.iky = (1j * self.ky)[None, :, None]
        self.ikz = (1j * self.kz)[:, None, None]
        mx = dealias_factor(self.Nx, self.dealias)
        my = dealias_factor(self.Ny, self.dealias)
        mz = dealias_factor(self.Nz, self.dealias)
        # outer products in z*y*x order (phrike/grid.py:123,141)
        self.dealias_mask = mz[:, None, None] * my[None, :, None] * mx[None, None, :]
        on, p, alpha = _filter_settings(self.filter_params)
        if on:
            sx = filter_factor(self.Nx, p, alpha)
            sy = filter_factor(self.Ny, p, alpha)
            sz = filter_factor(self.Nz, p, alpha)
            self.filter_sigma = sz[:, None, None] * sy[None, :, None] * sx[None, None, :]
        else:
            self.filter_sigma = np.ones((self.Nz, self.Ny, self.Nx))
        if self.fft_workers is None:
            self.fft_workers = default_workers()

    # transforms --------------------------------------------------------------
    def fftn(self, f: np.ndarray) -> np.ndarray:
        if _sfft is not None:
            return _sfft.fftn(f, axes=(-3, -2, -1), workers=self.fft_workers)
        return np.fft.fftn(f, axes=(-3, -2, -1))

    def ifftn(self, F: np.ndarray) -> np.ndarray:
        if _sfft is not None:
            return _sfft.ifftn(F, axes=(-3, -2, -1), workers=self.fft_workers).real
        return np.fft.ifftn(F, axes=(-3, -2, -1)).real

    def _apply_masks(self, F: np.ndarray) -> np.ndarray:
        F *= self.dealias_mask
        F *= self.filter_sigma
        return F

    def _deriv(self, f: np.ndarray, ik: np.ndarray) -> np.ndarray:
        F = self._apply_masks(self.fftn(f))
        return self.ifftn(F * ik)

    def dx1(self, f):
        return self._deriv(f, self.ikx)

    def dy1(self, f):
        return self._deriv(f, self.iky)

    def dz1(self, f):
        return self._deriv(f, self.ikz)

    def apply_spectral_filter(self, f: np.ndarray) -> np.ndarray:
        return self.ifftn(self._apply_masks(self.fftn(f)))

    def xyz_mesh(self):
        Z, Y, X = np.meshgrid(self.z, self.y, self.x, indexing="ij")
        return X, Y, Z


# ----------------------------------------------------------------------------
# 2D Fourier x Fourier grid (phrike/grid.py:775-1287)
# ----------------------------------------------------------------------------

@dataclass
class Grid2DOracle:
    Nx: int
    Ny: int
    Lx: float
    Ly: float
    dealias: bool = True
    filter_params: Optional[Dict] = None
    fft_workers: Optional[int] = None

    def __post_init__(self) -> None:
        self.dx = self.Lx / self.Nx
        self.dy = self.Ly / self.Ny
        self.dx_min = None
        self.dy_min = None
        self.x = np.linspace(0.0, self.Lx, self.Nx, endpoint=False)
        self.y = np.linspace(0.0, self.Ly, self.Ny, endpoint=False)
        self.kx = wavenumbers(self.Nx, self.Lx)
        self.ky = wavenumbers(self.Ny, self.Ly)
        self.ikx = (1j * self.kx)[None, :]
        self.iky = (1j * self.ky)[:, None]
        mx = dealias_factor(self.Nx, self.dealias)
        my = dealias_factor(self.Ny, self.dealias)
        self.dealias_mask = my[:, None] * mx[None, :]
        on, p, alpha = _filter_settings(self.filter_params)
        if on:
            sx = filter_factor(self.Nx, p, alpha)
            sy = filter_factor(self.Ny, p, alpha)
            self.filter_sigma = sy[:, None] * sx[None, :]
        else:
            self.filter_sigma = np.ones((self.Ny, self.Nx))
        if self.fft_workers is None:
            self.fft_workers = default_workers()

    def fft2(self, f):
        if _sfft is not None:
            return _sfft.fft2(f, axes=(-2, -1), workers=self.fft_workers)
        return np.fft.fft2(f, axes=(-2, -1))

    def ifft2(self, F):
        if _sfft is not None:
            return _sfft.ifft2(F, axes=(-2, -1), workers=self.fft_workers).real
        return np.fft.ifft2(F, axes=(-2, -1)).real

    def _apply_masks(self, F):
        F *= self.dealias_mask
        F *= self.filter_sigma
        return F

    def dx1(self, f):
        return self.ifft2(self._apply_masks(self.fft2(f)) * self.ikx)

    def dy1(self, f):
        return self.ifft2(self._apply_masks(self.fft2(f)) * self.iky)

    def apply_spectral_filter(self, f):
        return self.ifft2(self._apply_masks(self.fft2(f)))

    def xy_mesh(self):
        X, Y = np.meshgrid(self.x, self.y, indexing="xy")
        return X, Y


# ----------------------------------------------------------------------------
# Euler equations (phrike/equations.py:104-211, 322-381, 552-637)
# ----------------------------------------------------------------------------

@dataclass
class Euler3DOracle:
    gamma: float = 1.4

    def primitive(self, U):
        rho = U[0]
        ux = U[1] / rho
        uy = U[2] / rho
        uz = U[3] / rho
        kin = 0.5 * rho * (ux * ux + uy * uy + uz * uz)
        p = (self.gamma - 1.0) * (U[4] - kin)
        return rho, ux, uy, uz, p

    def conservative(self, rho, ux, uy, uz, p):
        E = p / (self.gamma - 1.0) + 0.5 * rho * (ux * ux + uy * uy + uz * uz)
        return np.array([rho, rho * ux, rho * uy, rho * uz, E])

    def flux(self, U):
        rho, ux, uy, uz, p = self.primitive(U)
        mx, my, mz = rho * ux, rho * uy, rho * uz
        Ep = U[4] + p
        Fx = np.array([mx, mx * ux + p, mx * uy, mx * uz, Ep * ux])
        Fy = np.array([my, my * ux, my * uy + p, my * uz, Ep * uy])
        Fz = np.array([mz, mz * ux, mz * uy, mz * uz + p, Ep * uz])
        return Fx, Fy, Fz

    def max_wave_speed(self, U) -> float:
        rho, ux, uy, uz, p = self.primitive(U)
        a = np.sqrt(self.gamma * p / rho)
        c = np.maximum(np.maximum(np.abs(ux), np.abs(uy)), np.abs(uz)) + a
        return float(np.max(c)) if c.size else 0.0

    def conserved_quantities(self, U) -> Dict[str, float]:
        rho, ux, uy, uz, p = self.primitive(U)
        return {
            "mass": float(rho.sum()),
            "momentum_x": float((rho * ux).sum()),
            "momentum_y": float((rho * uy).sum()),
            "momentum_z": float((rho * uz).sum()),
            "energy": float((p / (self.gamma - 1.0)
                             + 0.5 * rho * (ux * ux + uy * uy + uz * uz)).sum()),
        }


@dataclass
class Euler2DOracle:
    gamma: float = 1.4

    def primitive(self, U):
        rho = U[0]
        ux = U[1] / rho
        uy = U[2] / rho
        kin = 0.5 * rho * (ux * ux + uy * uy)
        p = (self.gamma - 1.0) * (U[3] - kin)
        return rho, ux, uy, p

    def conservative(self, rho, ux, uy, p):
        E = p / (self.gamma - 1.0) + 0.5 * rho * (ux * ux + uy * uy)
        return np.array([rho, rho * ux, rho * uy, E])

    def flux(self, U):
        rho, ux, uy, p = self.primitive(U)
        mx, my = rho * ux, rho * uy
        Ep = U[3] + p
        Fx = np.array([mx, mx * ux + p, mx * uy, Ep * ux])
        Fy = np.array([my, my * ux, my * uy + p, Ep * uy])
        return Fx, Fy

    def max_wave_speed(self, U) -> float:
        # the reference's 2D CFL speed uses |ux| + a only (phrike/equations.py:137-147)
        rho, ux, uy, p = self.primitive(U)
        a = np.sqrt(self.gamma * p / rho)
        return float(np.max(np.abs(ux) + a)) if ux.size else 0.0

    def conserved_quantities(self, U) -> Dict[str, float]:
        # phrike/equations.py:372-381 (uses ux only in the kinetic term)
        rho, u, p, _ = self.primitive(U)
        return {
            "mass": float(rho.sum()),
            "momentum": float((rho * u).sum()),
            "energy": float((p / (self.gamma - 1.0) + 0.5 * rho * u * u).sum()),
        }


# ----------------------------------------------------------------------------
# RHS, RK4 and solvers (phrike/solver.py)
# ----------------------------------------------------------------------------

def rhs3d(grid: Grid3DOracle, eqs: Euler3DOracle, U: np.ndarray) -> np.ndarray:
    """phrike/solver.py:1177-1184 (AV and gravity off on this path)."""
    Fx, Fy, Fz = eqs.flux(U)
    return -(grid.dx1(Fx) + grid.dy1(Fy) + grid.dz1(Fz))


def rk4_3d(grid, eqs, U, dt: float) -> np.ndarray:
    """phrike/solver.py:1220-1230."""
    k1 = rhs3d(grid, eqs, U)
    k2 = rhs3d(grid, eqs, U + 0.5 * dt * k1)
    k3 = rhs3d(grid, eqs, U + 0.5 * dt * k2)
    k4 = rhs3d(grid, eqs, U + dt * k3)
    return U + (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)


class Solver3DOracle:
    """Fixed-step RK4 + post-step filter every step (phrike/solver.py:1237-1337)."""

    def __init__(self, grid: Grid3DOracle, equations: Euler3DOracle, cfl: float = 0.25):
        self.grid = grid
        self.equations = equations
        self.cfl = cfl
        self.t = 0.0

    def compute_dt(self, U) -> float:
        s = self.equations.max_wave_speed(U)
        if s <= 0.0:
            return 1e-6
        return self.cfl * min(self.grid.dx, self.grid.dy, self.grid.dz) / s

    def step(self, U, dt: float):
        Un = rk4_3d(self.grid, self.equations, U, dt)
        return self.grid.apply_spectral_filter(Un)


def rhs2d(grid: Grid2DOracle, eqs: Euler2DOracle, U):
    """phrike/solver.py:428-443 (BC hook is a no-op for Fourier grids)."""
    Fx, Fy = eqs.flux(U)
    return -(grid.dx1(Fx) + grid.dy1(Fy))


def clamp2d(U, eqs: Euler2DOracle, rho_min: float = 1e-8, p_min: float = 1e-10):
    """Positivity clamp via prim -> cons round trip (phrike/solver.py:467-488)."""
    rho, ux, uy, p = eqs.primitive(U)
    return eqs.conservative(np.maximum(rho, rho_min), ux, uy, np.maximum(p, p_min))


def rk4_2d(grid, eqs, U, dt: float):
    """phrike/solver.py:502-513."""
    k1 = rhs2d(grid, eqs, U)
    U2 = clamp2d(U + 0.5 * dt * k1, eqs)
    k2 = rhs2d(grid, eqs, U2)
    U3 = clamp2d(U + 0.5 * dt * k2, eqs)
    k3 = rhs2d(grid, eqs, U3)
    U4 = clamp2d(U + dt * k3, eqs)
    k4 = rhs2d(grid, eqs, U4)
    return clamp2d(U + (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4), eqs)


class Solver2DOracle:
    """Fixed-step RK4 with clamp, filter at cadence (phrike/solver.py:527-812)."""

    def __init__(self, grid: Grid2DOracle, equations: Euler2DOracle, cfl: float = 0.3):
        self.grid = grid
        self.equations = equations
        self.cfl = cfl
        self.t = 0.0
        self._step_counter = 0
        self.filter_interval = 1
        fp = grid.filter_params
        if isinstance(fp, dict):
            try:
                self.filter_interval = max(1, int(fp.get("interval", 1)))
            except Exception:
                self.filter_interval = 1

    def compute_dt(self, U) -> float:
        s = self.equations.max_wave_speed(U)
        if s <= 0.0:
            return 1e-6
        dxe = self.grid.dx_min or self.grid.dx
        dye = self.grid.dy_min or self.grid.dy
        return self.cfl * min(dxe, dye) / s

    def step(self, U, dt: float):
        Un = rk4_2d(self.grid, self.equations, U, dt)
        if self._step_counter % self.filter_interval == 0:
            Un = self.grid.apply_spectral_filter(Un)
        self._step_counter += 1
        return Un

    def run_to(self, U0, t_end: float, t0: float = 0.0, max_steps: Optional[int] = None):
        """Run-loop semantics of phrike/solver.py:764-797 without I/O: dt from
        CFL, recomputed after each step; the last step clipped to t_end."""
        U = U0.copy()
        self.t = float(t0)
        dt = self.compute_dt(U)
        n = 0
        while self.t < t_end - 1e-12:
            dt = min(dt, t_end - self.t)
            U = self.step(U, dt)
            self.t += dt
            dt = self.compute_dt(U)
            n += 1
            if max_steps is not None and n >= max_steps:
                break
        return U, n
