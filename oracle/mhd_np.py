"""numpy/scipy restatement of the reference's 3D ideal-MHD pseudo-spectral path.

TEST INFRASTRUCTURE ONLY (the oracle).  The shipped ``phrike_amd`` package never
imports this module.

Follows, in the reference's operation order:

* MHDEquations3D primitive / conservative / flux / max_wave_speed /
  conserved_quantities (phrike/equations.py:911-1100; CFL 1041-1061);
* Grid3D.project_solenoidal (phrike/grid.py:1648-1684): full complex fftn of
  B, Nyquist planes zeroed, B^ -= k (k.B^)/k^2 (k^2 = 1 at k = 0), ifftn;
  Grid3D.divergence (phrike/grid.py:1635-1646);
* SpectralSolverMHD in 3D with the reference defaults (resistivity and
  viscosity 0, projection after every substage, filter interval 1;
  phrike/solver.py:840-1050): rhs = -(dx Fx + dy Fy + dz Fz) with the masked,
  filtered derivatives of the Euler path; RK4 stages U_s = proj(clamp(U + c dt
  k)); Un = proj(clamp(U + dt/6 (k1 + 2 k2 + 2 k3 + k4))); then the spectral
  filter and proj again; proj recomputes E from the projected B at fixed
  thermal pressure; clamp rho >= 1e-10, p >= 1e-12 through prim -> cons.

Parity pinning: golden vectors written from the reference
(tests/golden/make_golden.py, mhd3d.npz) and closed forms
(tests/test_oracle_mhd.py: the exact circularly polarised Alfven wave,
div B = 0 after projection, projection idempotent and the identity on
solenoidal fields, fluxes reducing to the Euler fluxes at B = 0).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict

import numpy as np

from .spectral_np import Grid3DOracle

RHO_MIN = 1e-10  # phrike/solver.py:877 (mhd_config default)
P_MIN = 1e-12    # phrike/solver.py:878


@dataclass
class MHD3DOracle:
    gamma: float = 5.0 / 3.0

    def primitive(self, U):
        rho = U[0]
        ux = U[1] / rho
        uy = U[2] / rho
        uz = U[3] / rho
        E = U[4]
        Bx, By, Bz = U[5], U[6], U[7]
        kinetic = 0.5 * rho * (ux * ux + uy * uy + uz * uz)
        magnetic = 0.5 * (Bx * Bx + By * By + Bz * Bz)
        p = (self.gamma - 1.0) * (E - kinetic - magnetic)
        return rho, ux, uy, uz, p, Bx, By, Bz

    def conservative(self, rho, ux, uy, uz, p, Bx, By, Bz):
        E = (p / (self.gamma - 1.0) + 0.5 * rho * (ux * ux + uy * uy + uz * uz)
             + 0.5 * (Bx * Bx + By * By + Bz * Bz))
        return np.array([rho, rho * ux, rho * uy, rho * uz, E, Bx, By, Bz])

    def flux(self, U):
        rho, ux, uy, uz, p, Bx, By, Bz = self.primitive(U)
        E = U[4]
        uB = ux * Bx + uy * By + uz * Bz
        pstar = p + 0.5 * (Bx * Bx + By * By + Bz * Bz)
        mx, my, mz = rho * ux, rho * uy, rho * uz
        zero = np.zeros_like(rho)
        Fx = np.array([mx, mx * ux + pstar - Bx * Bx, mx * uy - Bx * By, mx * uz - Bx * Bz,
                       (E + pstar) * ux - Bx * uB, zero, ux * By - uy * Bx, ux * Bz - uz * Bx])
        Fy = np.array([my, my * ux - By * Bx, my * uy + pstar - By * By, my * uz - By * Bz,
                       (E + pstar) * uy - By * uB, uy * Bx - ux * By, zero, uy * Bz - uz * By])
        Fz = np.array([mz, mz * ux - Bz * Bx, mz * uy - Bz * By, mz * uz + pstar - Bz * Bz,
                       (E + pstar) * uz - Bz * uB, uz * Bx - ux * Bz, uz * By - uy * Bz, zero])
        return Fx, Fy, Fz

    def max_wave_speed(self, U) -> float:
        rho, ux, uy, uz, p, Bx, By, Bz = self.primitive(U)
        a2 = self.gamma * p / rho
        b2 = (Bx * Bx + By * By + Bz * Bz) / rho
        fast = np.sqrt(np.maximum(a2 + b2, 0.0))
        comp = np.maximum(np.maximum(np.abs(ux), np.abs(uy)), np.abs(uz)) + fast
        return float(np.max(comp))

    def conserved_quantities(self, U) -> Dict[str, float]:
        rho, ux, uy, uz, p, Bx, By, Bz = self.primitive(U)
        b2 = Bx * Bx + By * By + Bz * Bz
        total = p / (self.gamma - 1.0) + 0.5 * rho * (ux * ux + uy * uy + uz * uz) + 0.5 * b2
        return {
            "mass": float(rho.sum()),
            "momentum_x": float((rho * ux).sum()),
            "momentum_y": float((rho * uy).sum()),
            "momentum_z": float((rho * uz).sum()),
            "energy": float(total.sum()),
            "magnetic_energy": float((0.5 * b2).sum()),
            "cross_helicity": float((ux * Bx + uy * By + uz * Bz).sum()),
        }


def project_solenoidal(grid: Grid3DOracle, Bx, By, Bz):
    """phrike/grid.py:1648-1684."""
    F = [grid.fftn(Bx), grid.fftn(By), grid.fftn(Bz)]
    for G in F:
        if grid.Nx % 2 == 0:
            G[:, :, grid.Nx // 2] = 0
        if grid.Ny % 2 == 0:
            G[:, grid.Ny // 2, :] = 0
        if grid.Nz % 2 == 0:
            G[grid.Nz // 2, :, :] = 0
    kx = grid.kx[None, None, :]
    ky = grid.ky[None, :, None]
    kz = grid.kz[:, None, None]
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k2_safe = np.where(k2 == 0, 1.0, k2)
    coef = (kx * F[0] + ky * F[1] + kz * F[2]) / k2_safe
    return grid.ifftn(F[0] - kx * coef), grid.ifftn(F[1] - ky * coef), grid.ifftn(F[2] - kz * coef)


def divergence(grid: Grid3DOracle, Vx, Vy, Vz):
    """phrike/grid.py:1635-1646 (masked derivative operator)."""
    Fx = grid._apply_masks(grid.fftn(Vx))
    Fy = grid._apply_masks(grid.fftn(Vy))
    Fz = grid._apply_masks(grid.fftn(Vz))
    return grid.ifftn(grid.ikx * Fx + grid.iky * Fy + grid.ikz * Fz)


def rhs_mhd3d(grid: Grid3DOracle, eqs: MHD3DOracle, U):
    """phrike/solver.py:932-954 (resistivity and viscosity 0)."""
    Fx, Fy, Fz = eqs.flux(U)
    return -(grid.dx1(Fx) + grid.dy1(Fy) + grid.dz1(Fz))


class SolverMHD3DOracle:
    """SpectralSolverMHD, 3D, reference defaults (phrike/solver.py:840-1050, 1173)."""

    def __init__(self, grid: Grid3DOracle, equations: MHD3DOracle, cfl: float = 0.3):
        self.grid = grid
        self.equations = equations
        self.cfl = cfl

    def clamp(self, U):
        """phrike/solver.py:956-964."""
        rho, ux, uy, uz, p, Bx, By, Bz = self.equations.primitive(U)
        return self.equations.conservative(np.maximum(rho, RHO_MIN), ux, uy, uz,
                                           np.maximum(p, P_MIN), Bx, By, Bz)

    def project_B(self, U):
        """phrike/solver.py:966-986: project B, E from the projected B at the
        thermal pressure of the unprojected state."""
        rho, ux, uy, uz, p, Bx, By, Bz = self.equations.primitive(U)
        Bx, By, Bz = project_solenoidal(self.grid, U[5], U[6], U[7])
        U = U.copy()
        U[5], U[6], U[7] = Bx, By, Bz
        U[4] = (p / (self.equations.gamma - 1.0) + 0.5 * rho * (ux * ux + uy * uy + uz * uz)
                + 0.5 * (Bx * Bx + By * By + Bz * Bz))
        return U

    def compute_dt(self, U) -> float:
        s = self.equations.max_wave_speed(U)
        if s <= 0.0:
            return 1e-6
        return self.cfl * min(self.grid.dx, self.grid.dy, self.grid.dz) / s

    def step(self, U, dt: float):
        """phrike/solver.py:988-1006, 1041-1050."""
        g, e = self.grid, self.equations
        proj, clamp = self.project_B, self.clamp
        k1 = rhs_mhd3d(g, e, U)
        U2 = proj(clamp(U + 0.5 * dt * k1))
        k2 = rhs_mhd3d(g, e, U2)
        U3 = proj(clamp(U + 0.5 * dt * k2))
        k3 = rhs_mhd3d(g, e, U3)
        U4 = proj(clamp(U + dt * k3))
        k4 = rhs_mhd3d(g, e, U4)
        Un = proj(clamp(U + (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)))
        Un = g.apply_spectral_filter(Un)
        return proj(Un)


def alfven_cp_3d(X, t=0.0, rho0=1.0, p0=0.1, B0=1.0, b1=0.1, m=1, Lx=1.0, direction=1,
                 gamma=5.0 / 3.0):
    """The exact circularly polarised Alfven wave along x (the reference's
    alfven3d initial condition, phrike/problems/problem_list/alfven3d.py:24-47):
    B = (B0, b1 cos phi, b1 sin phi), u = -direction B_perp / sqrt(rho0),
    phi = k x - k direction cA t, cA = B0 / sqrt(rho0)."""
    cA = B0 / np.sqrt(rho0)
    k = 2.0 * np.pi * m / Lx
    phi = k * X - k * direction * cA * t
    rho = np.full_like(X, float(rho0))
    p = np.full_like(X, float(p0))
    Bx = np.full_like(X, float(B0))
    By, Bz = b1 * np.cos(phi), b1 * np.sin(phi)
    inv = 1.0 / np.sqrt(rho0)
    ux = np.zeros_like(X)
    uy, uz = -float(direction) * By * inv, -float(direction) * Bz * inv
    return MHD3DOracle(gamma).conservative(rho, ux, uy, uz, p, Bx, By, Bz)
