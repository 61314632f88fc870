"""Euler equations objects of the ``hip`` backend.

Same names and methods as the reference's EulerEquations2D/3D
(phrike/equations.py:322-381, 552-637).  These pointwise helpers act on
device tensors with elementwise torch operations and are API conveniences
(initial conditions, diagnostics); the solver's RHS, step, CFL reduction and
conservation sums run in the HIP kernels (flux and primitive recovery are
fused into the x-forward pencil kernel).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict

import torch


@dataclass
class EulerEquations3D:
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
        return torch.stack([rho, rho * ux, rho * uy, rho * uz, E], dim=0)

    def flux(self, U):
        rho, ux, uy, uz, p = self.primitive(U)
        mx, my, mz = rho * ux, rho * uy, rho * uz
        Ep = U[4] + p
        Fx = torch.stack([mx, mx * ux + p, mx * uy, mx * uz, Ep * ux], dim=0)
        Fy = torch.stack([my, my * ux, my * uy + p, my * uz, Ep * uy], dim=0)
        Fz = torch.stack([mz, mz * ux, mz * uy, mz * uz + p, Ep * uz], dim=0)
        return Fx, Fy, Fz

    def max_wave_speed(self, U) -> float:
        rho, ux, uy, uz, p = self.primitive(U)
        a = torch.sqrt(self.gamma * p / rho)
        c = torch.maximum(torch.maximum(ux.abs(), uy.abs()), uz.abs()) + a
        return float(c.max().item())

    def conserved_quantities(self, U) -> Dict[str, float]:
        rho, ux, uy, uz, p = self.primitive(U)
        return {
            "mass": float(rho.sum().item()),
            "momentum_x": float((rho * ux).sum().item()),
            "momentum_y": float((rho * uy).sum().item()),
            "momentum_z": float((rho * uz).sum().item()),
            "energy": float((p / (self.gamma - 1.0)
                             + 0.5 * rho * (ux * ux + uy * uy + uz * uz)).sum().item()),
        }


@dataclass
class EulerEquations2D:
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
        return torch.stack([rho, rho * ux, rho * uy, E], dim=0)

    def flux(self, U):
        rho, ux, uy, p = self.primitive(U)
        mx, my = rho * ux, rho * uy
        Ep = U[3] + p
        Fx = torch.stack([mx, mx * ux + p, mx * uy, Ep * ux], dim=0)
        Fy = torch.stack([my, my * ux, my * uy + p, Ep * uy], dim=0)
        return Fx, Fy

    def max_wave_speed(self, U) -> float:
        # |ux| + a only, as phrike/equations.py:365-370
        rho, ux, uy, p = self.primitive(U)
        return float((ux.abs() + torch.sqrt(self.gamma * p / rho)).max().item())

    def conserved_quantities(self, U) -> Dict[str, float]:
        # phrike/equations.py:372-381 unpacks (rho, u, p, _) from (rho, ux, uy, p),
        # so its "p" is uy; kept as written there
        rho, u, p, _ = self.primitive(U)
        return {
            "mass": float(rho.sum().item()),
            "momentum": float((rho * u).sum().item()),
            "energy": float((p / (self.gamma - 1.0) + 0.5 * rho * u * u).sum().item()),
        }
