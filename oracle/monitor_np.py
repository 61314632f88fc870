"""Oracle (test infrastructure, never the product): NumPy restatement of the
run monitoring statistics of BaseProblem (phrike/problems/base.py:490-822),
in the torch branch a tensor state takes there: unbiased density standard
deviation, no pressure entries in the density statistics.

Pinned by tests/golden/monitor.npz (written from the reference's own methods
on torch tensors; tests/test_oracle_monitor.py).
"""
from __future__ import annotations

from typing import Dict

import numpy as np


def _prims(U, gamma):
    n = U.shape[0]
    rho = U[0]
    if n == 4:  # 2D Euler (phrike/equations.py:104-117)
        ux, uy = U[1] / rho, U[2] / rho
        p = (gamma - 1.0) * (U[3] - 0.5 * rho * (ux * ux + uy * uy))
        return rho, (ux, uy), p, None
    ux, uy, uz = U[1] / rho, U[2] / rho, U[3] / rho
    kin = 0.5 * rho * (ux * ux + uy * uy + uz * uz)
    if n == 5:  # 3D Euler (phrike/equations.py:150-164)
        return rho, (ux, uy, uz), (gamma - 1.0) * (U[4] - kin), None
    B = (U[5], U[6], U[7])  # 3D MHD (phrike/equations.py:922-933)
    b2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2]
    return rho, (ux, uy, uz), (gamma - 1.0) * (U[4] - kin - 0.5 * b2), B


def conserved(U, gamma) -> Dict[str, float]:
    """equations.conserved_quantities of each system (phrike/equations.py:
    372-381 for 2D, 622-637 3D, 1083-1100 MHD).  The 2D one unpacks
    (rho, u, p, _) from primitive()'s (rho, ux, uy, p), so its "p" is uy and
    its kinetic term has ux only: restated as written."""
    rho, u, p, B = _prims(U, gamma)
    if U.shape[0] == 4:
        ux, uy = u
        return {"mass": float(rho.sum()), "momentum": float((rho * ux).sum()),
                "energy": float((uy / (gamma - 1.0) + 0.5 * rho * ux * ux).sum())}
    q2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2]
    out = {"mass": float(rho.sum()), "momentum_x": float((rho * u[0]).sum()),
           "momentum_y": float((rho * u[1]).sum()), "momentum_z": float((rho * u[2]).sum())}
    if B is None:
        out["energy"] = float((p / (gamma - 1.0) + 0.5 * rho * q2).sum())
    else:
        b2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2]
        out["energy"] = float((p / (gamma - 1.0) + 0.5 * rho * q2 + 0.5 * b2).sum())
        out["magnetic_energy"] = float((0.5 * b2).sum())
        out["cross_helicity"] = float((u[0] * B[0] + u[1] * B[1] + u[2] * B[2]).sum())
    return out


def conservation_errors(U, U0, gamma) -> Dict[str, float]:
    """phrike/problems/base.py:490-664: signed relative change of mass and
    energy, absolute relative change of momenta (the difference itself when
    the initial value is 0)."""
    init, cur = conserved(U0, gamma), conserved(U, gamma)
    out = {}
    for k, v0 in init.items():
        d = (cur[k] - v0) / v0 if v0 != 0 else cur[k] - v0
        out[f"{k}_error"] = abs(d) if k.startswith("momentum") else d
    return out


def velocity_stats(U, gamma) -> Dict[str, float]:
    """phrike/problems/base.py:666-723: density-weighted rms, max, min |u|."""
    rho, u, p, _ = _prims(U, gamma)
    v = np.sqrt(sum(c * c for c in u))
    m = rho.sum()
    if m <= 0:
        return {"v_rms": 0.0, "v_max": 0.0, "v_min": 0.0}
    return {"v_rms": float(np.sqrt((rho * v * v).sum() / m)), "v_max": float(v.max()), "v_min": float(v.min())}


def density_stats(U, gamma) -> Dict[str, float]:
    """phrike/problems/base.py:725-766 (torch branch: unbiased std)."""
    rho = U[0]
    return {"rho_mean": float(rho.mean()), "rho_max": float(rho.max()), "rho_min": float(rho.min()),
            "rho_std": float(rho.std(ddof=1))}


def mhd_stats(U, gamma, max_div_b: float) -> Dict[str, float]:
    """phrike/problems/base.py:768-822."""
    if U.shape[0] != 8:
        return {}
    rho, u, p, B = _prims(U, gamma)
    b2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2]
    bmag = float(np.mean(0.5 * b2))
    return {"max|divB|": float(max_div_b), "b_rms": float(np.sqrt(np.mean(b2))),
            "magnetic_energy": float((0.5 * b2).sum()),
            "plasma_beta": float(np.mean(p)) / bmag if bmag > 0 else float("inf"),
            "cross_helicity": float((u[0] * B[0] + u[1] * B[1] + u[2] * B[2]).sum())}
