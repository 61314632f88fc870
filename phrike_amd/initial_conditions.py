"""Initial conditions for the benchmark workloads (host-side, numpy, FP64).

These are workload definitions, not part of the reference's numerics: the
reference's own problem generators use other profiles
(phrike/problems/problem_list/tgv3d.py:17-63 constant pressure,
turb3d.py:82-172 band spectrum, khi2d.py:16-116 Gaussian perturbation), so
the pressure, spectrum and layer profiles below are stated here and fed to
the unchanged step/grid/equation semantics.

Layout contract (phrike/grid.py:1728-1744): arrays are (ncomp, [Nz,] Ny, Nx)
with x fastest; conservative components are [rho, rho*u, rho*v, (rho*w,) E].
"""
from __future__ import annotations

import os

import numpy as np


def _conservative3d(rho, ux, uy, uz, p, gamma):
    E = p / (gamma - 1.0) + 0.5 * rho * (ux * ux + uy * uy + uz * uz)
    return np.array([rho, rho * ux, rho * uy, rho * uz, E])


def _conservative2d(rho, ux, uy, p, gamma):
    E = p / (gamma - 1.0) + 0.5 * rho * (ux * ux + uy * uy)
    return np.array([rho, rho * ux, rho * uy, E])


def periodic_axis(n: int, length: float) -> np.ndarray:
    return np.linspace(0.0, length, n, endpoint=False)


def taylor_green_3d(nx: int, ny: int, nz: int, lx: float = 2 * np.pi,
                    ly: float = 2 * np.pi, lz: float = 2 * np.pi,
                    gamma: float = 1.4) -> np.ndarray:
    """3D Taylor-Green vortex with the benchmark pressure:
    rho = 1, u = sin x cos y cos z, v = -cos x sin y cos z, w = 0,
    p = 100 + (1/16)(cos 2x + cos 2y)(cos 2z + 2)."""
    x = periodic_axis(nx, lx)
    y = periodic_axis(ny, ly)
    z = periodic_axis(nz, lz)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rho = np.ones_like(X)
    ux = np.sin(X) * np.cos(Y) * np.cos(Z)
    uy = -np.cos(X) * np.sin(Y) * np.cos(Z)
    uz = np.zeros_like(X)
    p = 100.0 + (1.0 / 16.0) * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    return _conservative3d(rho, ux, uy, uz, p, gamma)


def decaying_turbulence_3d(n: int, seed: int = 1234, density_seed: int = 4321,
                           mach_rms: float = 0.3, gamma: float = 1.4,
                           length: float = 2 * np.pi, k_peak: float = 4.0) -> np.ndarray:
    """Seeded random-phase solenoidal velocity with shell spectrum
    E(k) ~ k^4 exp(-2 (k/k_peak)^2), scaled to rms Mach ``mach_rms`` against
    c0 = sqrt(gamma p0 / rho0) = 1 (rho0 = 1, p0 = 1/gamma), where rms Mach is
    sqrt(<|u|^2>)/c0; rho = 1 + 0.01 N(0,1) from a separate seed; uniform p0.

    Velocity: white Gaussian noise per component (``default_rng(seed)``),
    shaped in Fourier space by sqrt(E(k) / (4 pi k^2)), Leray-projected
    (u_hat -= k (k.u_hat)/k^2, as phrike/grid.py:1648-1684), mean removed."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    try:  # multi-threaded transforms for large grids (host-side set-up only)
        import scipy.fft as _fft
        kw = {"workers": os.cpu_count() or 1}
    except ImportError:  # pragma: no cover
        _fft, kw = np.fft, {}
    k1 = np.fft.fftfreq(n, d=length / n) * 2 * np.pi
    kh = np.fft.rfftfreq(n, d=length / n) * 2 * np.pi  # last axis: half spectrum
    KZ, KY, KX = k1[:, None, None], k1[None, :, None], kh[None, None, :]
    k2 = (KX ** 2 + KY ** 2) + KZ ** 2
    kmag = np.sqrt(k2)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(kmag > 0, np.sqrt(kmag ** 4 * np.exp(-2.0 * (kmag / k_peak) ** 2)
                                         / (4 * np.pi * np.maximum(k2, 1e-300))), 0.0)
    del kmag
    uh = [_fft.rfftn(rng.standard_normal(shape), **kw) * amp for _ in range(3)]
    del amp
    k2s = np.where(k2 == 0, 1.0, k2)
    div = (KX * uh[0] + KY * uh[1] + KZ * uh[2]) / k2s
    del k2, k2s
    uh = [uh[0] - KX * div, uh[1] - KY * div, uh[2] - KZ * div]
    del div
    u = [_fft.irfftn(c, s=shape, **kw) for c in uh]
    del uh
    u = [c - c.mean() for c in u]
    rms = np.sqrt(np.mean(u[0] ** 2 + u[1] ** 2 + u[2] ** 2))
    p0 = 1.0 / gamma
    c0 = np.sqrt(gamma * p0 / 1.0)
    scale = mach_rms * c0 / rms if rms > 0 else 0.0
    u = [c * scale for c in u]
    drng = np.random.default_rng(density_seed)
    rho = 1.0 + 0.01 * drng.standard_normal(shape)
    p = np.full(shape, p0)
    return _conservative3d(rho, u[0], u[1], u[2], p, gamma)


def _shear_bracket(yc: np.ndarray, half_width: float = 0.25, delta: float = 0.05) -> np.ndarray:
    """tanh((y+h)/d) - tanh((y-h)/d): 2 inside |y|<h, 0 outside."""
    return np.tanh((yc + half_width) / delta) - np.tanh((yc - half_width) / delta)


def khi_2d_periodic(nx: int, ny: int, seed: int = 0, gamma: float = 1.4,
                    lx: float = 1.0, ly: float = 1.0) -> np.ndarray:
    """2D Kelvin-Helmholtz on [0,1)^2 (Fourier x Fourier): shear layers at
    y = 0.25 and 0.75 (the centred profile shifted by 1/2), u = +-0.5 with
    tanh width 0.05, rho 1 outside / 2 inside, p = 2.5,
    v = 0.01 sin(4 pi x) + 1e-4 N(0,1) from ``default_rng(seed)``."""
    x = periodic_axis(nx, lx)
    y = periodic_axis(ny, ly)
    X, Y = np.meshgrid(x, y, indexing="xy")
    br = _shear_bracket(Y - 0.5 * ly)
    ux = -0.5 + 0.5 * br
    rho = 1.0 + 0.5 * br
    rng = np.random.default_rng(seed)
    uy = 0.01 * np.sin(4 * np.pi * X / lx) + 1e-4 * rng.standard_normal(X.shape)
    p = np.full_like(X, 2.5)
    return _conservative2d(rho, ux, uy, p, gamma)


def khi_2d_gll(nx: int, y_nodes: np.ndarray, seed: int = 0, gamma: float = 1.4,
               lx: float = 1.0) -> np.ndarray:
    """2D Kelvin-Helmholtz with x in [0,1) Fourier and y on given nodes in
    [-1/2, 1/2] (LGL): u = -0.5 + 0.5[tanh((y+.25)/.05) - tanh((y-.25)/.05)],
    rho = 1 + 0.5[...], p = 2.5, v = 0.01 sin(4 pi x) + 1e-4 N(0,1)."""
    x = periodic_axis(nx, lx)
    X, Y = np.meshgrid(x, np.asarray(y_nodes, dtype=float), indexing="xy")
    br = _shear_bracket(Y)
    ux = -0.5 + 0.5 * br
    rho = 1.0 + 0.5 * br
    rng = np.random.default_rng(seed)
    uy = 0.01 * np.sin(4 * np.pi * X / lx) + 1e-4 * rng.standard_normal(X.shape)
    p = np.full_like(X, 2.5)
    return _conservative2d(rho, ux, uy, p, gamma)


def smooth_random_state_3d(nx: int, ny: int, nz: int, seed: int = 7,
                           kmax: int = 3, gamma: float = 1.4) -> np.ndarray:
    """Band-limited random state (|m| <= kmax per axis) with positive rho, p:
    used for operator-level parity checks on [0, 2pi)^3."""
    rng = np.random.default_rng(seed)
    x = periodic_axis(nx, 2 * np.pi)
    y = periodic_axis(ny, 2 * np.pi)
    z = periodic_axis(nz, 2 * np.pi)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")

    def field():
        f = np.zeros_like(X)
        for _ in range(6):
            a, b, c = rng.integers(-kmax, kmax + 1, size=3)
            ph = rng.uniform(0, 2 * np.pi)
            f += rng.uniform(-1, 1) * np.cos(a * X + b * Y + c * Z + ph)
        return f / 6.0

    rho = 1.0 + 0.2 * field()
    ux, uy, uz = 0.5 * field(), 0.5 * field(), 0.5 * field()
    p = 1.0 + 0.2 * field()
    return _conservative3d(rho, ux, uy, uz, p, gamma)


def smooth_random_state_2d(nx: int, ny: int, seed: int = 7, kmax: int = 3,
                           gamma: float = 1.4) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = periodic_axis(nx, 1.0)
    y = periodic_axis(ny, 1.0)
    X, Y = np.meshgrid(x, y, indexing="xy")

    def field():
        f = np.zeros_like(X)
        for _ in range(6):
            a, b = rng.integers(-kmax, kmax + 1, size=2)
            ph = rng.uniform(0, 2 * np.pi)
            f += rng.uniform(-1, 1) * np.cos(2 * np.pi * (a * X + b * Y) + ph)
        return f / 6.0

    rho = 1.0 + 0.2 * field()
    ux, uy = 0.5 * field(), 0.5 * field()
    p = 1.0 + 0.2 * field()
    return _conservative2d(rho, ux, uy, p, gamma)


def taylor_green_3d_torch(n: int, device, dtype=None, gamma: float = 1.4):
    """The benchmark TGV of ``taylor_green_3d`` evaluated with torch on the
    device (same formula; avoids a host meshgrid of n^3 at large n)."""
    import torch

    dtype = dtype or torch.float64
    x = torch.arange(n, dtype=torch.float64, device=device) * (2 * np.pi / n)
    Z = x.view(n, 1, 1)
    Y = x.view(1, n, 1)
    X = x.view(1, 1, n)
    ux = torch.sin(X) * torch.cos(Y) * torch.cos(Z)
    uy = -torch.cos(X) * torch.sin(Y) * torch.cos(Z)
    p = 100.0 + (1.0 / 16.0) * (torch.cos(2 * X) + torch.cos(2 * Y)) * (torch.cos(2 * Z) + 2.0)
    U = torch.empty((5, n, n, n), dtype=torch.float64, device=device)
    U[0] = 1.0
    U[1] = ux
    U[2] = uy
    U[3] = 0.0
    U[4] = p / (gamma - 1.0) + 0.5 * (ux * ux + uy * uy)
    return U.to(dtype)
