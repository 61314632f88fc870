"""phrike_amd: MI355X (gfx950) ``hip`` backend for PHRIKE's pseudo-spectral
Euler right-hand side and RK4 step.

Public names follow the reference package (phrike/grid.py, phrike/equations.py,
phrike/solver.py, phrike/basis/legendre.py).  Every object requires
``backend="hip"`` and a ROCm GPU; the compute runs in libphrike_hip.so
(hand-written HIP kernels, C ABI in include/phrike_hip.h).
"""
from __future__ import annotations

__version__ = "0.1.0"
BACKENDS = ("hip",)


def hip_available() -> bool:
    """True when the library loads and PyTorch sees a GPU."""
    try:
        import torch

        from . import _lib

        _lib.load()
        return bool(torch.cuda.is_available())
    except Exception:
        return False


def __getattr__(name):  # lazy imports keep `import phrike_amd` torch-free
    if name in ("Grid2D", "Grid3D"):
        from . import grid

        return getattr(grid, name)
    if name in ("EulerEquations2D", "EulerEquations3D"):
        from . import equations

        return getattr(equations, name)
    if name in ("SpectralSolver2D", "SpectralSolver3D"):
        from . import solver

        return getattr(solver, name)
    raise AttributeError(name)
