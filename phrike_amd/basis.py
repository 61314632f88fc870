"""Legendre-Gauss-Lobatto basis of the ``hip`` backend.

Tables follow LegendreLobattoBasis1D (phrike/basis/legendre.py:681-729):
nodes +-1 and the roots of P'_{N-1} (Gauss-Jacobi(1,1) nodes of order N-2 via
scipy), weights 2/(N(N-1)P_{N-1}^2) in 32-digit mpmath, orientation x = (1-y)
L/2 ascending, D_y from the Legendre-Vandermonde solve D_y V = dV and
D_x = (-2/L) D_y.  Tables are built once in FP64 on the host; the apply
out[b,i,x] = sum_j D[i,j] F[b,j,x] runs on matrix cores (ph_gll_apply):
FP64 MFMA, or FP32 MFMA with D (and FP64 data) split into hi + lo parts.
"""
from __future__ import annotations

import numpy as np
import numpy.polynomial.legendre as L
from scipy.special import eval_legendre, roots_jacobi

from . import _lib


def _lgl_reference_nodes(n: int):
    """ascending nodes on [-1, 1] and weights"""
    if n < 2:
        raise ValueError("LGL basis needs N >= 2")
    if n == 2:
        return np.array([-1.0, 1.0]), np.array([1.0, 1.0])
    if n == 3:
        return np.array([-1.0, 0.0, 1.0]), np.array([1.0 / 3.0, 4.0 / 3.0, 1.0 / 3.0])
    inner, _ = roots_jacobi(n - 2, 1.0, 1.0)
    pv = eval_legendre(n - 1, inner)
    try:
        import mpmath

        with mpmath.workdps(32):
            w_in = [float(2 / (mpmath.mpf(n * (n - 1)) * mpmath.mpf(float(v)) ** 2)) for v in pv]
    except ImportError:  # pragma: no cover
        w_in = list(2.0 / (n * (n - 1) * pv ** 2))
    w_end = 2.0 / (n * (n - 1))
    return np.r_[-1.0, inner, 1.0], np.r_[w_end, w_in, w_end]


def _vandermonde_derivative(y: np.ndarray) -> np.ndarray:
    n = y.size
    V = L.legvander(y, n - 1)
    dV = np.column_stack([L.legval(y, L.legder(np.eye(n)[k])) for k in range(n)])
    # D V = dV  <=>  V^T D^T = dV^T
    return np.linalg.solve(V.T, dV.T).T


def split_hi_lo(a: np.ndarray):
    hi = a.astype(np.float32)
    lo = (a - hi.astype(np.float64)).astype(np.float32)
    return hi, lo


class LegendreLobattoBasis1D:
    """LGL collocation on [0, Lx] with device-side D application."""

    def __init__(self, N: int, Lx: float, *, bc: str = "dirichlet", precision: str = "double",
                 use_parallel: bool = False, torch_device: str = "cuda"):
        self.N = int(N)
        self.Lx = float(Lx)
        self.bc = bc
        self.precision = precision
        y, w = _lgl_reference_nodes(self.N)
        order = np.argsort(-y)
        self._y = y[order]
        self._w_ref = w[order]
        self._x = (1.0 - self._y) * (self.Lx / 2.0)
        self._w_x = self._w_ref * (self.Lx / 2.0)
        self._D_x = np.ascontiguousarray((-2.0 / self.Lx) * _vandermonde_derivative(self._y))
        self._V = L.legvander(self._y, self.N - 1)
        self._proj_scale = (2.0 * np.arange(self.N) + 1.0) / 2.0
        self.torch_device = torch_device
        self._dev = {}

    def nodes(self):
        return self._x

    def quadrature_weights(self):
        return self._w_x.copy()

    @property
    def D_x(self):
        return self._D_x

    def nodal_filter(self, p: int = 8, alpha: float = 36.0) -> np.ndarray:
        """Fn = V diag(sigma*(2k+1)/2) V^T diag(w_ref) (phrike/basis/legendre.py:817-829)."""
        k = np.arange(self.N, dtype=float)
        sig = np.exp(-float(alpha) * (k / max(float(self.N - 1), 1.0)) ** int(p))
        return ((self._V * (sig * self._proj_scale)[None, :]) @ self._V.T) * self._w_ref[None, :]

    def device_tables(self, matrix: np.ndarray, key: str):
        import torch

        if key not in self._dev:
            dev = torch.device(self.torch_device)
            m = np.ascontiguousarray(matrix, dtype=np.float64)  # row-major for the kernel
            hi, lo = split_hi_lo(m)
            self._dev[key] = (torch.from_numpy(m).to(dev).contiguous(),
                              torch.from_numpy(np.ascontiguousarray(hi)).to(dev).contiguous(),
                              torch.from_numpy(np.ascontiguousarray(lo)).to(dev).contiguous())
        return self._dev[key]

    def apply(self, F, mode: str = "f64", matrix: str = "D", add=None, out=None):
        """out[..., i, x] = sum_j M[i, j] F[..., j, x] on matrix cores, with M the
        derivative matrix ("D") or the nodal filter ("filter16"); F is a GPU
        tensor (..., N, nx).  mode: "f64" (FP64 MFMA), "split" (FP32 MFMA, hi/lo
        split of M, and of F when F is float64).  With ``add`` the epilogue
        gives -(add + M.F)."""
        import torch

        if matrix == "D":
            M, key = self._D_x, "D"
        elif matrix.startswith("filter"):
            M, key = self.nodal_filter(int(matrix[6:] or 8), 36.0), matrix
        else:
            raise ValueError(matrix)
        Md, Mh, Ml = self.device_tables(M, key)
        if F.shape[-2] != self.N:
            raise ValueError("F must have N rows along axis -2")
        nx = F.shape[-1]
        Fc = F.contiguous()
        nb = Fc.numel() // (self.N * nx)
        if out is None:
            out = torch.empty_like(Fc)
        if mode == "f64":
            if Fc.dtype != torch.float64:
                raise TypeError("mode 'f64' needs float64 data")
            m = _lib.PH_GLL_F64
        elif mode == "split":
            m = _lib.PH_GLL_SPLIT_F64IO if Fc.dtype == torch.float64 else _lib.PH_GLL_SPLIT_F32
            if matrix == "D":  # rows sum to zero: apply to F - F[row 0]
                m |= _lib.PH_GLL_OFFSET
        else:
            raise ValueError(mode)
        addc = add.contiguous() if add is not None else None
        _lib.call("ph_gll_apply", m, self.N, nx, nb, _lib.ptr(Md), _lib.ptr(Mh), _lib.ptr(Ml),
                  _lib.ptr(Fc), _lib.ptr(addc), _lib.ptr(out), _lib.stream_handle(torch))
        return out

    def dx(self, f):
        """Derivative along the last axis (phrike/basis/legendre.py:787-802)."""
        import torch

        ft = f.transpose(-1, -2).contiguous()
        return self.apply(ft).transpose(-1, -2)
