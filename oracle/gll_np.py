"""Legendre-Gauss-Lobatto (LGL) tables: nodes, weights, differentiation matrix,
nodal filter and the dense apply out[c,i,x] = sum_j D[i,j] F[c,j,x].

TEST INFRASTRUCTURE ONLY (the oracle).  The shipped package builds its own
tables (``phrike_amd.basis``); this module is what they are checked against.

Follows phrike/basis/legendre.py:
* nodes: endpoints +-1 and the roots of P'_{N-1}, i.e. the Gauss-Jacobi(1,1)
  nodes of order N-2 (scipy ``roots_jacobi``), weights 2/(N(N-1) P_{N-1}^2)
  evaluated in 32-digit mpmath from double nodes (legendre.py:523-602,
  628-671);
* orientation: nodes sorted by descending y, x = (1 - y) L/2 so x ascends
  (legendre.py:706-717);
* D_y from the Legendre Vandermonde solve D_y V = dV (legendre.py:605-624),
  D_x = (-2/L) D_y;
* nodal filter Fn = (V diag(sigma*(2k+1)/2) V^T) diag(w_ref)
  (legendre.py:817-829);
* apply along the second-to-last axis (legendre.py:1020-1028, torch path; the
  numpy fallback at 1036-1038 returns a transposed array and is not followed).

Parity pinning: D x^m = m x^(m-1) for m <= N-1, zero row sums, corner
entries -+N(N-1)/4 * (2/L), agreement with the closed-form
D_ij = P(y_i)/(P(y_j)(y_i-y_j)) (tests/test_oracle_gll.py) and the reference's
own tables written to tests/golden/ by tests/golden/make_golden.py.
"""
from __future__ import annotations

import numpy as np
import numpy.polynomial.legendre as npleg
from scipy.special import eval_legendre, roots_jacobi

try:
    import mpmath

    _HAVE_MPMATH = True
except Exception:  # pragma: no cover
    _HAVE_MPMATH = False


def lgl_nodes_weights_ref(n: int):
    """Nodes ascending in [-1,1] and weights (legendre.py:523-602, 628-671)."""
    if n <= 0:
        raise ValueError("N must be positive")
    if n == 1:
        return np.array([0.0]), np.array([2.0])
    if n == 2:
        return np.array([-1.0, 1.0]), np.array([1.0, 1.0])
    if n == 3:
        return np.array([-1.0, 0.0, 1.0]), np.array([1.0 / 3.0, 4.0 / 3.0, 1.0 / 3.0])
    xi, _ = roots_jacobi(n - 2, 1.0, 1.0)
    pv = eval_legendre(n - 1, xi)
    if _HAVE_MPMATH:
        old = mpmath.mp.dps
        mpmath.mp.dps = 32
        try:
            wi = [float(mpmath.mpf(2) / (n * (n - 1) * mpmath.mpf(float(p)) ** 2)) for p in pv]
            we = float(mpmath.mpf(2) / (n * (n - 1)))
        finally:
            mpmath.mp.dps = old
    else:  # pragma: no cover
        wi = list(2.0 / (n * (n - 1) * pv ** 2))
        we = 2.0 / (n * (n - 1))
    y = np.concatenate([[-1.0], xi, [1.0]])
    w = np.concatenate([[we], wi, [we]])
    return y, w


def stable_diff_matrix(y: np.ndarray) -> np.ndarray:
    """D_y with D_y V = dV on the nodes y (legendre.py:605-624)."""
    n = y.shape[0]
    V = npleg.legvander(y, n - 1)
    dV = np.empty_like(V)
    eye = np.eye(n)
    for k in range(n):
        dV[:, k] = npleg.legval(y, npleg.legder(eye[k]))
    return np.linalg.solve(V.T, dV.T).T


class LGLBasisOracle:
    """LegendreLobattoBasis1D tables on [0, L] (legendre.py:681-729)."""

    def __init__(self, n: int, length: float):
        self.N = n
        self.L = float(length)
        y, w = lgl_nodes_weights_ref(n)
        order = np.argsort(-y)
        self.y = y[order]
        self.w_ref = w[order]
        self.x = (1.0 - self.y) * (self.L / 2.0)
        self.w_x = self.w_ref * (self.L / 2.0)
        self.D_y = stable_diff_matrix(self.y)
        self.D_x = (-2.0 / self.L) * self.D_y
        self.V = npleg.legvander(self.y, n - 1)
        self.proj_scale = (2.0 * np.arange(n) + 1.0) / 2.0

    def sigma(self, p: int, alpha: float) -> np.ndarray:
        k = np.arange(self.N, dtype=float)
        eta = k / max(float(self.N - 1), 1.0)
        return np.exp(-float(alpha) * eta ** int(p))

    def nodal_filter(self, p: int, alpha: float) -> np.ndarray:
        sp = self.sigma(p, alpha) * self.proj_scale
        return ((self.V * sp[None, :]) @ self.V.T) * self.w_ref[None, :]


def apply_along_y(D: np.ndarray, F: np.ndarray) -> np.ndarray:
    """out[..., i, x] = sum_j D[i, j] F[..., j, x]."""
    return np.einsum("ij,...jx->...ix", D, F)


def apply_brute(D: np.ndarray, F: np.ndarray) -> np.ndarray:
    """Triple loop, for tiny shapes only: the definition itself."""
    out = np.zeros(F.shape, dtype=np.float64)
    nc, ny, nx = F.shape
    for c in range(nc):
        for i in range(ny):
            for x in range(nx):
                s = 0.0
                for j in range(ny):
                    s += D[i, j] * F[c, j, x]
                out[c, i, x] = s
    return out
