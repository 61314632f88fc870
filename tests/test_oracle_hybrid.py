"""Pins the hybrid Fourier-x / LGL-y oracle (oracle/hybrid_np.py).

There is no reference solver for this grid, so each piece is checked against
what fixes it: closed-form derivatives, the reference-pinned Fourier 2D oracle
on states that do not vary along y, exact polynomial derivatives and the
modal definition of the nodal filter."""
import numpy as np
import numpy.polynomial.legendre as npleg
import pytest

from oracle.hybrid_np import HybridGridOracle, HybridSolverOracle, rhs_hybrid
from oracle.spectral_np import Euler2DOracle, Grid2DOracle, Solver2DOracle

FP = {"enabled": True, "p": 16, "alpha": 36.0}


def test_dx_closed_form():
    g = HybridGridOracle(48, 16, lx=2.0)
    X = np.broadcast_to(g.x, (16, 48))
    for m in (1, 5, 16):
        f = np.sin(2 * np.pi * m * X / 2.0)
        exact = (2 * np.pi * m / 2.0) * np.cos(2 * np.pi * m * X / 2.0)
        np.testing.assert_allclose(g.dx1(f), exact, atol=1e-12 * m)
    # m = 17 > 48 // 3 is removed by the dealias mask
    assert np.abs(g.dx1(np.sin(2 * np.pi * 17 * X / 2.0))).max() < 1e-12
    # filter: sigma(m) on each mode
    gf = HybridGridOracle(48, 16, lx=2.0, filter_params=FP)
    s = np.exp(-36.0 * (5 / 24) ** 16)
    f = np.sin(2 * np.pi * 5 * X / 2.0)
    np.testing.assert_allclose(gf.dx1(f), s * (2 * np.pi * 5 / 2.0) * np.cos(2 * np.pi * 5 * X / 2.0),
                               atol=1e-12)


def test_dx_matches_fourier_oracle_rows():
    g = HybridGridOracle(32, 16, filter_params=FP)
    g2 = Grid2DOracle(32, 8, 1.0, 1.0, dealias=True, filter_params=FP)
    rng = np.random.default_rng(0)
    row = rng.standard_normal(32)
    f = np.broadcast_to(row, (8, 32)).copy()
    np.testing.assert_allclose(g.dx1(f[:1])[0], g2.dx1(f)[0], rtol=0, atol=1e-12)


def test_dy_exact_on_polynomials():
    g = HybridGridOracle(16, 24, ly=1.0)
    y = g.y
    for deg in (0, 1, 7, 23):
        c = np.zeros(deg + 1)
        c[-1] = 1.0
        # polynomial in t = 2*(y - y0)/ly - 1 on [-1, 1]
        t = 2.0 * (y + 0.5) - 1.0
        f = npleg.legval(t, c)
        df = 2.0 * npleg.legval(t, npleg.legder(c))
        F = np.repeat(f[:, None], 16, axis=1)
        scale = max(1.0, np.abs(df).max())
        np.testing.assert_allclose(g.dy1(F[None])[0][:, 0], df, atol=1e-9 * scale)


def test_nodal_filter_modal_definition():
    g = HybridGridOracle(16, 20, filter_params=FP)
    t = 2.0 * (g.y + 0.5) - 1.0
    sig = g.basis.sigma(16, 36.0)
    for k in (0, 3, 12, 19):
        c = np.zeros(k + 1)
        c[-1] = 1.0
        f = npleg.legval(t, c)
        out = g.Fn @ f
        np.testing.assert_allclose(out, sig[k] * f, atol=1e-12)


@pytest.mark.parametrize("y_bc,uy", [("none", -0.1), ("wall", 0.0)])
def test_rhs_zero_on_uniform_state(y_bc, uy):
    g = HybridGridOracle(32, 16, filter_params=FP, y_bc=y_bc)
    eqs = Euler2DOracle(1.4)
    U = eqs.conservative(np.full((16, 32), 1.3), np.full((16, 32), 0.2), np.full((16, 32), uy),
                         np.full((16, 32), 2.5))
    r = rhs_hybrid(g, eqs, U)
    assert np.abs(r).max() < 1e-12


@pytest.mark.parametrize("y_bc,uy0", [("none", 0.2), ("wall", 0.0)])
def test_step_matches_fourier_solver_on_x_only_state(y_bc, uy0):
    """A state that is uniform along y has D.Fy = 0 and is left alone by the
    nodal filter (and, with uy = 0, by the wall rule), so a hybrid step equals
    the Fourier 2D solver's step (pinned by the reference's goldens): same
    RK4/clamp/filter composition."""
    nx = 32
    g = HybridGridOracle(nx, 16, filter_params=FP, y_bc=y_bc)
    g2 = Grid2DOracle(nx, 8, 1.0, 1.0, dealias=True, filter_params=FP)
    eqs = Euler2DOracle(1.4)
    x = g.x
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x) + 0.1 * np.cos(6 * np.pi * x)
    ux = 0.4 * np.cos(2 * np.pi * x)
    uy = uy0 * (1.0 + 0.25 * np.sin(4 * np.pi * x))
    p = 2.5 + 0.2 * np.cos(2 * np.pi * x)
    row = eqs.conservative(rho, ux, uy, p)
    Uh = np.repeat(row[:, None, :], 16, axis=1)
    Uf = np.repeat(row[:, None, :], 8, axis=1)
    dt = 1e-3
    a = HybridSolverOracle(g, eqs).step(Uh, dt)
    b = Solver2DOracle(g2, eqs).step(Uf, dt)
    for c in range(4):
        sc = max(np.abs(b[c]).max(), 1e-3 * np.abs(b).max())  # my is ~0 when uy = 0
        assert np.abs(a[c] - b[c][:1]).max() <= 1e-12 * sc


def test_cfl_uses_min_spacing():
    g = HybridGridOracle(32, 64)
    eqs = Euler2DOracle(1.4)
    U = eqs.conservative(np.ones((64, 32)), np.zeros((64, 32)), np.zeros((64, 32)), np.full((64, 32), 1.0))
    dt = HybridSolverOracle(g, eqs, cfl=0.5).compute_dt(U)
    a = np.sqrt(1.4)
    assert g.dy_min < g.dx
    np.testing.assert_allclose(dt, 0.5 * g.dy_min / a, rtol=1e-15)


def test_wall_rule_entries():
    """phrike/grid.py:621-738 "wall" along y: normal momentum zero on both
    boundary rows, density and energy copied from the adjacent row, the
    tangential momentum and all interior rows untouched."""
    g = HybridGridOracle(16, 16)
    U = np.random.default_rng(1).standard_normal((4, 16, 16)) + 3.0
    W = g.walls(U)
    np.testing.assert_array_equal(W[:, 1:-1], U[:, 1:-1])
    np.testing.assert_array_equal(W[1], U[1])
    for row, nb in ((0, 1), (-1, -2)):
        np.testing.assert_array_equal(W[0, row], U[0, nb])
        np.testing.assert_array_equal(W[3, row], U[3, nb])
        assert not W[2, row].any()
    assert g.walls(W) is not W and np.array_equal(g.walls(W), W)  # idempotent
    assert HybridGridOracle(16, 16, y_bc="none").walls(U) is U


def test_walls_keep_khi_bounded():
    """Without the walls the collocation blows up at the boundary rows of the
    KHI shear flow within ~100 steps at 256 x 128; with them a short run at a
    smaller size stays bounded (smoke check of the boundary treatment)."""
    from oracle.hybrid_np import HybridSolverOracle
    from phrike_amd.initial_conditions import khi_2d_gll
    g = HybridGridOracle(64, 48, filter_params=FP)
    eqs = Euler2DOracle(1.4)
    s = HybridSolverOracle(g, eqs, cfl=0.4)
    U = khi_2d_gll(64, g.y, seed=2)
    dt = s.compute_dt(U)
    for _ in range(150):
        U = s.step(U, dt)
    rho, ux, uy, p = eqs.primitive(U)
    assert np.isfinite(U).all() and np.abs(uy).max() < 0.05 and p.min() > 2.0
