"""Pins the hybrid Fourier-x / LGL-y oracle (oracle/hybrid_np.py).

There is no reference solver for this grid, so each piece is checked against
what fixes it: closed-form derivatives, the reference-pinned Fourier 2D oracle
on states that do not vary along y, exact polynomial derivatives and the
modal definition of the nodal filter."""
import math

import numpy as np
import numpy.polynomial.legendre as npleg
import numpy.polynomial as npp
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


# ---------------------------------------------------------------------------
# The D.Fy leg on y-varying states.  Each expectation below comes from the
# mathematics of the state (polynomial calculus in y, single Fourier modes in
# x), not from the oracle's own operators, so a sign flip, a dropped term or
# a transposed D fails.
# ---------------------------------------------------------------------------
def _legendre_poly(g, coeffs):
    """f(y) = sum_k c_k P_k(t) with t = 2 (y - y_bottom) / ly - 1 on the grid's
    nodes, and its exact derivative df/dy."""
    t = 2.0 * (g.y - g.y[0]) / g.ly - 1.0
    f = npleg.legval(t, coeffs)
    df = (2.0 / g.ly) * npleg.legval(t, npleg.legder(coeffs))
    return f, df


@pytest.mark.parametrize("filt", [None, FP])
def test_rhs_entropy_wave_closed_form(filt):
    """rho = 1 + eps P(y) + delta sin(k x), u = u0, v = v0, p = p0: every
    flux is linear in rho, so rhs = -c (u0 dx rho + v0 dy rho) with
    c = (1, u0, v0, (u0^2 + v0^2)/2), dx rho = delta k sigma(k) cos(k x)
    (mask * sigma of the x derivative), dy rho = eps P'(y) exactly (P has
    degree < Ny)."""
    nx, ny = 32, 16
    g = HybridGridOracle(nx, ny, filter_params=filt, y_bc="none")
    eqs = Euler2DOracle(1.4)
    u0, v0, p0, eps, delta, m = 0.3, -0.45, 2.5, 0.25, 0.02, 10  # m = nx // 3: kept
    coeffs = np.array([0.0, 0.7, -0.4, 0.25, 0.1, -0.3, 0.2, 0.15, -0.1, 0.05])
    P, dP = _legendre_poly(g, coeffs)
    k = 2 * np.pi * m / g.lx
    sig = np.exp(-36.0 * (m / (nx / 2)) ** 16) if filt else 1.0
    X = g.x[None, :]
    rho = 1.0 + eps * P[:, None] + delta * np.sin(k * X)
    shape = rho.shape
    U = eqs.conservative(rho, np.full(shape, u0), np.full(shape, v0), np.full(shape, p0))
    drho_x = delta * k * sig * np.cos(k * X) * np.ones((ny, 1))
    drho_y = eps * dP[:, None] * np.ones((1, nx))
    c = np.array([1.0, u0, v0, 0.5 * (u0 * u0 + v0 * v0)])
    expect = -c[:, None, None] * (u0 * drho_x + v0 * drho_y)[None]
    r = rhs_hybrid(g, eqs, U)
    scale = np.abs(expect).max()
    assert np.abs(r - expect).max() <= 1e-11 * scale
    # the y leg alone carries most of the rhs: a transposed D or a flipped
    # sign of D.Fy would be off by O(scale)
    assert np.abs(v0 * drho_y).max() > 0.5 * np.abs(u0 * drho_x).max()
    gt = HybridGridOracle(nx, ny, filter_params=filt, y_bc="none")
    gt.D = g.D.T.copy()
    assert np.abs(rhs_hybrid(gt, eqs, U) - expect).max() > 1e-2 * scale


def _wall_state(g, eqs, rho0=1.3, garbage=True):
    """x-uniform state with constant rho, polynomial mx(y), my(y), E(y) such
    that the wall rule turns it into a polynomial state: my = 0 on both walls,
    E(y0) = E(y1) and E(y_{N-2}) = E(y_{N-1}).  With `garbage`, the rows the
    rule overwrites (rho, my, E on both walls) hold other values."""
    y = g.y
    Pm = npp.Polynomial
    ya, yb = y[0], y[-1]
    mx = Pm([0.4, -0.3, 0.5])
    my = 0.6 * Pm.fromroots([ya, yb]) * Pm([1.0, 0.8])
    E = 3.0 + 5.0 * Pm.fromroots([y[0], y[1], y[-2], y[-1]])
    p = (eqs.gamma - 1.0) * (E - (mx * mx + my * my) / (2.0 * rho0))
    Fy = [my, mx * my / rho0, my * my / rho0 + p, (E + p) * my / rho0]
    nx = g.nx
    U = np.empty((4, len(y), nx))
    for f, poly in enumerate((Pm([rho0]), mx, my, E)):
        U[f] = poly(y)[:, None]
    if garbage:
        U[0, 0] += 0.21
        U[0, -1] -= 0.17
        U[2, 0] = 0.33
        U[2, -1] = -0.29
        U[3, 0] += 0.5
        U[3, -1] -= 0.4
    dFy = np.stack([q.deriv()(y) for q in Fy])
    return U, -np.repeat(dFy[:, :, None], nx, axis=2)


@pytest.mark.parametrize("y_bc,garbage", [("wall", True), ("wall", False), ("none", False)])
def test_rhs_polynomial_state_with_walls(y_bc, garbage):
    """rhs = -d/dy Fy(y) of an x-uniform state whose fluxes are polynomials of
    degree < Ny (rho constant), every component non-zero (rho u v included);
    with the wall rule, the input's boundary rows hold values the rule must
    overwrite first (phrike/grid.py:621-738)."""
    g = HybridGridOracle(16, 16, ly=1.0, y0=-0.5, y_bc=y_bc)
    eqs = Euler2DOracle(1.4)
    U, expect = _wall_state(g, eqs, garbage=garbage)
    r = rhs_hybrid(g, eqs, U)
    scale = np.abs(expect).max()
    for f in range(4):
        assert np.abs(expect[f]).max() > 1e-3 * scale
    assert np.abs(r - expect).max() <= 1e-10 * scale
    if garbage:  # without the rule the same input gives another rhs
        gn = HybridGridOracle(16, 16, ly=1.0, y0=-0.5, y_bc="none")
        assert np.abs(rhs_hybrid(gn, eqs, U) - expect).max() > 1e-2 * scale


def test_rk4_step_entropy_wave_exact():
    """One RK4 step of the entropy wave (no filter, no walls): the state
    stays (rho, u0, v0, p0) and rho advances by the RK4 polynomial of the
    operator L = -(u0 dx + v0 D): sum_{j<=4} (-v0 dt)^j / j! P^(j)(y) for the
    y part (D exact on polynomials), Im(R(z) e^{ikx}) with
    z = -i u0 k dt, R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24 for the x part."""
    nx, ny = 32, 16
    g = HybridGridOracle(nx, ny, y_bc="none")
    eqs = Euler2DOracle(1.4)
    u0, v0, p0, eps, delta, m = 0.3, -0.45, 2.5, 0.05, 0.07, 3
    coeffs = np.array([0.0, 0.7, -0.4, 0.25, 0.1, -0.3, 0.2])
    t = 2.0 * (g.y - g.y[0]) / g.ly - 1.0
    P = npleg.legval(t, coeffs)
    k = 2 * np.pi * m / g.lx
    X = g.x[None, :]
    rho = 1.0 + eps * P[:, None] + delta * np.sin(k * X)
    shape = rho.shape
    U = eqs.conservative(rho, np.full(shape, u0), np.full(shape, v0), np.full(shape, p0))
    dt = 0.02
    ypart = np.zeros(ny)
    c = coeffs.copy()
    for j in range(5):
        ypart += (-v0 * dt) ** j / math.factorial(j) * npleg.legval(t, c)
        c = (2.0 / g.ly) * npleg.legder(c) if len(c) > 1 else np.zeros(1)
    z = -1j * u0 * k * dt
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    rho_new = 1.0 + eps * ypart[:, None] + delta * np.imag(R * np.exp(1j * k * X))
    expect = eqs.conservative(rho_new, np.full(shape, u0), np.full(shape, v0), np.full(shape, p0))
    got = HybridSolverOracle(g, eqs).step(U, dt, apply_filter=False)
    for f in range(4):
        assert np.abs(got[f] - expect[f]).max() <= 1e-12 * np.abs(expect[f]).max()
    # the step moved rho by much more than the tolerance along y
    assert np.abs(ypart - P).max() > 1e-3


# ---------------------------------------------------------------------------
# Against the reference's own 1D Legendre solver with the wall rule
# (tests/golden/wall1d.npz, written by tests/golden/make_golden.py from
# Grid1D(basis="legendre", bc="wall"), _compute_rhs and _rk4_step;
# phrike/grid.py:144-738, phrike/solver.py:33-131): an x-uniform hybrid state
# with zero x momentum is that 1D problem along y.
# ---------------------------------------------------------------------------
def _as_hybrid(U1, nx):
    rho, m, E = U1
    return np.stack([np.repeat(v[:, None], nx, axis=1) for v in (rho, np.zeros_like(rho), m, E)])


def _as_1d(Uh):
    return np.stack([Uh[0, :, 0], Uh[2, :, 0], Uh[3, :, 0]])


@pytest.mark.parametrize("n", [16, 33, 64])
@pytest.mark.parametrize("ftag", ["f", "nof"])
def test_wall_composition_vs_reference_1d(n, ftag):
    from conftest import golden
    d = golden("wall1d.npz")
    key = f"N{n}_{ftag}"
    nx = 16
    g = HybridGridOracle(nx, n, lx=1.0, ly=1.0, y0=0.0, filter_params=FP if ftag == "f" else None,
                         y_bc="wall")
    np.testing.assert_allclose(g.y, d[f"{key}_x"], rtol=0, atol=1e-15)
    eqs = Euler2DOracle(1.4)
    U0 = _as_hybrid(d[f"{key}_U0"], nx)
    assert np.abs(U0[2][[0, -1]]).min() > 1e-3  # the rule has work to do
    r = rhs_hybrid(g, eqs, U0)
    ref = d[f"{key}_rhs0"]
    assert np.abs(r[1]).max() == 0.0
    np.testing.assert_allclose(_as_1d(r), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    s = HybridSolverOracle(g, eqs)
    dt = float(d[f"{key}_dt"])
    U = U0
    for kstep in range(1, 6):
        U = s.step(U, dt, apply_filter=(ftag == "f"))
        if kstep in (1, 5):
            refU = d[f"{key}_U{kstep}"]
            for f in range(3):
                err = np.abs(_as_1d(U)[f] - refU[f]).max()
                assert err <= 1e-12 * np.abs(refU[f]).max(), (kstep, f, err)
