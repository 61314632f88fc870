"""Hybrid Fourier-x / LGL-y 2D Euler on the GPU vs the oracle
(oracle/hybrid_np.py): FP64 with the FP64 matrix-core D to 1e-11 per field;
split-precision D (FP32 matrix cores on hi/lo parts) and FP32 state reported
against the FP64 result and bounded."""
import numpy as np
import pytest

from conftest import field_errors

pytestmark = pytest.mark.gpu

FP = {"enabled": True, "p": 16, "alpha": 36.0}


def _setup(nx, ny, dev, precision="double", gll="f64", batch=1, y_bc="wall"):
    from oracle.hybrid_np import HybridGridOracle, HybridSolverOracle
    from oracle.spectral_np import Euler2DOracle
    from phrike_amd.equations import EulerEquations2D
    from phrike_amd.hybrid import HybridGrid2D, HybridSolver2D
    og = HybridGridOracle(nx, ny, filter_params=FP, y_bc=y_bc)
    osol = HybridSolverOracle(og, Euler2DOracle(1.4), cfl=0.4)
    g = HybridGrid2D(nx, ny, filter_params=FP, precision=precision, gll=gll, batch=batch,
                     torch_device=str(dev), y_bc=y_bc)
    sol = HybridSolver2D(g, EulerEquations2D(1.4), cfl=0.4)
    return og, osol, g, sol


@pytest.mark.parametrize("nx,ny,nsteps,y_bc", [(64, 32, 10, "wall"), (256, 128, 3, "wall"), (32, 48, 5, "wall"),
                                             (64, 32, 10, "none")])
def test_hybrid_f64_vs_oracle(dev, nx, ny, nsteps, y_bc):
    import torch
    from oracle.hybrid_np import rhs_hybrid
    from phrike_amd.initial_conditions import khi_2d_gll
    og, osol, g, sol = _setup(nx, ny, dev, y_bc=y_bc)
    np.testing.assert_allclose(g.y_nodes, og.y, rtol=0, atol=1e-15)
    U = khi_2d_gll(nx, og.y, seed=11)
    Ut = torch.as_tensor(U, device=dev)
    r = sol.rhs(Ut).cpu().numpy()
    e_r = field_errors(r, rhs_hybrid(og, osol.eqs, U))
    print(f"hybrid {nx}x{ny} {y_bc} rhs errors {e_r}", flush=True)
    assert e_r.max() < 1e-11
    Uo = U
    for k in range(nsteps):
        dt = osol.compute_dt(Uo)
        dtd = float(sol.dt_runs(Ut)[0])
        assert abs(dtd - dt) <= 1e-13 * dt
        Uo = osol.step(Uo, dt)
        Ut = sol.step(Ut, dt)
    e = field_errors(Ut.cpu().numpy(), Uo)
    print(f"hybrid {nx}x{ny} {y_bc} {nsteps} steps errors {e}", flush=True)
    assert e.max() < 1e-11


def test_hybrid_batch_equals_alone(dev):
    import torch
    from phrike_amd.initial_conditions import khi_2d_gll
    og, osol, g, sol = _setup(64, 32, dev, batch=3)
    U = np.stack([khi_2d_gll(64, og.y, seed=s) for s in (1, 2, 3)])
    Ut = torch.as_tensor(U, device=dev)
    both = sol.advance(Ut, 4)
    for b in range(3):
        alone = sol.advance(Ut[b], 4)
        assert torch.equal(both[b], alone)


def test_hybrid_split_precision(dev):
    import torch
    from phrike_amd.initial_conditions import khi_2d_gll
    nx, ny, n = 128, 64, 50
    og, _, g64, s64 = _setup(nx, ny, dev)
    _, _, gs, ss = _setup(nx, ny, dev, gll="split")
    _, _, g32, s32 = _setup(nx, ny, dev, precision="single")
    U = torch.as_tensor(khi_2d_gll(nx, og.y, seed=5), device=dev)
    dt = s64.compute_dt(U)
    a = s64.advance(U, n, dt=dt).cpu().numpy()
    b = ss.advance(U, n, dt=dt).cpu().numpy()
    c = s32.advance(U.float(), n, dt=dt).double().cpu().numpy()
    eb, ec = field_errors(b, a), field_errors(c, a)
    # y-momentum is ~1e-2 of the others: measure it on the momentum scale too
    mom = max(np.linalg.norm(a[1]), np.linalg.norm(a[2]))
    eb2 = np.linalg.norm(b[2] - a[2]) / mom
    ec2 = np.linalg.norm(c[2] - a[2]) / mom
    print(f"hybrid {nx}x{ny} {n} steps: split(f64 state) {eb} (my/momentum {eb2:.2e}) "
          f"f32 {ec} (my/momentum {ec2:.2e})", flush=True)
    others = [0, 1, 3]
    assert eb[others].max() < 1e-5 and eb2 < 1e-5
    assert ec[others].max() < 1e-5 and ec2 < 1e-5


def test_hybrid_rejects_bad_shapes(dev):
    from phrike_amd._lib import HipBackendError
    from phrike_amd.hybrid import HybridGrid2D
    with pytest.raises(HipBackendError):
        HybridGrid2D(64, 40, torch_device=str(dev))._plan()
    with pytest.raises(HipBackendError):
        HybridGrid2D(48, 32, torch_device=str(dev))._plan()


@pytest.mark.parametrize("n", [16, 64])
@pytest.mark.parametrize("ftag", ["f", "nof"])
def test_hybrid_wall_vs_reference_1d(dev, n, ftag):
    """x-uniform state with zero x momentum = the reference's 1D Legendre
    solver with the wall rule (tests/golden/wall1d.npz; phrike/grid.py:144-738,
    phrike/solver.py:33-131): RHS and 1 / 5 steps, FP64 matrix-core D."""
    import torch
    from conftest import golden
    from phrike_amd.equations import EulerEquations2D
    from phrike_amd.hybrid import HybridGrid2D, HybridSolver2D
    d = golden("wall1d.npz")
    key = f"N{n}_{ftag}"
    nx = 16
    g = HybridGrid2D(nx, n, 1.0, 1.0, y0=0.0, filter_params=FP if ftag == "f" else None,
                     torch_device=str(dev), y_bc="wall")
    np.testing.assert_allclose(g.y_nodes, d[f"{key}_x"], rtol=0, atol=1e-15)
    s = HybridSolver2D(g, EulerEquations2D(1.4), cfl=0.4)
    rho, m, E = d[f"{key}_U0"]
    U0 = np.stack([np.repeat(v[:, None], nx, axis=1) for v in (rho, np.zeros_like(rho), m, E)])
    Ut = torch.as_tensor(U0, device=dev)
    r = s.rhs(Ut).cpu().numpy()
    ref = d[f"{key}_rhs0"]
    got = np.stack([r[0, :, 0], r[2, :, 0], r[3, :, 0]])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"wall1d {key} rhs {err:.2e}", flush=True)
    assert err < 1e-11 and np.abs(r[1]).max() < 1e-13 * np.abs(ref).max()
    dt = float(d[f"{key}_dt"])
    for k in range(1, 6):
        Ut = s.step(Ut, dt, apply_filter=(ftag == "f"))
        if k in (1, 5):
            U = Ut.cpu().numpy()
            refU = d[f"{key}_U{k}"]
            got = np.stack([U[0, :, 0], U[2, :, 0], U[3, :, 0]])
            e = [np.abs(got[f] - refU[f]).max() / np.abs(refU[f]).max() for f in range(3)]
            print(f"wall1d {key} step {k} {e}", flush=True)
            assert max(e) < 1e-11
            # every column is the same 1D solution
            assert np.abs(U - U[:, :, :1]).max() <= 1e-13 * np.abs(U).max()


def test_hybrid_entropy_wave_rhs_closed_form(dev):
    """rho = 1 + eps P(y) + delta sin(kx) at constant u, v, p: the HIP RHS
    against the exact -c (u0 sigma dx rho + v0 dy rho) (no walls)."""
    import numpy.polynomial.legendre as npleg
    import torch
    from oracle.spectral_np import Euler2DOracle
    og, osol, g, sol = _setup(32, 16, dev, y_bc="none")
    eqs = Euler2DOracle(1.4)
    u0, v0, p0, eps, delta, m = 0.3, -0.45, 2.5, 0.25, 0.02, 10
    coeffs = np.array([0.0, 0.7, -0.4, 0.25, 0.1, -0.3, 0.2, 0.15, -0.1, 0.05])
    t = 2.0 * (og.y - og.y[0]) / og.ly - 1.0
    P = npleg.legval(t, coeffs)
    dP = (2.0 / og.ly) * npleg.legval(t, npleg.legder(coeffs))
    k = 2 * np.pi * m / og.lx
    sig = np.exp(-36.0 * (m / 16.0) ** 16)
    X = og.x[None, :]
    rho = 1.0 + eps * P[:, None] + delta * np.sin(k * X)
    sh = rho.shape
    U = eqs.conservative(rho, np.full(sh, u0), np.full(sh, v0), np.full(sh, p0))
    c = np.array([1.0, u0, v0, 0.5 * (u0 * u0 + v0 * v0)])
    drho = u0 * delta * k * sig * np.cos(k * X) + v0 * eps * dP[:, None]
    expect = -c[:, None, None] * drho[None]
    r = sol.rhs(torch.as_tensor(U, device=dev)).cpu().numpy()
    err = np.abs(r - expect).max() / np.abs(expect).max()
    print(f"entropy wave rhs {err:.2e}", flush=True)
    assert err < 1e-11


def test_hybrid_c4_ny256_vs_oracle(dev):
    """The C4 size 256 x 256 (KHI, walls, filter) against the oracle, 3 steps
    with the per-step CFL dt."""
    import torch
    from oracle.hybrid_np import rhs_hybrid
    from phrike_amd.initial_conditions import khi_2d_gll
    og, osol, g, sol = _setup(256, 256, dev)
    U = khi_2d_gll(256, og.y, seed=11)
    Ut = torch.as_tensor(U, device=dev)
    e_r = field_errors(sol.rhs(Ut).cpu().numpy(), rhs_hybrid(og, osol.eqs, U))
    assert e_r.max() < 1e-11
    Uo = U
    for _ in range(3):
        dt = osol.compute_dt(Uo)
        Uo = osol.step(Uo, dt)
        Ut = sol.step(Ut, dt)
    e = field_errors(Ut.cpu().numpy(), Uo)
    print(f"hybrid 256x256 3 steps errors {e}", flush=True)
    assert e.max() < 1e-11


def test_hybrid_graph_replay_equals_eager(dev):
    """advance(..., graph_steps=4): the steps replayed from a captured HIP
    graph give the eager steps' states bit for bit, over several replays with
    eager work and synchronisations between them, FP64 and split-FP32 forms,
    fixed dt and the per-step device dt."""
    import torch
    from phrike_amd.initial_conditions import khi_2d_gll
    for prec, gll in (("double", "f64"), ("single", "split")):
        og, osol, g, sol = _setup(64, 32, dev, precision=prec, gll=gll, batch=2)
        U0 = torch.stack([torch.as_tensor(khi_2d_gll(64, g.y_nodes, seed=3 + i), device=dev)
                          for i in range(2)]).to(g.dtype)
        dt = sol.compute_dt(U0[0])
        for fixed in (dt, None):
            print("case graph replay", prec, gll, "fixed dt" if fixed else "device dt", flush=True)
            ref = sol.advance(U0, 13, dt=fixed)
            got = sol.advance(U0, 13, dt=fixed, graph_steps=4)  # 3 replays + 1 eager step
            torch.cuda.synchronize()
            assert torch.equal(ref, got)
            extra = sol.step(got, dt) * 0.5  # eager work between captures and replays
            torch.cuda.synchronize()
            again = sol.advance(U0, 13, dt=fixed, graph_steps=4)
            assert torch.equal(ref, again) and torch.isfinite(extra).all()
