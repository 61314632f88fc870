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
