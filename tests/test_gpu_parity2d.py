"""Batched 2D Fourier hip backend vs the reference's golden vectors."""
import numpy as np
import pytest

from conftest import golden, rel_l2, field_errors, group_errors

pytestmark = pytest.mark.gpu
FILT = {"enabled": True, "p": 16, "alpha": 36.0}


def _setup(nx, ny, interval=1, precision="double", batch=1):
    from phrike_amd.grid import Grid2D
    from phrike_amd.equations import EulerEquations2D
    from phrike_amd.solver import SpectralSolver2D
    g = Grid2D(nx, ny, 1.0, 1.0, dealias=True, filter_params=dict(FILT, interval=interval),
               backend="hip", precision=precision, batch=batch)
    e = EulerEquations2D(1.4)
    return g, e, SpectralSolver2D(g, e, scheme="rk4", cfl=0.4)


@pytest.mark.parametrize("tag,nx,ny,interval", [("sq64", 64, 64, 1), ("rect32x16", 32, 16, 1),
                                                 ("cad2_32", 32, 32, 2)])
def test_khi_steps_vs_reference(dev, tag, nx, ny, interval):
    import torch
    d = golden("khi2d.npz")
    g, e, s = _setup(nx, ny, interval)
    U = torch.as_tensor(d[f"{tag}_U0"], device=dev)
    print(f"case {tag}", flush=True)
    _v = rel_l2(s.rhs(U).cpu(), d[f"{tag}_rhs0"])
    print("check", 'rel_l2(s.rhs(U).cpu(), d[f"{tag}_rhs0"])'[:60], float(_v), flush=True)
    assert _v < 1e-11
    for k in range(10):
        dt = s.compute_dt(U)
        assert abs(dt / d[f"{tag}_dts"][k] - 1) < 1e-12
        U = s.step(U, float(d[f"{tag}_dts"][k]))
        if k == 0:
            _v = field_errors(U.cpu(), d[f"{tag}_U1"]).max()
            print("check", 'field_errors(U.cpu(), d[f"{tag}_U1"]).max()'[:60], float(_v), flush=True)
            assert _v < 1e-12
    _v = field_errors(U.cpu(), d[f"{tag}_U10"]).max()
    print("check", 'field_errors(U.cpu(), d[f"{tag}_U10"]).max()'[:60], float(_v), flush=True)
    assert _v < 1e-11


def test_batch_equals_alone(dev):
    """A run gives the same bits alone and inside a batch of other runs."""
    import torch
    from phrike_amd.initial_conditions import khi_2d_periodic
    g1, e1, s1 = _setup(64, 64)
    gb, eb, sb = _setup(64, 64, batch=3)
    runs = torch.stack([torch.as_tensor(khi_2d_periodic(64, 64, seed=k), device=dev) for k in range(3)])
    dts = torch.tensor([1e-3, 2e-3, 1.5e-3], dtype=torch.float64, device=dev)
    outb = torch.empty_like(runs)
    sb.step_into(runs, outb, 0.0, dt_dev=dts)
    for k in range(3):
        alone = s1.step(runs[k], float(dts[k].item()))
        assert torch.equal(alone, outb[k]), k
    # inactive runs are left untouched
    active = torch.tensor([1, 0, 1], dtype=torch.int32, device=dev)
    outc = torch.full_like(runs, 7.0)
    sb.step_into(runs, outc, 0.0, dt_dev=dts, active_dev=active)
    assert torch.equal(outc[0], outb[0]) and bool((outc[1] == 7.0).all())
    # per-run CFL dt on device equals the one-run compute_dt
    dtr = sb.dt_runs(runs)
    for k in range(3):
        assert abs(float(dtr[k]) / s1.compute_dt(runs[k]) - 1) < 1e-15


def test_fp32_vs_fp64(dev):
    import torch
    d = golden("khi2d.npz")
    g, e, s = _setup(64, 64, precision="single")
    U = torch.as_tensor(d["sq64_U0"], device=dev, dtype=torch.float32)
    for k in range(10):
        U = s.step(U, float(d["sq64_dts"][k]))
    errs = group_errors(U.double().cpu(), d["sq64_U10"], [(0,), (1, 2), (3,)])
    print("fp32 KHI 64^2 x 10 per-field rel L2:", field_errors(U.double().cpu(), d["sq64_U10"]),
          "group-normalised:", errs, flush=True)
    assert errs.max() < 1e-5


@pytest.mark.parametrize("nx,ny", [(1024, 32), (512, 16)])
def test_long_x_lines_vs_oracle(dev, nx, ny):
    """x lines of 512 and 1024 points (the widest x kernels: in FP64 one
    thread per x position of a row pair, in FP32 one per two) against the
    oracle: two steps of a periodic KHI, two runs in one batch; FP32 bounded
    against the same oracle."""
    import torch
    from oracle import spectral_np as sn
    from phrike_amd.initial_conditions import khi_2d_periodic
    og = sn.Grid2DOracle(nx, ny, 1.0, 1.0, filter_params=FILT, fft_workers=1)
    osol = sn.Solver2DOracle(og, sn.Euler2DOracle(1.4), cfl=0.4)
    V = [khi_2d_periodic(nx, ny, seed=s) for s in (3, 4)]
    dt = osol.compute_dt(V[0])
    ref = [osol.step(osol.step(v, dt), dt) for v in V]
    for prec, tol in (("double", 1e-12), ("single", 1e-5)):
        g, e, s = _setup(nx, ny, precision=prec, batch=2)
        U = torch.as_tensor(np.stack(V), device=dev).to(g.dtype)
        print(f"case {nx}x{ny} {prec}", flush=True)
        for _ in range(2):
            U = s.step(U, dt)
        for k in range(2):
            if prec == "double":
                err = field_errors(U[k].cpu(), ref[k]).max()
            else:
                err = group_errors(U[k].double().cpu(), ref[k], [(0,), (1, 2), (3,)]).max()
            print(f"  run {k}: {err:.2e}", flush=True)
            assert err < tol, (prec, k, err)
