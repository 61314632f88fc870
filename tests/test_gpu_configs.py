"""The benchmark configurations at their own sizes:

* C2: TGV 512^3 FP64, 200 RK4 steps at the frozen CFL dt: mass and energy
  drift at round-off and the TGV's reflection symmetries kept;
* C3: decaying turbulence 256^3, FP32 bulk (correction-form filter) against
  FP64 after 200 steps, every field within 1e-5 relative L2; one FP64 step
  at 128^3 against the oracle;
* C5: the ensemble server in FP32 on 128^2 and 1024^2 requests: every result
  the same bits alone and in the mix, and a 128^2 run against the oracle's
  run loop within 1e-5."""
import numpy as np
import pytest

from conftest import field_errors, group_errors

pytestmark = pytest.mark.gpu

FILT = {"enabled": True, "p": 16, "alpha": 36.0}


def _mirror(a, dim):
    """a(-x) on the periodic grid x_i = 2 pi i / N: index N - i (mod N)."""
    import torch
    return torch.roll(torch.flip(a, dims=(dim,)), shifts=1, dims=dim)


def test_c2_tgv512_200_steps_invariants(dev):
    import torch
    from phrike_amd.equations import EulerEquations3D
    from phrike_amd.grid import Grid3D
    from phrike_amd.initial_conditions import taylor_green_3d
    from phrike_amd.solver import SpectralSolver3D
    n = 512
    g = Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, filter_params=FILT, backend="hip", torch_device=str(dev))
    s = SpectralSolver3D(g, EulerEquations3D(1.4), cfl=0.4)
    U = torch.as_tensor(taylor_green_3d(n, n, n), device=dev)
    dt = s.compute_dt(U)
    c0 = s.conserved_sums(U)
    out = torch.empty_like(U)
    for _ in range(200):
        s.step_into(U, out, dt)
        U, out = out, U
    c1 = s.conserved_sums(U)
    mass_drift = abs(c1[0] / c0[0] - 1)
    energy_drift = abs(c1[4] / c0[4] - 1)
    print(f"TGV 512^3 x 200: mass drift {mass_drift:.2e}, energy drift {energy_drift:.2e}", flush=True)
    assert mass_drift < 1e-13 and energy_drift < 1e-6
    assert torch.isfinite(U).all()
    # reflections x -> -x, y -> -y, z -> -z: the reflected velocity component
    # changes sign, every other field is even
    scale = U.abs().amax(dim=(1, 2, 3))
    for dim, odd in ((3, 1), (2, 2), (1, 3)):
        for f in range(5):
            m = _mirror(U[f], dim - 1)
            d = (U[f] + m) if f == odd else (U[f] - m)
            assert float(d.abs().max()) <= 1e-11 * float(scale[f]) + 1e-300, (dim, f)


def test_c3_turbulence256_fp32_vs_fp64_200_steps(dev):
    import torch
    from phrike_amd.equations import EulerEquations3D
    from phrike_amd.grid import Grid3D
    from phrike_amd.initial_conditions import decaying_turbulence_3d
    from phrike_amd.solver import SpectralSolver3D
    n = 256
    U0 = decaying_turbulence_3d(n)
    res = {}
    for prec in ("double", "single"):
        g = Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, filter_params=FILT, backend="hip",
                   precision=prec, torch_device=str(dev))
        s = SpectralSolver3D(g, EulerEquations3D(1.4), cfl=0.4)
        U = torch.as_tensor(U0, device=dev).to(g.dtype)
        if prec == "double":
            dt = s.compute_dt(U)
        for _ in range(200):
            U = s.step(U, dt)  # the solver's own output: FP32 filters in correction form
        res[prec] = U.double().cpu().numpy()
        del U, s, g
        torch.cuda.empty_cache()
    e = field_errors(res["single"], res["double"])
    print(f"turbulence 256^3 x 200, FP32 vs FP64 rel L2 per field: {e}", flush=True)
    assert e.max() < 1e-5


def test_c3_fp64_step_128_vs_oracle(dev):
    import torch
    from oracle import spectral_np as sn
    from phrike_amd.equations import EulerEquations3D
    from phrike_amd.grid import Grid3D
    from phrike_amd.initial_conditions import decaying_turbulence_3d
    from phrike_amd.solver import SpectralSolver3D
    n = 128
    U0 = decaying_turbulence_3d(n)
    og = sn.Grid3DOracle(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, filter_params=FILT, fft_workers=8)
    osv = sn.Solver3DOracle(og, sn.Euler3DOracle(1.4), cfl=0.4)
    dt = osv.compute_dt(U0)
    ref = osv.step(U0, dt)
    g = Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, filter_params=FILT, backend="hip", torch_device=str(dev))
    s = SpectralSolver3D(g, EulerEquations3D(1.4), cfl=0.4)
    got = s.step(torch.as_tensor(U0, device=dev), dt).cpu().numpy()
    e = field_errors(got, ref)
    print(f"turbulence 128^3 one FP64 step vs oracle: {e}", flush=True)
    assert e.max() < 1e-12


def test_c5_fp32_server_128_1024(dev):
    from oracle import spectral_np as sn
    from phrike_amd.initial_conditions import khi_2d_periodic
    from phrike_amd.serving import EnsembleServer, RunRequest
    reqs = [RunRequest(0, 128, 21, 0.02), RunRequest(1, 1024, 22, 0.0015), RunRequest(2, 128, 23, 0.012),
            RunRequest(3, 1024, 24, 0.001), RunRequest(4, 128, 25, 0.006)]

    def server(caps):
        return EnsembleServer(caps, precision="single", steps_per_replay=4, keep_states=True)

    mixed = server({128: 2, 1024: 2}).serve(reqs, realtime=False)
    assert [r.rid for r in mixed] == [0, 1, 2, 3, 4] and not any(r.failed for r in mixed)
    for k in (0, 1):
        alone = server({reqs[k].n: 1}).serve([reqs[k]], realtime=False)[0]
        np.testing.assert_array_equal(alone.state, mixed[k].state)
        assert alone.steps == mixed[k].steps and alone.t_final == mixed[k].t_final
    # the 128^2 run against the oracle's run loop (FP64, phrike/solver.py:764-797)
    og = sn.Grid2DOracle(128, 128, 1.0, 1.0, filter_params=dict(FILT, interval=1), fft_workers=4)
    osv = sn.Solver2DOracle(og, sn.Euler2DOracle(1.4), cfl=0.4)
    ref, nsteps = osv.run_to(khi_2d_periodic(128, 128, seed=21), 0.02)
    print(f"C5 128^2: {nsteps} oracle steps, {mixed[0].steps} server steps", flush=True)
    assert nsteps == mixed[0].steps
    # y momentum is ~1e-2 of the x momentum: measured on the momentum scale
    e = group_errors(mixed[0].state, ref, [(0,), (1, 2), (3,)])
    print(f"C5 128^2 FP32 server vs FP64 oracle: {e} (per field {field_errors(mixed[0].state, ref)})",
          flush=True)
    assert e.max() < 1e-5
