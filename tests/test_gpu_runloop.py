"""The run loops on the device (ph_run3d / ph_run2d) against the reference's
host loop restated on the oracle (phrike/solver.py:1395-1513 for 3D, 706-812
for 2D): the same steps (dt from the CFL condition after every step, the
last step clipped to t_end), outputs at the same times, the same state;
callbacks get a copy of each step's state."""
import numpy as np
import pytest

from conftest import field_errors

pytestmark = pytest.mark.gpu

FILT = {"enabled": True, "p": 16, "alpha": 36.0}


def _oracle_loop(osv, U0, t0, t_end, output_interval):
    """phrike/solver.py:1395-1513 on the oracle: returns the states and dts
    of every step and the output times."""
    U, t, dt = U0, t0, osv.compute_dt(U0)
    next_output = t0 + output_interval
    states, dts, outs = [], [], [t0]
    while t < t_end - 1e-12:
        dt = min(dt, t_end - t)
        U = osv.step(U, dt)
        t += dt
        states.append(U)
        dts.append(dt)
        dt = osv.compute_dt(U)
        if t + 1e-12 >= next_output:
            outs.append(t)
            next_output += output_interval
    outs.append(t)
    return U, t, states, dts, outs


def test_run3d_device_loop_vs_oracle(dev, tmp_path):
    import torch
    from oracle import spectral_np as sn
    from phrike_amd.equations import EulerEquations3D
    from phrike_amd.grid import Grid3D
    from phrike_amd.initial_conditions import taylor_green_3d
    from phrike_amd.io import load_checkpoint
    n, cfl = 32, 0.4
    U0 = taylor_green_3d(n, n, n)
    og = sn.Grid3DOracle(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, filter_params=FILT, fft_workers=4)
    osv = sn.Solver3DOracle(og, sn.Euler3DOracle(1.4), cfl=cfl)
    dt0 = osv.compute_dt(U0)
    t_end, oi = 11.3 * dt0, 3.2 * dt0
    Uo, to, states, dts, outs = _oracle_loop(osv, U0, 0.0, t_end, oi)
    g = Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, filter_params=FILT, backend="hip", torch_device=str(dev))
    from phrike_amd.solver import SpectralSolver3D
    s = SpectralSolver3D(g, EulerEquations3D(1.4), cfl=cfl)
    got = []
    h = s.run(torch.as_tensor(U0, device=dev), 0.0, t_end, output_interval=oi,
              on_step=lambda k, dt, U: got.append((k, dt, U)), checkpoint_interval=5.0 * dt0,
              outdir=str(tmp_path))
    print(f"3D run: {len(dts)} oracle steps, {len(got)} device steps, outputs {len(outs)}", flush=True)
    assert [k for k, _, _ in got] == list(range(len(dts)))
    np.testing.assert_allclose([d for _, d, _ in got], dts, rtol=1e-13, atol=0)
    # every callback state is its own tensor: stored states are each step's
    for (k, _, U), ref in zip(got, states):
        assert field_errors(U.cpu().numpy(), ref).max() < 1e-11, k
    assert abs(s.t - to) <= 1e-14 * to
    np.testing.assert_allclose(h["time"], outs, rtol=1e-13, atol=0)
    assert field_errors(s.U.cpu().numpy(), Uo).max() < 1e-11
    np.testing.assert_allclose(h["mass"][0], float(U0[0].sum()), rtol=1e-14)
    assert max(abs(m / h["mass"][0] - 1) for m in h["mass"]) < 1e-13
    # checkpoints at 5 dt0 and 10 dt0 (the reference's next_checkpoint rule)
    import phrike_amd.io as pio
    if pio._default_writer is not None:
        pio._default_writer.wait()
    files = sorted(p.name for p in tmp_path.iterdir())
    assert len(files) == 2, files
    ck = load_checkpoint(str(tmp_path / files[0]))
    k = int(np.argmin([abs(sum(dts[:i + 1]) - ck["t"]) for i in range(len(dts))]))
    assert field_errors(ck["U"], states[k]).max() < 1e-11


def test_run3d_without_callbacks_same_result(dev):
    """Chunked launches (no on_step) give the same bits as one step per
    launch."""
    import torch
    from phrike_amd.equations import EulerEquations3D
    from phrike_amd.grid import Grid3D
    from phrike_amd.initial_conditions import taylor_green_3d
    from phrike_amd.solver import SpectralSolver3D
    n = 32
    g = Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, filter_params=FILT, backend="hip", torch_device=str(dev))
    U0 = torch.as_tensor(taylor_green_3d(n, n, n), device=dev)
    a = SpectralSolver3D(g, EulerEquations3D(1.4), cfl=0.4)
    ha = a.run(U0, 0.0, 0.05, output_interval=0.013)
    b = SpectralSolver3D(g, EulerEquations3D(1.4), cfl=0.4)
    hb = b.run(U0, 0.0, 0.05, output_interval=0.013, on_step=lambda k, dt, U: None)
    assert ha == hb and a.t == b.t
    assert torch.equal(a.U, b.U)


def test_run2d_device_loop_vs_oracle(dev):
    import torch
    from oracle import spectral_np as sn
    from phrike_amd.equations import EulerEquations2D
    from phrike_amd.grid import Grid2D
    from phrike_amd.initial_conditions import khi_2d_periodic
    from phrike_amd.solver import SpectralSolver2D
    n, cfl = 64, 0.4
    U0 = khi_2d_periodic(n, n, seed=4)
    og = sn.Grid2DOracle(n, n, 1.0, 1.0, filter_params=FILT, fft_workers=1)
    osv = sn.Solver2DOracle(og, sn.Euler2DOracle(1.4), cfl=cfl)
    t_end = 0.02
    Uo, to, states, dts, outs = _oracle_loop(osv, U0, 0.0, t_end, 0.007)
    g = Grid2D(n, n, 1.0, 1.0, filter_params=FILT, backend="hip", torch_device=str(dev))
    s = SpectralSolver2D(g, EulerEquations2D(1.4), cfl=cfl)
    got = []
    h = s.run(torch.as_tensor(U0, device=dev), 0.0, t_end, output_interval=0.007,
              on_step=lambda k, dt, U: got.append(dt))
    print(f"2D run: {len(dts)} steps, outputs {h['time']}", flush=True)
    np.testing.assert_allclose(got, dts, rtol=1e-13, atol=0)
    np.testing.assert_allclose(h["time"], outs, rtol=1e-13, atol=0)
    assert field_errors(s.U.cpu().numpy(), Uo).max() < 1e-11
