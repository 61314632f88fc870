"""Snapshots in the reference's .npz layout (phrike/io.py:103-236), written
from device state on a side stream."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_snapshot_roundtrip(dev, tmp_path):
    import torch
    from phrike_amd.grid import Grid3D, Grid2D
    from phrike_amd.equations import EulerEquations3D, EulerEquations2D
    from phrike_amd.io import save_solution_snapshot, load_checkpoint, SnapshotWriter
    from phrike_amd.initial_conditions import taylor_green_3d, khi_2d_periodic
    g = Grid3D(16, 16, 16, 2 * np.pi, 2 * np.pi, 2 * np.pi, backend="hip")
    e = EulerEquations3D(1.4)
    U = torch.as_tensor(taylor_green_3d(16, 16, 16), device=dev)
    f = save_solution_snapshot(str(tmp_path), 0.125, U, g, e)
    assert f.endswith("snapshot_t0.125000.npz")
    d = load_checkpoint(f)
    np.testing.assert_array_equal(d["U"], U.cpu().numpy())
    assert d["meta"] == {"Nx": 16, "Ny": 16, "Nz": 16, "Lx": 2 * np.pi, "Ly": 2 * np.pi,
                         "Lz": 2 * np.pi, "gamma": 1.4}
    assert set(d["primitive_vars"]) == {"rho", "ux", "uy", "uz", "p"}
    with np.load(f, allow_pickle=False) as z:
        assert {"t", "U", "x", "y", "z", "rho", "ux", "uy", "uz", "p", "meta", "created"} <= set(z.files)
    # asynchronous 2D write
    g2 = Grid2D(32, 16, 1.0, 1.0, backend="hip")
    V = torch.as_tensor(khi_2d_periodic(32, 16, seed=1), device=dev)
    V0 = V.clone()
    w = SnapshotWriter()
    f2 = w.save(str(tmp_path), 0.5, V, g2, EulerEquations2D(1.4), tag="run7", blocking=False)
    V.add_(1.0)  # later device work does not affect the snapshot of the earlier state
    w.wait()
    d2 = load_checkpoint(f2)
    assert f2.endswith("snapshot_t0.500000_run7.npz")
    np.testing.assert_array_equal(d2["U"], V0.cpu().numpy())


def test_restart_continues_bit_for_bit(dev, tmp_path):
    """A run stopped at a checkpoint and restarted from the file
    (load_restart_data / validate_restart_data, phrike/problems/base.py:
    238-290) ends in the same bits as the same run continued from memory:
    the file holds the state exactly and the restarted loop takes its first
    dt from it.  A grid of another size is refused."""
    import torch
    from phrike_amd.equations import EulerEquations3D
    from phrike_amd.grid import Grid3D
    from phrike_amd.initial_conditions import taylor_green_3d
    from phrike_amd.io import restart_state
    from phrike_amd.solver import SpectralSolver3D
    n = 32
    fp = {"enabled": True, "p": 16, "alpha": 36.0}
    g = Grid3D(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, filter_params=fp, backend="hip", torch_device=str(dev))
    e = EulerEquations3D(1.4)
    U0 = torch.as_tensor(taylor_green_3d(n, n, n), device=dev)
    first = SpectralSolver3D(g, e, cfl=0.4)
    first.run(U0, 0.0, 0.03, output_interval=1.0, checkpoint_interval=0.03, outdir=str(tmp_path))
    files = sorted(p.name for p in tmp_path.iterdir())
    assert files == ["snapshot_t0.030000.npz"], files
    U1, t1 = restart_state(str(tmp_path / files[0]), g, e)
    assert t1 == first.t and torch.equal(U1, first.U)
    second = SpectralSolver3D(g, e, cfl=0.4)
    second.run(U1, t1, 0.06, output_interval=1.0)
    ref = SpectralSolver3D(g, e, cfl=0.4)
    ref.run(U0, 0.0, 0.03, output_interval=1.0)
    ref.run(ref.U, ref.t, 0.06, output_interval=1.0)
    assert second.t == ref.t and torch.equal(second.U, ref.U)
    with pytest.raises(ValueError):
        restart_state(str(tmp_path / files[0]),
                      Grid3D(16, 16, 16, 2 * np.pi, 2 * np.pi, 2 * np.pi, backend="hip", torch_device=str(dev)), e)


def test_mhd_snapshot_layout(dev, tmp_path):
    """MHD snapshots carry the reference's MHD keys (phrike/io.py:122-167):
    B components, a div(B) diagnostic and meta with mhd=True."""
    import torch
    from phrike_amd.grid import Grid3D
    from phrike_amd.io import load_checkpoint, save_solution_snapshot
    from phrike_amd.mhd import MHDEquations3D, SpectralSolverMHD3D
    g = Grid3D(16, 16, 16, 1.0, 1.0, 1.0, backend="hip", torch_device=str(dev))
    e = MHDEquations3D(5.0 / 3.0)
    s = SpectralSolverMHD3D(g, e, cfl=0.3)
    X, _, _ = g.xyz_mesh()
    k = 2 * np.pi
    by, bz = 0.1 * torch.cos(k * X), 0.1 * torch.sin(k * X)
    U = e.conservative(torch.ones_like(X), torch.zeros_like(X), -by, -bz, torch.full_like(X, 0.1),
                       torch.ones_like(X), by, bz)
    U = s._project_B(U)
    f = save_solution_snapshot(str(tmp_path), 0.25, U, g, e)
    d = load_checkpoint(f)
    assert d["meta"] == {"mhd": True, "gamma": 5.0 / 3.0, "Lx": 1.0, "Nx": 16, "Ny": 16, "Nz": 16, "Ly": 1.0,
                         "Lz": 1.0}
    np.testing.assert_array_equal(d["Bx"], U[5].cpu().numpy())
    assert np.abs(d["div_b"]).max() < 1e-12
