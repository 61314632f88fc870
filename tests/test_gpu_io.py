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
