"""Snapshots written by the reference itself (tests/golden/ref_snapshots,
from phrike/io.py:103-236; meta is a pickled dict there) load without
unpickling, with grid and gamma recovered from the plain arrays, and restart
validation follows phrike/problems/base.py:257-290."""
import os
import types

import numpy as np
import pytest

from conftest import GOLDEN

SNAP = os.path.join(GOLDEN, "ref_snapshots")


def _grid(**kw):
    return types.SimpleNamespace(**kw)


@pytest.mark.parametrize("name,meta", [
    ("euler3d_snapshot_t0.125000.npz", {"Nx": 8, "Ny": 8, "Nz": 16, "Lx": 2 * np.pi, "Ly": np.pi,
                                        "Lz": 4 * np.pi, "gamma": 1.4}),
    ("euler2d_snapshot_t0.500000_khi.npz", {"Nx": 16, "Ny": 8, "Lx": 1.0, "Ly": 0.5, "gamma": 5.0 / 3.0}),
    ("mhd3d_snapshot_t0.250000.npz", {"Nx": 8, "Ny": 8, "Nz": 8, "Lx": 1.0, "Ly": 1.0, "Lz": 1.0,
                                      "gamma": 5.0 / 3.0, "mhd": True}),
])
def test_reference_snapshot_meta_recovered(name, meta):
    from phrike_amd.io import load_checkpoint
    path = os.path.join(SNAP, name)
    with np.load(path, allow_pickle=False) as d:
        with pytest.raises(ValueError):  # the reference's meta is a pickle
            d["meta"]
    c = load_checkpoint(path)
    assert c["meta"].pop("recovered") is True
    assert set(c["meta"]) == set(meta)
    for k, v in meta.items():
        if isinstance(v, float):
            assert abs(c["meta"][k] - v) <= 1e-12 * abs(v), k
        else:
            assert c["meta"][k] == v, k
    assert c["t"] == float(name.split("_t")[1][:8])
    assert c["U"].shape[0] in (4, 5, 8)


def test_validate_restart_data():
    from phrike_amd.io import load_restart_data, validate_restart_data
    d = load_restart_data(os.path.join(SNAP, "euler3d_snapshot_t0.125000.npz"))
    assert d["t0"] == 0.125
    eq = types.SimpleNamespace(gamma=1.4)
    assert validate_restart_data(d, _grid(Nx=8, Ny=8, Nz=16), eq) == []
    w = validate_restart_data(d, _grid(Nx=8, Ny=8, Nz=16), types.SimpleNamespace(gamma=5.0 / 3.0))
    assert len(w) == 1 and "Gamma mismatch" in w[0]
    with pytest.raises(ValueError, match="Nz mismatch"):
        validate_restart_data(d, _grid(Nx=8, Ny=8, Nz=8), eq)
    with pytest.raises(ValueError, match="Nx mismatch"):
        validate_restart_data(d, _grid(Nx=16, Ny=8, Nz=16), eq)


def test_own_snapshot_meta_json(tmp_path):
    """Files written here carry meta twice: the reference's dict (pickled,
    for the reference's loader) and meta_json, which this loader reads."""
    from phrike_amd.io import load_checkpoint
    meta = {"Nx": 4, "Ny": 2, "Lx": 1.0, "Ly": 2.0, "gamma": 1.4}
    import json
    U = np.ones((4, 2, 4))
    np.savez(tmp_path / "s.npz", t=0.5, U=U, x=np.arange(4) / 4, y=np.arange(2.0), rho=U[0], ux=U[1],
             uy=U[2], p=U[3], meta=meta, meta_json=json.dumps(meta))
    c = load_checkpoint(str(tmp_path / "s.npz"))
    assert c["meta"] == meta and c["t"] == 0.5
