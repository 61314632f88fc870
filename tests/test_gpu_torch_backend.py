"""The reference's torch-backend arithmetic, restated on torch.fft (bench.py
torch_backend_step, the second GPU reference point of the benchmark), pinned
against the reference's own TGV 32^3 run (10 steps, written by the reference's
NumPy path) and against the HIP step."""
import os
import sys

import numpy as np
import pytest

from conftest import field_errors, golden

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def test_torch_backend_matches_reference_and_hip(dev):
    import torch
    from bench import torch_backend_step
    from phrike_amd.equations import EulerEquations3D
    from phrike_amd.grid import Grid3D
    from phrike_amd.solver import SpectralSolver3D
    d = golden("tgv3d_32.npz")
    dt = float(d["dt"])
    step = torch_backend_step(32, dev)
    U = torch.as_tensor(d["U0"], device=dev)
    g = Grid3D(32, 32, 32, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
               filter_params={"enabled": True, "p": 16, "alpha": 36.0}, backend="hip")
    s = SpectralSolver3D(g, EulerEquations3D(1.4), scheme="rk4", cfl=0.4)
    V = U.clone()
    for _ in range(10):
        U = step(U, dt)
        V = s.step(V, dt)
    e_ref = field_errors(U.cpu().numpy(), d["U_10"])
    e_hip = field_errors(U.cpu().numpy(), V.cpu().numpy())
    print("torch backend vs reference:", e_ref, "vs hip:", e_hip, flush=True)
    assert e_ref.max() < 1e-12 and e_hip.max() < 1e-12
