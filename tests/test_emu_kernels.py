"""The HIP kernel sources, compiled for the CPU against the emulation header
(tests/emu), checked against the reference's golden vectors at tiny sizes.
Runs without a GPU; catches index and bounds mistakes before a GPU run."""
import os
import shutil
import sys

import numpy as np
import pytest

from conftest import golden, field_errors, group_errors

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu"))

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def emu():
    from emu_lib import load
    return load()


def test_box_step_fp64(emu):
    from emu_lib import Emu3D
    d = golden("rhs3d.npz")
    e = Emu3D(emu, 16, 8, 32)
    U = d["box16x8x32_U"]
    assert field_errors(e.rhs(U), d["box16x8x32_rhs"]).max() < 1e-11
    assert field_errors(e.step(U, float(d["box16x8x32_dt"])), d["box16x8x32_step"]).max() < 1e-12


def test_cube_step_fp32(emu):
    from emu_lib import Emu3D
    d = golden("rhs3d.npz")
    e = Emu3D(emu, 16, 16, 16, dtype=np.float32)
    got = e.step(d["cube16_U"], float(d["cube16_dt"])).astype(np.float64)
    assert group_errors(got, d["cube16_step"], [(0,), (1, 2, 3), (4,)]).max() < 1e-5


def test_khi_rect_step_and_batch(emu):
    from emu_lib import Emu2D
    d = golden("khi2d.npz")
    e = Emu2D(emu, 32, 16, batch=2)
    U0 = d["rect32x16_U0"]
    dt = float(d["rect32x16_dts"][0])
    assert field_errors(e.step(U0, dt), d["rect32x16_U1"]).max() < 1e-12
    pair = np.stack([U0, U0[:, ::-1, :].copy()])
    both = e.step(pair, dt)
    np.testing.assert_array_equal(both[0], e.step(U0, dt))
