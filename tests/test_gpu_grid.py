"""Grid3D spectral helpers of the drop-in facade (phrike/grid.py:1396-1401,
1497-1518): fftn / ifftn against numpy's transforms, k2 / k2_safe, and
_apply_masks against the oracle's mask and filter tables."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FILT = {"enabled": True, "p": 16, "alpha": 36.0}


@pytest.mark.parametrize("shape,prec", [((16, 8, 32), "double"), ((32, 32, 32), "double"), ((16, 16, 16), "single")])
def test_fftn_ifftn_vs_numpy(dev, shape, prec):
    import torch
    from phrike_amd.grid import Grid3D
    nz, ny, nx = shape
    g = Grid3D(nx, ny, nz, 2 * np.pi, 1.0, 3.0, filter_params=FILT, backend="hip", precision=prec,
               torch_device=str(dev))
    rng = np.random.default_rng(5)
    f = rng.standard_normal((3, nz, ny, nx))
    F = g.fftn(torch.as_tensor(f, device=dev)).cpu().numpy()
    ref = np.fft.fftn(f, axes=(-3, -2, -1))
    tol = 1e-12 if prec == "double" else 2e-5
    assert np.abs(F - ref).max() <= tol * np.abs(ref).max()
    # ifftn of an arbitrary (non-Hermitian) spectrum: the real part
    G = rng.standard_normal((2, nz, ny, nx)) + 1j * rng.standard_normal((2, nz, ny, nx))
    r = g.ifftn(torch.as_tensor(G, device=dev)).cpu().numpy()
    refr = np.fft.ifftn(G, axes=(-3, -2, -1)).real
    assert np.abs(r - refr).max() <= tol * np.abs(refr).max()
    # round trip
    back = g.ifftn(g.fftn(torch.as_tensor(f, device=dev))).cpu().numpy()
    assert np.abs(back - f).max() <= tol * np.abs(f).max()


def test_k2_and_masks(dev):
    import torch
    from oracle import spectral_np as sn
    from phrike_amd.grid import Grid3D
    nx, ny, nz = 16, 8, 32
    g = Grid3D(nx, ny, nz, 2 * np.pi, 1.0, 3.0, filter_params=FILT, backend="hip", torch_device=str(dev))
    og = sn.Grid3DOracle(nx, ny, nz, 2 * np.pi, 1.0, 3.0, filter_params=FILT, fft_workers=1)
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=2 * np.pi / nx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=1.0 / ny)
    kz = 2 * np.pi * np.fft.fftfreq(nz, d=3.0 / nz)
    k2 = kx[None, None, :] ** 2 + ky[None, :, None] ** 2 + kz[:, None, None] ** 2
    np.testing.assert_allclose(g.k2.cpu().numpy(), k2, rtol=1e-14)
    ks = g.k2_safe.cpu().numpy()
    assert ks[0, 0, 0] == 1.0 and np.array_equal(ks.ravel()[1:], k2.ravel()[1:])
    rng = np.random.default_rng(2)
    F = rng.standard_normal((2, nz, ny, nx)) + 1j * rng.standard_normal((2, nz, ny, nx))
    got = g._apply_masks(torch.as_tensor(F, device=dev)).cpu().numpy()
    ref = og._apply_masks(F.copy())
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=0)
