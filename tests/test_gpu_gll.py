"""LGL D.U on matrix cores vs the oracle (FP64 MFMA to 1e-13 of |D||F|;
split FP32 reported and bounded)."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ny,nx,nb", [(16, 16, 1), (128, 256, 4), (256, 256, 4), (64, 48, 3)])
def test_gll_apply_modes(dev, ny, nx, nb):
    import torch
    from oracle import gll_np as gn
    from phrike_amd.basis import LegendreLobattoBasis1D
    b = LegendreLobattoBasis1D(ny, 1.0)
    ob = gn.LGLBasisOracle(ny, 1.0)
    np.testing.assert_allclose(b.D_x, ob.D_x, rtol=0, atol=1e-13 * np.abs(ob.D_x).max())
    rng = np.random.default_rng(ny + nx)
    y = ob.x
    F = np.stack([np.cos(2 * np.pi * (k + 1) * y)[:, None] * rng.standard_normal(nx)[None, :]
                  for k in range(nb)])
    ref = gn.apply_along_y(ob.D_x, F)
    scale = np.abs(ob.D_x).sum(axis=1).max() * np.abs(F).max()
    Ft = torch.as_tensor(F, device=dev)
    got = b.apply(Ft, "f64").cpu().numpy()
    e64 = np.abs(got - ref).max() / scale
    got_s = b.apply(Ft, "split").cpu().numpy()
    es = np.abs(got_s - ref).max() / scale
    got_32 = b.apply(Ft.float(), "split").double().cpu().numpy()
    e32 = np.abs(got_32 - ref).max() / scale
    print(f"gll ny={ny} nx={nx}: f64 {e64:.2e} split(f64 io) {es:.2e} split(f32) {e32:.2e}", flush=True)
    assert e64 < 1e-14
    assert es < 1e-6
    assert e32 < 1e-6
    # epilogue: -(add + D F)
    A = torch.ones_like(Ft)
    got_e = b.apply(Ft, "f64", add=A).cpu().numpy()
    assert np.abs(got_e + (1.0 + ref)).max() / scale < 1e-14


def test_gll_reference_tables_on_device(dev):
    import torch
    d = golden("lgl.npz")
    from phrike_amd.basis import LegendreLobattoBasis1D
    b = LegendreLobattoBasis1D(128, 1.0)
    D = d["N128_D"]
    F = np.random.default_rng(0).standard_normal((2, 128, 32))
    got = b.apply(torch.as_tensor(F, device=dev), "f64").cpu().numpy()
    ref = np.einsum("ij,bjx->bix", D, F)
    assert np.abs(got - ref).max() / (np.abs(D).sum(axis=1).max() * np.abs(F).max()) < 1e-13


@pytest.mark.parametrize("ny,nx,nb", [(256, 256, 16), (128, 64, 128)])
def test_gll_tiled_batch_matches_single(dev, ny, nx, nb):
    """A batch large enough for the LDS-tiled kernel gives, field by field,
    exactly what single-field launches (the per-wave kernel) give: both walk
    k in the same order.  Also checked against the oracle."""
    import torch
    from oracle import gll_np as gn
    from phrike_amd.basis import LegendreLobattoBasis1D
    b = LegendreLobattoBasis1D(ny, 1.0)
    ob = gn.LGLBasisOracle(ny, 1.0)
    rng = np.random.default_rng(7 * ny + nb)
    F = rng.standard_normal((nb, ny, nx)) + 3.0
    Ft = torch.as_tensor(F, device=dev)
    A = torch.as_tensor(rng.standard_normal((nb, ny, nx)), device=dev)
    ref = gn.apply_along_y(ob.D_x, F)
    scale = np.abs(ob.D_x).sum(axis=1).max() * np.abs(F).max()
    for mode, X, add in (("f64", Ft, None), ("split", Ft, None), ("split", Ft.float(), None),
                         ("f64", Ft, A), ("split", Ft.float(), A.float())):
        whole = b.apply(X, mode, add=add)
        for i in (0, nb // 2, nb - 1):
            one = b.apply(X[i:i + 1].contiguous(), mode,
                          add=None if add is None else add[i:i + 1].contiguous())
            assert torch.equal(whole[i:i + 1], one), (mode, X.dtype, add is None, i)
        if add is None:
            err = np.abs(whole.double().cpu().numpy() - ref).max() / scale
            assert err < (1e-14 if mode == "f64" else 1e-6), (mode, X.dtype, err)
