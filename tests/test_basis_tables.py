"""Host-side LGL tables of the shipped package vs the reference's own values."""
import numpy as np
import pytest

from conftest import golden


@pytest.mark.parametrize("n", [8, 16, 32, 128])
def test_tables_match_reference(n):
    from phrike_amd.basis import LegendreLobattoBasis1D, split_hi_lo
    d = golden("lgl.npz")
    b = LegendreLobattoBasis1D(n, 1.0)
    np.testing.assert_array_equal(b.nodes(), d[f"N{n}_x"])
    np.testing.assert_array_equal(b.quadrature_weights(), d[f"N{n}_w"])
    np.testing.assert_allclose(b.D_x, d[f"N{n}_D"], rtol=0, atol=1e-13 * np.abs(d[f"N{n}_D"]).max())
    np.testing.assert_allclose(b.nodal_filter(16, 36.0), d[f"N{n}_filter_p16_a36"], rtol=0, atol=1e-13)
    hi, lo = split_hi_lo(b.D_x)
    # hi + lo carries ~48 bits of D
    assert np.abs(hi.astype(np.float64) + lo - b.D_x).max() <= 2.0 ** -46 * np.abs(b.D_x).max()
