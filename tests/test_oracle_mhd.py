"""The 3D ideal-MHD oracle (oracle/mhd_np.py) against golden vectors written
from the reference (tests/golden/mhd3d.npz) and against closed forms."""
import numpy as np

from conftest import golden, field_errors


def _grids():
    from oracle import spectral_np as sn
    alf = sn.Grid3DOracle(32, 8, 8, 1.0, 1.0, 1.0, dealias=True,
                          filter_params={"enabled": True, "p": 8, "alpha": 36.0}, fft_workers=1)
    rnd = sn.Grid3DOracle(16, 16, 16, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=True,
                          filter_params={"enabled": True, "p": 16, "alpha": 36.0}, fft_workers=1)
    return alf, rnd


def test_oracle_mhd_vs_reference_goldens():
    from oracle import mhd_np as mn
    d = golden("mhd3d.npz")
    alf, rnd = _grids()
    e = mn.MHD3DOracle(5.0 / 3.0)
    U = d["rnd_U"]
    Fx, Fy, Fz = e.flux(U)
    for a, b in ((Fx, d["rnd_Fx"]), (Fy, d["rnd_Fy"]), (Fz, d["rnd_Fz"])):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-14 * np.abs(b).max())
    assert field_errors(mn.rhs_mhd3d(rnd, e, U), d["rnd_rhs"]).max() < 1e-12
    P = np.array(mn.project_solenoidal(rnd, U[5], U[6], U[7]))
    assert field_errors(P, d["rnd_proj"]).max() < 1e-13
    np.testing.assert_allclose(mn.divergence(rnd, U[5], U[6], U[7]), d["rnd_div"], atol=1e-12)
    assert abs(e.max_wave_speed(U) - float(d["rnd_maxc"])) < 1e-14 * float(d["rnd_maxc"])
    cq = e.conserved_quantities(U)
    ref = dict(zip([str(k) for k in d["rnd_cons_keys"]], d["rnd_cons"]))
    for k, v in ref.items():
        assert abs(cq[k] - v) <= 1e-12 * max(abs(v), 1.0), k
    s = mn.SolverMHD3DOracle(rnd, e, cfl=0.3)
    dt = s.compute_dt(U)
    assert abs(dt - float(d["rnd_dt"])) < 1e-15 * dt
    V1 = s.step(U, dt)
    assert field_errors(V1, d["rnd_U1"]).max() < 1e-12
    V3 = s.step(s.step(V1, dt), dt)
    assert field_errors(V3, d["rnd_U3"]).max() < 1e-12
    sa = mn.SolverMHD3DOracle(alf, e, cfl=0.3)
    A = d["alf_U0"]
    assert field_errors(mn.rhs_mhd3d(alf, e, A), d["alf_rhs"]).max() < 1e-12
    dt = float(d["alf_dt"])
    A1 = sa.step(A, dt)
    assert field_errors(A1, d["alf_U1"]).max() < 1e-12
    for _ in range(4):
        A1 = sa.step(A1, dt)
    assert field_errors(A1, d["alf_U5"]).max() < 1e-12


def test_oracle_mhd_closed_forms():
    """The circularly polarised Alfven wave is an exact nonlinear solution;
    projection gives div B = 0, is idempotent and the identity on solenoidal
    fields; the fluxes reduce to the Euler fluxes at B = 0."""
    from oracle import mhd_np as mn, spectral_np as sn
    alf, rnd = _grids()
    e = mn.MHD3DOracle(5.0 / 3.0)
    X, _, _ = alf.xyz_mesh()
    U = mn.alfven_cp_3d(X)
    s = mn.SolverMHD3DOracle(alf, e, cfl=0.3)
    dt = s.compute_dt(U)
    t = 0.0
    for _ in range(10):
        U = s.step(U, dt)
        t += dt
    ex = mn.alfven_cp_3d(X, t=t)
    assert np.abs(U - ex).max() < 5e-8  # RK4 error of the wave's phase at this dt
    rng = np.random.default_rng(3)
    B = rng.standard_normal((3, 16, 16, 16))
    P = np.array(mn.project_solenoidal(rnd, *B))
    assert np.abs(mn.divergence(rnd, *P)).max() < 1e-12
    P2 = np.array(mn.project_solenoidal(rnd, *P))
    np.testing.assert_allclose(P2, P, atol=1e-13)
    Xr, Yr, Zr = rnd.xyz_mesh()
    sol = np.array([np.sin(Yr) * np.cos(Zr), np.sin(Zr) * np.cos(Xr), np.sin(Xr) * np.cos(Yr)])
    np.testing.assert_allclose(np.array(mn.project_solenoidal(rnd, *sol)), sol, atol=1e-14)
    H = np.concatenate([1.0 + 0.1 * rng.random((1, 4, 4, 4)), 0.1 * rng.standard_normal((3, 4, 4, 4)),
                        2.0 + rng.random((1, 4, 4, 4))])
    Fe = sn.Euler3DOracle(5.0 / 3.0).flux(H)
    Fm = e.flux(np.concatenate([H, np.zeros((3, 4, 4, 4))]))
    for a, b in zip(Fm, Fe):
        np.testing.assert_allclose(a[:5], b, rtol=1e-15, atol=1e-15)
        assert np.all(a[5:] == 0)


def test_rhs_of_b_is_solenoidal():
    """The premise of the fused MHD step: the ideal-MHD RHS of B has zero
    spectral divergence (the induction flux is antisymmetric and every
    spectral factor is common to the components), even for a state whose own
    B is not solenoidal.  Checked on the reference's own RHS golden and on the
    oracle, against the divergence of B itself (which is far from zero); and a
    stage state U + c rhs of a projected U is left unchanged by the projection
    up to round-off."""
    from oracle import mhd_np as mn
    d = golden("mhd3d.npz")
    alf, rnd = _grids()
    e = mn.MHD3DOracle(5.0 / 3.0)
    U = d["rnd_U"]
    div_b = np.abs(mn.divergence(rnd, U[5], U[6], U[7])).max()
    assert div_b > 1e-3
    R = d["rnd_rhs"]
    scale = np.abs(R[5:]).max() * (16 // 3)  # |rhs_B| times the largest kept |k|
    div_ref = np.abs(mn.divergence(rnd, R[5], R[6], R[7])).max()
    div_orc = np.abs(mn.divergence(rnd, *mn.rhs_mhd3d(rnd, e, U)[5:])).max()
    assert div_ref < 1e-13 * scale and div_orc < 1e-13 * scale, (div_ref, div_orc, scale)
    s = mn.SolverMHD3DOracle(rnd, e, cfl=0.3)
    V = s.project_B(U)
    S = V + 0.01 * mn.rhs_mhd3d(rnd, e, V)
    P = np.array(mn.project_solenoidal(rnd, S[5], S[6], S[7]))
    assert field_errors(P, S[5:]).max() < 1e-14
