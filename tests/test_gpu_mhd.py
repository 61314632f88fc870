"""3D ideal MHD on the GPU (phrike_amd.mhd) against the reference's golden
vectors, the oracle and the exact Alfven wave.  FP64 to relative L2 <= 1e-11
per field after the stated steps; FP32 bulk to 1e-4 of each field group's
scale over 20 steps."""
import numpy as np
import pytest

from conftest import golden, field_errors, group_errors

pytestmark = pytest.mark.gpu


def _setup(nx, ny, nz, L=(2 * np.pi,) * 3, fp=None, precision="double", cfl=0.3):
    from phrike_amd.grid import Grid3D
    from phrike_amd.mhd import MHDEquations3D, SpectralSolverMHD3D
    g = Grid3D(nx, ny, nz, *L, dealias=True,
               filter_params=dict(fp or {"enabled": True, "p": 16, "alpha": 36.0}), backend="hip",
               precision=precision)
    g.mhd_config = {"resistivity": 0.0, "viscosity": 0.0, "divergence_clean": "projection",
                    "project_each_substage": True}
    e = MHDEquations3D(5.0 / 3.0)
    return g, e, SpectralSolverMHD3D(g, e, cfl=cfl)


def _t(a, dev, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device=dev, dtype=dtype or torch.float64)


def test_mhd_vs_reference_goldens(dev):
    d = golden("mhd3d.npz")
    g, e, s = _setup(16, 16, 16)
    U = _t(d["rnd_U"], dev)
    print("case rnd16", flush=True)
    assert field_errors(s._compute_rhs(U).cpu().numpy(), d["rnd_rhs"]).max() < 1e-12
    V = s._project_B(U).cpu().numpy()
    assert field_errors(V[5:], d["rnd_proj"]).max() < 1e-13
    assert abs(s.max_wave_speed(U) - float(d["rnd_maxc"])) < 1e-14 * float(d["rnd_maxc"])
    dt = s.compute_dt(U)
    assert abs(dt - float(d["rnd_dt"])) < 1e-14 * dt
    cq = s.conserved_quantities(U)
    for k, v in zip([str(x) for x in d["rnd_cons_keys"]], d["rnd_cons"]):
        assert abs(cq[k] - v) <= 1e-12 * max(abs(v), 1.0), k
    W = U
    for k in range(3):
        W = s.step(W, float(d["rnd_dt"]))
        if k == 0:
            err1 = field_errors(W.cpu().numpy(), d["rnd_U1"])
    err3 = field_errors(W.cpu().numpy(), d["rnd_U3"])
    print("rnd16 step 1 / 3 per-field rel L2:", err1, err3, flush=True)
    assert err1.max() < 1e-11 and err3.max() < 1e-11
    print("case alfven 32x8x8", flush=True)
    g, e, s = _setup(32, 8, 8, L=(1.0, 1.0, 1.0), fp={"enabled": True, "p": 8, "alpha": 36.0})
    A = _t(d["alf_U0"], dev)
    assert field_errors(s._compute_rhs(A).cpu().numpy(), d["alf_rhs"]).max() < 1e-12
    dt = float(d["alf_dt"])
    for k in range(5):
        A = s.step(A, dt)
        if k == 0:
            assert field_errors(A.cpu().numpy(), d["alf_U1"]).max() < 1e-11
    e5 = field_errors(A.cpu().numpy(), d["alf_U5"])
    print("alfven 5 steps per-field rel L2:", e5, flush=True)
    assert e5.max() < 1e-11


def test_mhd_vs_oracle_64(dev):
    """A band-limited random state with a non-solenoidal B on 64^3 (the
    first projection removes the longitudinal part), 3 steps, vs the oracle."""
    from oracle import mhd_np as mn, spectral_np as sn
    from phrike_amd.initial_conditions import smooth_random_state_3d
    n = 64
    H = smooth_random_state_3d(n, n, n, seed=31, kmax=6)
    Bn = smooth_random_state_3d(n, n, n, seed=32, kmax=5)
    B = np.stack([0.3 * Bn[1] / Bn[0], 0.3 * Bn[2] / Bn[0], 1.0 + 0.3 * Bn[3] / Bn[0]])
    U0 = np.concatenate([H, B])
    U0[4] = H[4] + 0.5 * (B ** 2).sum(axis=0)
    og = sn.Grid3DOracle(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, filter_params={"enabled": True, "p": 16,
                                                                                  "alpha": 36.0})
    os_ = mn.SolverMHD3DOracle(og, mn.MHD3DOracle(5.0 / 3.0), cfl=0.3)
    dt = os_.compute_dt(U0)
    ref = U0
    for _ in range(3):
        ref = os_.step(ref, dt)
    g, e, s = _setup(n, n, n)
    U = _t(U0, dev)
    for _ in range(3):
        U = s.step(U, dt)
    err = field_errors(U.cpu().numpy(), ref)
    print("64^3 x 3 steps per-field rel L2:", err, "max|div B|", s.max_div_b(U), flush=True)
    assert err.max() < 1e-11
    assert s.max_div_b(U) < 1e-10


def test_mhd_alfven_exact_and_run(dev):
    """run() on the exact circularly polarised Alfven wave (64 x 8 x 8 on the
    unit cube, t_end 0.2): the state stays within the RK4 phase error of the
    exact solution, div B stays at round-off, the history has the reference's
    keys and conserves mass and energy."""
    from oracle.mhd_np import alfven_cp_3d
    g, e, s = _setup(64, 8, 8, L=(1.0, 1.0, 1.0), fp={"enabled": True, "p": 8, "alpha": 36.0})
    X = np.broadcast_to(np.linspace(0, 1, 64, endpoint=False)[None, None, :], (8, 8, 64))
    U0 = _t(alfven_cp_3d(np.ascontiguousarray(X)), dev)
    hist = s.run(U0, 0.0, 0.2, output_interval=0.1)
    assert list(hist) == ["time", "mass", "momentum_x", "momentum_y", "momentum_z", "energy",
                          "magnetic_energy", "cross_helicity", "div_b_max"]
    assert abs(hist["time"][-1] - 0.2) < 1e-12
    ex = alfven_cp_3d(np.ascontiguousarray(X), t=0.2)
    err = np.abs(s.U.cpu().numpy() - ex).max()
    print("alfven t=0.2 max abs error vs exact:", err, "div_b_max", hist["div_b_max"], flush=True)
    assert err < 1e-6
    assert max(hist["div_b_max"]) < 1e-10
    assert abs(hist["mass"][-1] - hist["mass"][0]) < 1e-12 * abs(hist["mass"][0])
    assert abs(hist["energy"][-1] - hist["energy"][0]) < 1e-8 * abs(hist["energy"][0])


def test_mhd_fp32_vs_fp64(dev):
    import torch
    d = golden("mhd3d.npz")
    g, e, s = _setup(16, 16, 16)
    g32, e32, s32 = _setup(16, 16, 16, precision="single")
    U = _t(d["rnd_U"], dev)
    V = U.float()
    dt = float(d["rnd_dt"])
    for _ in range(20):
        U = s.step(U, dt)
        V = s32.step(V, dt)
    err = group_errors(V.double().cpu().numpy(), U.cpu().numpy(), [(0,), (1, 2, 3), (4,), (5, 6, 7)])
    print("fp32 vs fp64 MHD 16^3 x 20:", err, flush=True)
    assert np.isfinite(V.cpu().numpy()).all()
    assert err.max() < 1e-4


def test_mhd_fused_step_vs_substage_projection(dev):
    """ph_step_mhd3d_ex with PH_STEP_SOLENOIDAL (stage boundaries fused, one
    projection per step) against ph_step_mhd3d (projection after every
    substage) on a projected band-limited random state, 64^3 x 5 steps
    through the C ABI: FP64 equal to 1e-12 per field, max |div B| at
    round-off; the FP32 fused run within 1e-4 of each field group's scale."""
    import torch
    from phrike_amd import _lib
    from phrike_amd.initial_conditions import smooth_random_state_3d
    n = 64
    H = smooth_random_state_3d(n, n, n, seed=41, kmax=6)
    Bn = smooth_random_state_3d(n, n, n, seed=42, kmax=5)
    B = np.stack([0.3 * Bn[1] / Bn[0], 0.3 * Bn[2] / Bn[0], 1.0 + 0.3 * Bn[3] / Bn[0]])
    U0 = np.concatenate([H, B])
    U0[4] = H[4] + 0.5 * (B ** 2).sum(axis=0)
    g, e, s = _setup(n, n, n)
    V0 = s._project_B(_t(U0, dev))
    dt = s.compute_dt(V0)
    ws = _lib.ptr(s._ws)

    def run(step_fn, V, solver=s, wsp=ws):
        for _ in range(5):
            out = torch.empty_like(V)
            step_fn(V, out, solver, wsp)
            V = out
        return V

    def unfused(V, out, sv, wsp):
        _lib.call("ph_step_mhd3d", sv._plan, _lib.ptr(V), _lib.ptr(out), dt, wsp, sv._st())

    def fused(V, out, sv, wsp):
        _lib.call("ph_step_mhd3d_ex", sv._plan, _lib.ptr(V), _lib.ptr(out), dt, _lib.PH_STEP_SOLENOIDAL, wsp,
                  sv._st())

    print("case fused vs unfused 64^3 FP64", flush=True)
    ref = run(unfused, V0)
    got = run(fused, V0)
    err = field_errors(got.cpu().numpy(), ref.cpu().numpy())
    print("fused vs unfused per-field rel L2:", err, "max|div B|", s.max_div_b(got), flush=True)
    assert err.max() < 1e-12
    assert s.max_div_b(got) < 1e-10
    print("case fused FP32 64^3", flush=True)
    g32, e32, s32 = _setup(n, n, n, precision="single")
    V32 = run(fused, V0.float(), s32, _lib.ptr(s32._ws))
    err32 = group_errors(V32.double().cpu().numpy(), ref.cpu().numpy(), [(0,), (1, 2, 3), (4,), (5, 6, 7)])
    print("fused FP32 vs FP64 per-field:", err32, flush=True)
    assert np.isfinite(V32.cpu().numpy()).all() and err32.max() < 1e-4
