"""3D hip backend vs the reference's golden vectors and the oracle.

Tolerance: FP64 results must match to relative L2 <= 1e-11 per conserved
field (absolute, scaled by the state norm, for a field whose norm is ~0);
single operators to 1e-13.  FP32 bulk: <= 1e-5 against FP64."""
import numpy as np
import pytest

from conftest import golden, rel_l2, field_errors, scaled_err, group_errors

pytestmark = pytest.mark.gpu
FILT = {"enabled": True, "p": 16, "alpha": 36.0}


def _setup(nx, ny, nz, precision="double", fp=FILT, L=(2 * np.pi,) * 3):
    from phrike_amd.grid import Grid3D
    from phrike_amd.equations import EulerEquations3D
    from phrike_amd.solver import SpectralSolver3D
    g = Grid3D(nx, ny, nz, *L, dealias=True, filter_params=dict(fp), backend="hip",
               precision=precision)
    e = EulerEquations3D(1.4)
    return g, e, SpectralSolver3D(g, e, scheme="rk4", cfl=0.4)


def _t(a, dev, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device=dev, dtype=dtype or torch.float64)


@pytest.mark.parametrize("tag,shape", [("cube16", (16, 16, 16)), ("box16x8x32", (16, 8, 32))])
def test_operators_vs_reference(dev, tag, shape):
    d = golden("rhs3d.npz")
    g, e, s = _setup(*shape)
    U = _t(d[f"{tag}_U"], dev)
    print(f"case {tag}", flush=True)
    _v = scaled_err(g.dx1(U[0]).cpu(), d[f"{tag}_dx_rho"], np.linalg.norm(d[f"{tag}_U"][0]))
    print("check", 'rel_l2(g.dx1(U[0]).cpu(), d[f"{tag}_dx_rho"])'[:60], float(_v), flush=True)
    assert _v < 1e-12
    _v = scaled_err(g.dy1(U[0]).cpu(), d[f"{tag}_dy_rho"], np.linalg.norm(d[f"{tag}_U"][0]))
    print("check", 'rel_l2(g.dy1(U[0]).cpu(), d[f"{tag}_dy_rho"])'[:60], float(_v), flush=True)
    assert _v < 1e-12
    _v = scaled_err(g.dz1(U[0]).cpu(), d[f"{tag}_dz_rho"], np.linalg.norm(d[f"{tag}_U"][0]))
    print("check", 'rel_l2(g.dz1(U[0]).cpu(), d[f"{tag}_dz_rho"])'[:60], float(_v), flush=True)
    assert _v < 1e-12
    _v = rel_l2(g.apply_spectral_filter(U).cpu(), d[f"{tag}_filter"])
    print("check", 'rel_l2(g.apply_spectral_filter(U).cpu(), d[f"{tag}_filter"])'[:60], float(_v), flush=True)
    assert _v < 1e-12
    _v = rel_l2(s.rhs(U).cpu(), d[f"{tag}_rhs"])
    print("check", 'rel_l2(s.rhs(U).cpu(), d[f"{tag}_rhs"])'[:60], float(_v), flush=True)
    assert _v < 1e-11
    assert abs(s.max_wave_speed(U) - float(d[f"{tag}_maxc"])) <= 1e-14 * float(d[f"{tag}_maxc"])
    dt = s.compute_dt(U)
    assert abs(dt / float(d[f"{tag}_dt"]) - 1) < 1e-14
    _v = field_errors(s.step(U, float(d[f"{tag}_dt"])).cpu(), d[f"{tag}_step"]).max()
    print("check", 'field_errors(s.step(U, float(d[f"{tag}_dt"])).cpu(), d[f"{tag}_step"]).max()'[:60], float(_v), flush=True)
    assert _v < 1e-12


def test_filter_disabled_step(dev):
    d = golden("rhs3d.npz")
    g, e, s = _setup(16, 16, 16, fp={"enabled": False})
    U = _t(d["cube16_U"], dev)
    _v = field_errors(s.step(U, 1e-3).cpu(), d["nofilter_step"]).max()
    print("check", 'field_errors(s.step(U, 1e-3).cpu(), d["nofilter_step"]).max()'[:60], float(_v), flush=True)
    assert _v < 1e-12


@pytest.mark.parametrize("n", [16, 32])
def test_tgv_steps_vs_reference(dev, n):
    d = golden(f"tgv3d_{n}.npz")
    g, e, s = _setup(n, n, n)
    U = _t(d["U0"], dev)
    dt = float(d["dt"])
    assert abs(s.compute_dt(U) / dt - 1) < 1e-14
    for k in range(1, 11):
        U = s.step(U, dt)
        if k in (1, 10):
            _v = field_errors(U.cpu(), d[f"U_{k}"]).max()
            print("check", 'field_errors(U.cpu(), d[f"U_{k}"]).max()'[:60], float(_v), flush=True)
            assert _v < 1e-11, k


def test_tgv64_200_steps_vs_reference(dev):
    """200 RK4 steps at 64^3 against the reference's own run (stride-4
    subsample and conserved sums)."""
    d = golden("tgv3d_64_200.npz")
    from phrike_amd.initial_conditions import taylor_green_3d
    g, e, s = _setup(64, 64, 64)
    U = _t(taylor_green_3d(64, 64, 64), dev)
    dt = float(d["dt"])
    assert abs(s.compute_dt(U) / dt - 1) < 1e-14
    import torch
    out = torch.empty_like(U)
    for _ in range(200):
        s.step_into(U, out, dt)
        U, out = out, U
    errs = field_errors(U[:, ::4, ::4, ::4].cpu(), d["sub4"])
    print("64^3 x 200 per-field rel L2:", errs, flush=True)
    assert errs.max() <= 1e-11
    sums = s.conserved_sums(U)
    np.testing.assert_allclose(sums[[0, 4]], d["conserved"][[0, 4]], rtol=1e-12)


def test_tgv256_vs_reference(dev):
    """256^3 against the reference's own 256^3 run: stride-8 subsample and
    the Fourier modes |m| <= 8 after 1, 3 and (when recorded) 200 steps."""
    d = golden("tgv3d_256_200.npz")
    from phrike_amd.initial_conditions import taylor_green_3d
    import torch
    g, e, s = _setup(256, 256, 256)
    U = _t(taylor_green_3d(256, 256, 256), dev)
    dt = float(d["dt"])
    assert abs(s.compute_dt(U) / dt - 1) < 1e-14
    out = torch.empty_like(U)
    checks = [k for k in (1, 3, 200) if f"sub8_{k}" in d.files]
    for k in range(1, max(checks) + 1):
        s.step_into(U, out, dt)
        U, out = out, U
        if k in checks:
            sub = U[:, ::8, ::8, ::8].cpu().numpy()
            errs = field_errors(sub, d[f"sub8_{k}"])
            low = np.fft.rfftn(U.cpu().numpy(), axes=(-3, -2, -1))
            idx = np.r_[0:9, -8:0]
            low = low[:, idx][:, :, idx][:, :, :, :9]
            lerr = field_errors(low.view(np.float64), d[f"low8_{k}"].view(np.float64))
            sums = s.conserved_sums(U)
            cerr = np.abs(sums[[0, 4]] / d[f"conserved_{k}"][[0, 4]] - 1).max()
            print(f"256^3 step {k}: sub8 {errs.max():.2e} lowmodes {lerr.max():.2e} "
                  f"mass/energy {cerr:.1e}", flush=True)
            assert errs.max() <= 1e-11 and lerr.max() <= 1e-11 and cerr <= 1e-12


def test_nonuniform_box_vs_oracle(dev):
    from oracle import spectral_np as sn
    from phrike_amd.initial_conditions import smooth_random_state_3d
    nx, ny, nz = 64, 32, 128
    U0 = smooth_random_state_3d(nx, ny, nz, seed=3, kmax=6)
    og = sn.Grid3DOracle(nx, ny, nz, 2 * np.pi, 2 * np.pi, 2 * np.pi, filter_params=dict(FILT))
    os_ = sn.Solver3DOracle(og, sn.Euler3DOracle(1.4), cfl=0.4)
    dt = os_.compute_dt(U0)
    Uo = os_.step(os_.step(U0, dt), dt)
    g, e, s = _setup(nx, ny, nz)
    U = s.step(s.step(_t(U0, dev), dt), dt)
    _v = field_errors(U.cpu(), Uo).max()
    print("check", 'field_errors(U.cpu(), Uo).max()'[:60], float(_v), flush=True)
    assert _v < 1e-12


@pytest.mark.parametrize("nx,ny,nz", [(1024, 8, 16), (512, 16, 8), (16, 1024, 8), (8, 16, 1024)])
def test_long_lines_3d_vs_oracle(dev, nx, ny, nz):
    """3D lines of 1024 points along x, y or z (and 512 along x: the widest x
    kernels and the longest column transforms) against the oracle, two steps
    in FP64; FP32 bounded against the same oracle."""
    from oracle import spectral_np as sn
    from phrike_amd.initial_conditions import smooth_random_state_3d
    import torch
    U0 = smooth_random_state_3d(nx, ny, nz, seed=5, kmax=6)
    og = sn.Grid3DOracle(nx, ny, nz, 2 * np.pi, 2 * np.pi, 2 * np.pi, filter_params=dict(FILT))
    os_ = sn.Solver3DOracle(og, sn.Euler3DOracle(1.4), cfl=0.4)
    dt = os_.compute_dt(U0)
    Uo = os_.step(os_.step(U0, dt), dt)
    for prec in ("double", "single"):
        print(f"case {nx}x{ny}x{nz} {prec}", flush=True)
        g, e, s = _setup(nx, ny, nz, precision=prec)
        U = _t(U0, dev, torch.float64 if prec == "double" else torch.float32)
        U = s.step(s.step(U, dt), dt)
        if prec == "double":
            _v = field_errors(U.cpu(), Uo).max()
            assert _v < 1e-12, _v
        else:
            _v = group_errors(U.double().cpu(), Uo, [(0,), (1, 2, 3), (4,)]).max()
            assert _v < 1e-5, _v
        print(f"  {_v:.2e}", flush=True)


def test_conserved_sums_vs_oracle(dev):
    from oracle import spectral_np as sn
    from phrike_amd.initial_conditions import smooth_random_state_3d
    U0 = smooth_random_state_3d(32, 16, 64, seed=4)
    g, e, s = _setup(32, 16, 64)
    got = s.conserved_sums(_t(U0, dev))
    ref = sn.Euler3DOracle(1.4).conserved_quantities(U0)
    want = [ref["mass"], ref["momentum_x"], ref["momentum_y"], ref["momentum_z"], ref["energy"]]
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * abs(want[0]))


def test_fp32_bulk_vs_fp64(dev):
    import torch
    d = golden("tgv3d_32.npz")
    g, e, s = _setup(32, 32, 32, precision="single")
    U = _t(d["U0"], dev, torch.float32)
    dt = float(d["dt"])
    for _ in range(10):
        U = s.step(U, dt)
    errs = group_errors(U.double().cpu(), d["U_10"], [(0,), (1, 2, 3), (4,)])
    print("fp32 32^3 x 10 per-field rel L2:", field_errors(U.double().cpu(), d["U_10"]),
          "group-normalised:", errs, flush=True)
    assert errs.max() < 1e-5


def test_step_rejects_aliasing_and_bad_shapes(dev):
    """The host layer checks out (shape, dtype, contiguity, no shared memory)
    before any launch; the C ABI itself rejects outputs that overlap U, fully
    or in part."""
    import torch
    from phrike_amd import _lib
    from phrike_amd._lib import HipBackendError
    g, e, s = _setup(16, 16, 16)
    U = torch.ones((5, 16, 16, 16), dtype=torch.float64, device=dev)
    with pytest.raises(ValueError):
        s.step_into(U, U, 1e-3)
    with pytest.raises(ValueError):
        s.step_into(U, torch.empty((5, 16, 16, 16), dtype=torch.float32, device=dev), 1e-3)
    with pytest.raises(ValueError):
        s.step_into(U, torch.empty((5, 16, 16, 32), dtype=torch.float64, device=dev)[..., ::2], 1e-3)
    with pytest.raises(ValueError):
        s.step(torch.ones((4, 16, 16, 16), dtype=torch.float64, device=dev), 1e-3)
    big = torch.ones((6, 16, 16, 16), dtype=torch.float64, device=dev)
    ws = _lib.ptr(g.workspace())
    st = _lib.stream_handle(torch, dev)
    for a, b in ((big[:5], big[:5]), (big[:5], big[1:])):  # full and partial overlap
        with pytest.raises(HipBackendError, match="overlap"):
            _lib.call("ph_step3d_ex", s._plan, _lib.ptr(a), _lib.ptr(b), 1e-3, None, 0, ws, st)



def test_fp32_chained_steps_vs_fp64(dev):
    """Decaying turbulence 64^3 (rms Mach 0.3), 60 steps, FP32 against FP64:
    the solver chains its steps (the state's spectrum carried from step to
    step in the spectral RK4 sum, so no FP32 forward transform of the whole
    state) and stays within 1e-5; unchained steps (ph_step3d every step, which
    transforms its input forward) let that FP32 round trip accumulate."""
    import torch
    from phrike_amd import _lib
    from phrike_amd.initial_conditions import decaying_turbulence_3d
    U0 = decaying_turbulence_3d(64, seed=11, density_seed=12)
    g64, e64, s64 = _setup(64, 64, 64)
    U = _t(U0, dev)
    dt = s64.compute_dt(U)
    for _ in range(60):
        U = s64.step(U, dt)
    ref = U.cpu().numpy()
    g, e, s = _setup(64, 64, 64, precision="single")
    V = _t(U0, dev, torch.float32)
    for _ in range(60):
        V = s.step(V, dt)
    err_corr = group_errors(V.double().cpu().numpy(), ref, [(0,), (1, 2, 3), (4,)])
    W = _t(U0, dev, torch.float32)
    out = torch.empty_like(W)
    plan, ws, st = s._ctx()
    for _ in range(60):
        _lib.call("ph_step3d", plan, _lib.ptr(W), _lib.ptr(out), float(dt), None, ws, st)
        W, out = out, W
    err_direct = group_errors(W.double().cpu().numpy(), ref, [(0,), (1, 2, 3), (4,)])
    print("fp32 vs fp64, correction form:", err_corr, "direct form:", err_direct, flush=True)
    assert err_corr.max() < 1e-5
    assert err_corr[1:4].max() < 0.5 * err_direct[1:4].max()


def test_tgv1024_fp32_matches_512(dev):
    """The largest grid, 1024^3 in FP32 (spectral buffers of more than 2^31
    elements): two TGV steps at the dt of 512^3 keep mass to round-off, and
    the Fourier modes |m| <= 8 agree with the 512^3 run's (the TGV is resolved
    at both sizes and the filter is ~1 at these modes)."""
    import torch
    from phrike_amd.initial_conditions import taylor_green_3d_torch
    low = {}
    dt = None
    for n in (512, 1024):
        print(f"case TGV {n}^3 fp32", flush=True)
        g, e, s = _setup(n, n, n, precision="single")
        U = taylor_green_3d_torch(n, dev, torch.float32)
        if dt is None:
            dt = s.compute_dt(U)
        m0 = float(s.conserved_sums(U)[0])
        out = torch.empty_like(U)
        for _ in range(2):
            s.step_into(U, out, dt)
            U, out = out, U
        del out
        m1 = float(s.conserved_sums(U)[0])
        assert abs(m1 / m0 - 1) < 1e-6, (n, m0, m1)
        F = torch.fft.rfftn(U, dim=(-3, -2, -1)) / float(n) ** 3
        idx = torch.tensor(list(range(9)) + list(range(-8, 0)), device=dev) % n
        low[n] = F[:, idx][:, :, idx][:, :, :, :9].to(torch.complex128).cpu().numpy()
        del U, F, g, e, s
        torch.cuda.empty_cache()
    d = np.abs(low[1024] - low[512]).max() / np.abs(low[512]).max()
    print(f"1024^3 vs 512^3 low modes after 2 steps: {d:.2e}", flush=True)
    assert d < 1e-5


def test_step_without_dealiasing_vs_oracle(dev):
    """A plan without the 2/3 rule (RK4 accumulator in the x kernels, the
    post-step filter as its own passes): 3 steps of a 32 x 16 x 16 random
    state against the oracle, FP64 to 1e-11 per field; FP32 (the solver's
    chained steps, filter in correction form) within 1e-5 of each field
    group's scale of FP64."""
    import torch
    from oracle import spectral_np as sn
    from phrike_amd.equations import EulerEquations3D
    from phrike_amd.grid import Grid3D
    from phrike_amd.initial_conditions import smooth_random_state_3d
    from phrike_amd.solver import SpectralSolver3D
    nx, ny, nz = 32, 16, 16
    fp = {"enabled": True, "p": 8, "alpha": 36.0}
    U0 = smooth_random_state_3d(nx, ny, nz, seed=23, kmax=4)
    og = sn.Grid3DOracle(nx, ny, nz, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=False, filter_params=fp)
    osv = sn.Solver3DOracle(og, sn.Euler3DOracle(1.4), cfl=0.4)
    dt = 0.5 * osv.compute_dt(U0)
    ref = U0
    for _ in range(3):
        ref = osv.step(ref, dt)
    out = {}
    for prec, dt_ in (("double", torch.float64), ("single", torch.float32)):
        print("case no dealiasing", prec, flush=True)
        g = Grid3D(nx, ny, nz, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=False, filter_params=dict(fp),
                   backend="hip", precision=prec)
        s = SpectralSolver3D(g, EulerEquations3D(1.4), scheme="rk4", cfl=0.4)
        V = torch.as_tensor(U0, device=dev, dtype=dt_)
        for _ in range(3):
            V = s.step(V, dt)
        out[prec] = V.double().cpu().numpy()
    err = field_errors(out["double"], ref)
    err32 = group_errors(out["single"], ref, [(0,), (1, 2, 3), (4,)])
    print("no dealiasing: FP64 per-field rel L2", err, "FP32 groups", err32, flush=True)
    assert err.max() < 1e-11
    assert err32.max() < 1e-5


def test_chained_steps_agree_with_plain(dev):
    """Steps chained through the C ABI (PH_STEP_CHAIN: the last kernel also
    forms the next step's fluxes and leaves the filtered state's spectrum;
    PH_STEP_CHAINED: the next step uses both) agree with plain steps, which
    transform their input state, to rounding (1e-13 of each field's scale in
    FP64, 2e-6 in FP32) on a 64 x 32 x 16 box, with a derivative call between
    two of the steps (the step after it computes its own fluxes and
    spectrum); and the solver's own steps take the chained path and agree
    with the oracle."""
    import torch
    from oracle import spectral_np as sn
    from phrike_amd import _lib
    from phrike_amd.initial_conditions import smooth_random_state_3d
    CH, CD, BL = _lib.PH_STEP_CHAIN, _lib.PH_STEP_CHAINED, _lib.PH_STEP_BANDLIMITED
    nx, ny, nz = 64, 32, 16
    U0 = smooth_random_state_3d(nx, ny, nz, seed=11, kmax=4)
    for prec, dt_ in (("double", torch.float64), ("single", torch.float32)):
        print("case chained", prec, flush=True)
        g, e, s = _setup(nx, ny, nz, precision=prec)
        ws = _lib.ptr(g.workspace())
        st = _lib.stream_handle(torch, dev)

        def run(flags_of):
            V = torch.as_tensor(U0, device=dev, dtype=dt_)
            for k in range(4):
                W = torch.empty_like(V)
                _lib.call("ph_step3d_ex", s._plan, _lib.ptr(V), _lib.ptr(W), 2e-3, None, flags_of(k), ws, st)
                if k == 1:
                    g.dx1(W[0])  # writes the spectral buffers between steps 2 and 3
                V = W
            return V.cpu().numpy()

        plain = run(lambda k: BL if k else 0)
        chained = run(lambda k: CH | ((BL | CD) if k else 0))
        tol = 1e-13 if prec == "double" else 2e-6
        for f in range(5):
            scale = np.abs(plain[f]).max()
            assert np.abs(plain[f].astype(np.float64) - chained[f]).max() <= tol * scale, (prec, f)
    og = sn.Grid3DOracle(nx, ny, nz, 2 * np.pi, 2 * np.pi, 2 * np.pi,
                         filter_params={"enabled": True, "p": 16, "alpha": 36.0})
    osv = sn.Solver3DOracle(og, sn.Euler3DOracle(1.4), cfl=0.4)
    ref = U0
    for _ in range(3):
        ref = osv.step(ref, 2e-3)
    g, e, s = _setup(nx, ny, nz)
    V = torch.as_tensor(U0, device=dev)
    for _ in range(3):
        V = s.step(V, 2e-3)
    assert field_errors(V.cpu().numpy(), ref).max() < 1e-11
