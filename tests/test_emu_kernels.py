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


def test_fp32_filter_correction_form(emu):
    """FP32 step on a band-limited state (a previous step's output) whose kept
    modes carry a flat random spectrum, so the filter acts on every mode:
    with and without PH_STEP_BANDLIMITED (which only plans without the
    spectral RK4 sum use, for the correction-form filter) within FP32
    rounding of FP64; FP64 plans ignore the flag; a device dt (dt_dev[0])
    equals the host dt."""
    from emu_lib import Emu3D
    d = golden("tgv3d_16.npz")
    dt = float(d["dt"])
    rng = np.random.default_rng(3)
    U0 = d["U0"] + 0.05 * rng.standard_normal(d["U0"].shape) * np.array([1, 1, 1, 1, 10.0])[:, None, None, None]
    e64 = Emu3D(emu, 16, 16, 16)
    V1 = e64.step(U0, dt)  # band-limited
    e32 = Emu3D(emu, 16, 16, 16, dtype=np.float32)
    direct = e32.step(V1, dt).astype(np.float64)
    corr = e32.step(V1, dt, flags=1).astype(np.float64)
    ref = e64.step(V1, dt)
    print("corr vs direct", group_errors(corr, direct, [(0,), (1, 2, 3), (4,)]),
          "corr vs fp64", group_errors(corr, ref, [(0,), (1, 2, 3), (4,)]))
    assert group_errors(corr, ref, [(0,), (1, 2, 3), (4,)]).max() < 1e-6
    assert group_errors(direct, ref, [(0,), (1, 2, 3), (4,)]).max() < 1e-6
    np.testing.assert_array_equal(e64.step(V1, dt, flags=1), ref)
    np.testing.assert_array_equal(e64.step(V1, 0.0, dt_dev=np.array([dt])), ref)


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


def test_hybrid_rhs_and_step(emu):
    """Hybrid Fourier-x / LGL-y path (x-line kernel, RK combine, filters) vs the
    oracle; the matrix-core apply is stood in by a CPU loop here."""
    from emu_lib import EmuHybrid
    from oracle.hybrid_np import HybridGridOracle, HybridSolverOracle, rhs_hybrid
    from oracle.spectral_np import Euler2DOracle
    from phrike_amd.initial_conditions import khi_2d_gll

    fp = {"enabled": True, "p": 16, "alpha": 36.0}
    g = HybridGridOracle(32, 16, filter_params=fp)
    eqs = Euler2DOracle(1.4)
    sol = HybridSolverOracle(g, eqs, cfl=0.4)
    U = khi_2d_gll(32, g.y, seed=3)
    e = EmuHybrid(emu, 32, 16, g.D, g.Fn, g.dy_min, batch=2, ybc=1)
    assert field_errors(e.rhs(U), rhs_hybrid(g, eqs, U)).max() < 1e-12
    dt = sol.compute_dt(U)
    assert abs(e.dt(U, 0.4)[0] - dt) <= 1e-14 * dt
    ref = sol.step(U, dt)
    assert field_errors(e.step(U, dt), ref).max() < 1e-12
    pair = np.stack([U, U[:, ::-1, :].copy()])
    both = e.step(pair, dt)
    np.testing.assert_array_equal(both[0], e.step(U, dt))


def test_mhd_kernels_vs_reference(emu):
    """3D ideal MHD on the CPU emulation against the reference's goldens:
    RHS, projection (every mode kept, Nyquist planes zeroed, E at fixed
    thermal pressure), CFL speed, conserved sums and one step, FP64."""
    from emu_lib import EmuMHD3D
    d = golden("mhd3d.npz")
    e = EmuMHD3D(emu, 16, 16, 16)
    U = d["rnd_U"]
    assert field_errors(e.rhs(U), d["rnd_rhs"]).max() < 1e-12
    V = e.project(U)
    assert field_errors(V[5:], d["rnd_proj"]).max() < 1e-13
    np.testing.assert_array_equal(V[:4], U[:4])
    m, dt = e.max_speed_dt(U, 0.3)
    assert abs(m - float(d["rnd_maxc"])) < 1e-14 * m
    assert abs(dt - float(d["rnd_dt"])) < 1e-14 * dt
    keys = [str(k) for k in d["rnd_cons_keys"]]
    order = ["mass", "momentum_x", "momentum_y", "momentum_z", "energy", "magnetic_energy", "cross_helicity"]
    ref = dict(zip(keys, d["rnd_cons"]))
    got = e.sums(U)
    for i, k in enumerate(order):
        assert abs(got[i] - ref[k]) <= 1e-12 * max(abs(ref[k]), 1.0), k
    assert field_errors(e.step(U, float(d["rnd_dt"])), d["rnd_U1"]).max() < 1e-11
    a = EmuMHD3D(emu, 32, 8, 8, L=(1.0, 1.0, 1.0), filt=(1, 8, 36.0))
    assert field_errors(a.rhs(d["alf_U0"]), d["alf_rhs"]).max() < 1e-12
    assert field_errors(a.step(d["alf_U0"], float(d["alf_dt"])), d["alf_U1"]).max() < 1e-11


def test_mhd_fused_step(emu):
    """The fused MHD step (PH_STEP_SOLENOIDAL: stage boundaries in k_xmid, one
    projection per step) on a solenoidal state equals the step with a
    projection after every substage to round-off, and continues the
    reference's own runs: random 16^3 steps 2-3 after an unfused first step
    (which projects its non-solenoidal input), the Alfven wave's 5 steps."""
    from emu_lib import EmuMHD3D
    from phrike_amd import _lib
    SOL = _lib.PH_STEP_SOLENOIDAL
    d = golden("mhd3d.npz")
    e = EmuMHD3D(emu, 16, 16, 16)
    dt = float(d["rnd_dt"])
    V0 = e.project(d["rnd_U"])
    assert field_errors(e.step(V0, dt, SOL), e.step(V0, dt)).max() < 1e-13
    W = e.step(d["rnd_U"], dt)
    for _ in range(2):
        W = e.step(W, dt, SOL)
    assert field_errors(W, d["rnd_U3"]).max() < 1e-11
    a = EmuMHD3D(emu, 32, 8, 8, L=(1.0, 1.0, 1.0), filt=(1, 8, 36.0))
    A = d["alf_U0"]
    for _ in range(5):
        A = a.step(A, float(d["alf_dt"]), SOL)
    assert field_errors(A, d["alf_U5"]).max() < 1e-11
    with pytest.raises(RuntimeError, match="flags"):
        e.step(V0, dt, 4)


def test_plan_kind_checked(emu):
    """A plan handle of another kind is an argument error, not undefined
    behaviour: every entry point checks the plan's dynamic type."""
    from emu_lib import Emu3D, EmuMHD3D, P
    e3 = Emu3D(emu, 8, 8, 8)
    em = EmuMHD3D(emu, 8, 8, 8)
    U5 = np.ones((5, 8, 8, 8))
    U8 = np.ones((8, 8, 8, 8))
    out8 = np.zeros_like(U8)
    rc = emu.ph_step_mhd3d(e3.h, P(U8), P(out8), 1e-3, P(em.ws), None)
    assert rc != 0 and b"not an MHD plan" in emu.ph_last_error()
    out5 = np.zeros_like(U5)
    rc = emu.ph_step3d(em.h, P(U5), P(out5), 1e-3, None, P(e3.ws), None)
    assert rc != 0 and b"not a 3D plan" in emu.ph_last_error()
    # the right kinds still work
    assert np.isfinite(e3.step(U5, 1e-3)).all()


def test_gll_apply_rejects_in_place(emu):
    """ph_gll_apply with out overlapping F is an argument error (each output
    row reads the whole line of F)."""
    import ctypes
    from emu_lib import P
    F = np.ones((2, 16, 16))
    D = np.eye(16)
    emu.ph_gll_apply.restype = ctypes.c_int
    rc = emu.ph_gll_apply(0, 16, 16, 2, P(D), None, None, P(F), None, P(F[1:]), None)
    assert rc != 0 and b"must not overlap" in emu.ph_last_error()
    out = np.zeros_like(F)
    assert emu.ph_gll_apply(0, 16, 16, 2, P(D), None, None, P(F), None, P(out), None) == 0
    np.testing.assert_array_equal(out, F)


def test_step2d_noalloc_equals_workspace_step(emu):
    """ph_step2d_noalloc (static scratch, no workspace argument) gives the same
    bits as ph_step2d, on two calls in a row with the second smaller; a batch
    beyond the static scratch is an argument error."""
    from emu_lib import Emu2D
    from phrike_amd.initial_conditions import khi_2d_periodic
    for (n, nb, prec) in ((32, 3, np.float64), (16, 2, np.float32)):
        e = Emu2D(emu, n, n, batch=nb, dtype=prec)
        U = np.stack([khi_2d_periodic(n, n, seed=s) for s in range(nb)]).astype(prec)
        a = e.step(U, 1e-3)
        b = e.step(U, 1e-3, noalloc=True)
        assert np.array_equal(a, b)
    big = Emu2D(emu, 1024, 1024, batch=2, dtype=np.float64)
    with pytest.raises(RuntimeError, match="static scratch"):
        big.step(np.ones((2, 4, 1024, 1024)), 1e-3, noalloc=True)


def test_hybrid_inactive_run_untouched_with_filter(emu):
    """A masked run (active 0) of a batched hybrid step keeps its output
    buffer bit for bit with the filter on (the nodal-filter apply skips it),
    and the active runs equal their steps alone."""
    from emu_lib import EmuHybrid
    from oracle.hybrid_np import HybridGridOracle
    from phrike_amd.initial_conditions import khi_2d_gll
    FP = {"enabled": True, "p": 16, "alpha": 36.0}
    og = HybridGridOracle(16, 16, filter_params=FP)
    e = EmuHybrid(emu, 16, 16, og.D, og.Fn, og.dy_min, batch=3)
    U = np.stack([khi_2d_gll(16, og.y, seed=s) for s in (1, 2, 3)])
    active = np.array([1, 0, 1], dtype=np.int32)
    sentinel = np.full_like(U, 7.25)
    out = e.step(U, 1e-3, active=active, out=sentinel.copy())
    assert np.array_equal(out[1], sentinel[1])
    for b in (0, 2):
        assert np.array_equal(out[b], e.step(U[b:b + 1], 1e-3)[0])


def test_nan_state_gives_nan_dt_and_retires(emu):
    """A NaN state gives a NaN dt (the reference's cfl*h/max with a NaN max,
    phrike/solver.py:613-620), not 1e-6, and ph_runs_advance retires that run
    with t = NaN while the other run advances."""
    from emu_lib import Emu2D, P
    from phrike_amd.initial_conditions import khi_2d_periodic
    e = Emu2D(emu, 16, 16, batch=2)
    U = np.stack([khi_2d_periodic(16, 16, seed=1), khi_2d_periodic(16, 16, seed=2)])
    U[1, 0, 3, 5] = np.nan
    m, d = e.dt(U, 0.4)
    assert np.isfinite(d[0]) and d[0] > 0 and np.isnan(m[1]) and np.isnan(d[1])
    t = np.zeros(2)
    t_end = np.ones(2)
    active = np.ones(2, dtype=np.int32)
    steps = np.zeros(2, dtype=np.int32)
    assert emu.ph_runs_clip(2, P(t), P(t_end), P(d), P(active), None) == 0
    assert emu.ph_runs_advance(2, P(t), P(t_end), P(d), P(active), P(steps), None) == 0
    assert active[0] == 1 and active[1] == 0 and np.isnan(t[1]) and t[0] == d[0]


def test_device_run_loop_matches_host_loop(emu):
    """ph_run3d (device-resident dt, clip, advance and stop test) against the
    reference loop restated on the oracle (phrike/solver.py:1395-1513): same
    number of steps, same times, same state, with a stop time inside the run
    and the last step clipped to t_end."""
    from emu_lib import Emu3D, P
    from oracle import spectral_np as sn
    from phrike_amd.initial_conditions import taylor_green_3d
    n, cfl = 16, 0.4
    filt = {"enabled": True, "p": 16, "alpha": 36.0}
    U0 = taylor_green_3d(n, n, n)
    og = sn.Grid3DOracle(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, filter_params=filt, fft_workers=1)
    osv = sn.Solver3DOracle(og, sn.Euler3DOracle(1.4), cfl=cfl)
    t_end = 6.5 * osv.compute_dt(U0)
    # host loop of the reference
    U, t, dt, ts = U0, 0.0, osv.compute_dt(U0), []
    while t < t_end - 1e-12:
        dt = min(dt, t_end - t)
        U = osv.step(U, dt)
        t += dt
        ts.append(t)
        dt = osv.compute_dt(U)
    e = Emu3D(emu, n, n, n)
    a = np.ascontiguousarray(U0)
    b = np.zeros_like(a)
    rs = np.zeros(6)
    rf = np.zeros(2, dtype=np.int32)
    rs[2] = t_end
    rs[3] = ts[2] - 1e-9  # an "output" time between steps 2 and 3: stop after step 3
    mx = np.zeros(1)
    assert emu.ph_max_wave_speed3d(e.h, P(a), P(mx), P(rs[1:]), cfl, P(e.ws), None) == 0
    rf[0] = 1
    assert emu.ph_run3d(e.h, P(a), P(b), 5, P(rs), P(rf), cfl, 0, P(e.ws), None) == 0
    # stopped at the stop time, state in Ub (times equal up to the last bit of
    # the CFL speed: the two states differ by FP64 rounding)
    assert rf[1] == 3 and rf[0] == 0 and abs(rs[0] - ts[2]) <= 1e-14 * ts[2]
    rs[3] = np.inf
    rf[0] = 1
    assert emu.ph_run3d(e.h, P(b), P(a), 8, P(rs), P(rf), cfl, 1, P(e.ws), None) == 0
    assert rf[1] == len(ts) == 7 and rf[0] == 0 and abs(rs[0] - t_end) <= 1e-14 * t_end
    state = b if (rf[1] - 3) % 2 == 0 else a  # even: the second call's Ua (b)
    assert field_errors(state, U).max() < 1e-12


def test_step_without_dealiasing_vs_oracle(emu):
    """A plan without the 2/3 rule takes the step with the RK4 accumulator in
    the x kernels and the post-step filter as its own passes: FP64, 2 steps
    on a 16 x 8 x 16 random state against the oracle (1e-12 per field); FP32
    with the correction-form filter (PH_STEP_BANDLIMITED on the second step)
    within 1e-5 of each field group's scale of FP64."""
    from emu_lib import Emu3D
    from oracle import spectral_np as sn
    from phrike_amd.initial_conditions import smooth_random_state_3d
    nx, ny, nz = 16, 8, 16
    U0 = smooth_random_state_3d(nx, ny, nz, seed=21, kmax=3)
    og = sn.Grid3DOracle(nx, ny, nz, 2 * np.pi, 2 * np.pi, 2 * np.pi, dealias=False,
                         filter_params={"enabled": True, "p": 8, "alpha": 36.0}, fft_workers=1)
    osv = sn.Solver3DOracle(og, sn.Euler3DOracle(1.4), cfl=0.4)
    dt = 0.5 * osv.compute_dt(U0)
    ref = osv.step(osv.step(U0, dt), dt)
    e = Emu3D(emu, nx, ny, nz, dealias=False, filt=(1, 8, 36.0))
    got = e.step(e.step(U0, dt), dt)
    assert field_errors(got, ref).max() < 1e-12
    e32 = Emu3D(emu, nx, ny, nz, dtype=np.float32, dealias=False, filt=(1, 8, 36.0))
    g32 = e32.step(e32.step(U0.astype(np.float32), dt), dt, 1).astype(np.float64)
    assert group_errors(g32, ref, [(0,), (1, 2, 3), (4,)]).max() < 1e-5


def test_run3d_chained_calls_equal_one_call(emu):
    """ph_run3d with PH_STEP_CHAIN / PH_STEP_CHAINED: a run in three calls (2
    steps, a call on the stopped run, 3 steps), each starting from the buffer
    that holds the previous call's final state, equals one call of 5 steps
    bit for bit, FP64 and FP32; a call after a workspace write computes its
    own first passes and agrees to rounding."""
    from emu_lib import Emu3D, P
    from phrike_amd import _lib
    from phrike_amd.initial_conditions import smooth_random_state_3d
    CH, CD = _lib.PH_STEP_CHAIN, _lib.PH_STEP_CHAINED
    n, cfl = 16, 0.4
    U0 = smooth_random_state_3d(n, n, n, seed=9, kmax=4)
    for dtype, tol in ((np.float64, 1e-13), (np.float32, 2e-6)):
        e = Emu3D(emu, n, n, n, dtype=dtype)

        def fresh():
            a = np.ascontiguousarray(U0.astype(dtype))
            b = np.zeros_like(a)
            rs = np.zeros(6)
            rf = np.zeros(2, dtype=np.int32)
            rs[2], rs[3] = 1e30, np.inf
            assert emu.ph_max_wave_speed3d(e.h, P(a), None, P(rs[1:]), cfl, P(e.ws), None) == 0
            return a, b, rs, rf

        def call(x, y, k, rs, rf, active, flags):
            rf[0] = active
            assert emu.ph_run3d(e.h, P(x), P(y), k, P(rs), P(rf), cfl, flags, P(e.ws), None) == 0

        a, b, rs, rf = fresh()
        call(a, b, 5, rs, rf, 1, CH)
        one = b.copy()
        a, b, rs2, rf2 = fresh()
        call(a, b, 2, rs2, rf2, 1, CH)      # final state in a (2 steps)
        call(a, b, 1, rs2, rf2, 0, CH | CD)  # stopped run: no step
        call(a, b, 3, rs2, rf2, 1, CH | CD)  # final state in b
        assert rf2[1] == 5 and rs2[0] == rs[0]
        np.testing.assert_array_equal(one, b)
        a, b, rs3, rf3 = fresh()
        call(a, b, 2, rs3, rf3, 1, CH)
        e.deriv(0, a)  # writes the spectral buffers
        call(a, b, 3, rs3, rf3, 1, CH | CD)
        for f in range(5):
            assert np.abs(b[f].astype(np.float64) - one[f]).max() <= tol * np.abs(one[f]).max(), (dtype, f)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-5)])
def test_fft3d_half_spectrum(emu, dtype, tol):
    """ph_fft3d / ph_ifft3d of a plan that keeps every mode: the half spectra
    equal numpy's rfftn (kx >= 0) and the inverse returns the fields."""
    from emu_lib import Emu3D, P
    nx, ny, nz = 16, 8, 8
    e = Emu3D(emu, nx, ny, nz, dtype=dtype, dealias=False, filt=(0, 8, 36.0))
    rng = np.random.default_rng(4)
    f = np.ascontiguousarray(rng.standard_normal((2, nz, ny, nx)).astype(dtype))
    spec = np.zeros((2, nz, ny, nx // 2 + 1), dtype=np.complex128 if dtype == np.float64 else np.complex64)
    assert emu.ph_fft3d(e.h, 2, P(f), P(spec), P(e.ws), None) == 0
    ref = np.fft.rfftn(f.astype(np.float64), axes=(-3, -2, -1))
    assert np.abs(spec - ref).max() <= tol * np.abs(ref).max()
    back = np.zeros_like(f)
    assert emu.ph_ifft3d(e.h, 2, P(spec), P(back), P(e.ws), None) == 0
    assert np.abs(back - f).max() <= tol * np.abs(f).max()
    # a plan with the 2/3 rule refuses
    d = Emu3D(emu, nx, ny, nz, dtype=dtype)
    assert emu.ph_fft3d(d.h, 2, P(f), P(spec), P(d.ws), None) != 0


def test_fused_cfl_equals_standalone_max(emu):
    """The run loops take the next dt from wave-speed maxima the step's last
    kernel leaves per row (no second read of the state); that dt equals the
    standalone ph_max_wave_speed* of the written state bit for bit, in 3D
    (FP64 and FP32) and for batched 2D runs (an inactive run keeps its dt)."""
    from emu_lib import Emu2D, Emu3D, P
    from phrike_amd.initial_conditions import khi_2d_periodic, taylor_green_3d
    cfl = 0.4
    for dt_ in (np.float64, np.float32):
        e = Emu3D(emu, 16, 8, 32, dtype=dt_)
        a = np.ascontiguousarray(taylor_green_3d(16, 8, 32).astype(dt_))
        b = np.zeros_like(a)
        rs = np.zeros(6)
        rf = np.array([1, 0], dtype=np.int32)
        rs[2], rs[3] = 1.0, np.inf
        assert emu.ph_max_wave_speed3d(e.h, P(a), None, P(rs[1:]), cfl, P(e.ws), None) == 0
        assert emu.ph_run3d(e.h, P(a), P(b), 1, P(rs), P(rf), cfl, 0, P(e.ws), None) == 0
        d = np.zeros(1)
        assert emu.ph_max_wave_speed3d(e.h, P(b), None, P(d), cfl, P(e.ws), None) == 0
        assert rf[1] == 1 and rs[1] == d[0] and rs[4] == d[0], (dt_, rs, d)
    nb = 3
    e = Emu2D(emu, 32, 16, batch=nb, dtype=np.float32)
    U = np.ascontiguousarray(np.stack([khi_2d_periodic(32, 16, seed=s) for s in range(nb)]).astype(np.float32))
    V = np.zeros_like(U)
    rs = np.zeros(6 * nb)
    rf = np.zeros(2 * nb, dtype=np.int32)
    rf[:nb] = [1, 0, 1]
    rs[2 * nb:3 * nb] = 1.0
    rs[3 * nb:4 * nb] = np.inf
    _, d0 = e.dt(U, cfl)
    rs[nb:2 * nb] = d0
    rs[nb + 1] = 0.125  # the inactive run's dt stays
    assert emu.ph_run2d(e.h, nb, P(U), P(V), 1, P(rs), P(rf), cfl, P(e.ws), None) == 0
    _, d1 = e.dt(V, cfl)
    assert rs[nb] == d1[0] and rs[nb + 2] == d1[2] and rs[nb + 1] == 0.125


def test_chained_steps(emu):
    """PH_STEP_CHAIN / PH_STEP_CHAINED: steps whose last kernel forms the next
    step's fluxes (k_xend) and whose spectral RK4 sum starts from the previous
    step's filtered spectrum agree with plain steps (which transform their
    input state) to rounding: 1e-13 of each field's scale in FP64, 2e-6 in
    FP32; a call that writes the workspace between two chained steps makes
    the second compute its fluxes and spectrum itself (then bit for bit the
    plain steps), and an unknown flag is an argument error."""
    from emu_lib import Emu3D
    from phrike_amd import _lib
    from phrike_amd.initial_conditions import smooth_random_state_3d
    CH, CD, BL = _lib.PH_STEP_CHAIN, _lib.PH_STEP_CHAINED, _lib.PH_STEP_BANDLIMITED
    U0 = smooth_random_state_3d(16, 16, 16, seed=5, kmax=4)
    for dtype in (np.float64, np.float32):
        e = Emu3D(emu, 16, 16, 16, dtype=dtype)
        dt = 2e-3
        A = U0.astype(dtype)
        for k in range(3):
            A = e.step(A, dt, BL if k else 0)
        B = U0.astype(dtype)
        for k in range(3):
            B = e.step(B, dt, CH | ((BL | CD) if k else 0))
        tol = 1e-13 if dtype == np.float64 else 2e-6
        for f in range(5):
            scale = np.abs(A[f]).max()
            assert np.abs(A[f].astype(np.float64) - B[f]).max() <= tol * scale, (dtype, f)
        C = e.step(U0.astype(dtype), dt, CH)
        e.deriv(0, C)  # writes the spectral buffers: the held fluxes are gone
        C = e.step(C, dt, CH | CD | BL)
        D = e.step(e.step(U0.astype(dtype), dt), dt, BL)
        np.testing.assert_array_equal(C, D)
    with pytest.raises(RuntimeError, match="flags"):
        e.step(U0.astype(np.float32), dt, 16)


def test_run2d_chained_steps(emu):
    """ph_run2d chains its steps (each but the last forms the next step's
    fluxes in k_xend): one call of 3 steps equals three calls of one step bit
    for bit (states, scalars, flags), with one run inactive throughout and one
    that stops after its second (clipped) step."""
    from emu_lib import Emu2D, P
    from phrike_amd.initial_conditions import khi_2d_periodic
    cfl, nb = 0.4, 3
    e = Emu2D(emu, 32, 16, batch=nb, dtype=np.float32)
    U0 = np.ascontiguousarray(np.stack([khi_2d_periodic(32, 16, seed=s + 3) for s in range(nb)]).astype(np.float32))
    _, d0 = e.dt(U0, cfl)

    def fresh():
        rs = np.zeros(6 * nb)
        rf = np.zeros(2 * nb, dtype=np.int32)
        rf[:nb] = [1, 0, 1]
        rs[nb:2 * nb] = d0
        rs[2 * nb:3 * nb] = [1.0, 1.0, 1.5 * d0[2]]
        rs[3 * nb:4 * nb] = np.inf
        return U0.copy(), np.zeros_like(U0), rs, rf

    U, V, rs, rf = fresh()
    assert emu.ph_run2d(e.h, nb, P(U), P(V), 3, P(rs), P(rf), cfl, P(e.ws), None) == 0
    U1, V1, rs1, rf1 = fresh()
    for k in range(3):
        a, b = (U1, V1) if k % 2 == 0 else (V1, U1)
        assert emu.ph_run2d(e.h, nb, P(a), P(b), 1, P(rs1), P(rf1), cfl, P(e.ws), None) == 0
    np.testing.assert_array_equal(rf, rf1)
    assert list(rf[nb:]) == [3, 0, 2]
    np.testing.assert_array_equal(rs, rs1)
    np.testing.assert_array_equal(U, U1)
    np.testing.assert_array_equal(V, V1)


def test_column_tile_pairing_is_a_permutation(emu):
    """The XCD pairing of column tiles (xcd_pair_tile: blocks i and i + 8 of
    each run of 16 take tiles 2i and 2i + 1) forced at every tile width, with
    2-column tiles so that a row of the grid holds more than 8 tiles (12 at
    the sizes here): only the block-to-tile assignment changes, so 3D steps
    and batched 2D steps, FP64 and FP32, give the default build's bits."""
    from emu_lib import Emu2D, Emu3D, load
    from phrike_amd.initial_conditions import khi_2d_periodic, taylor_green_3d
    paired = load(defines=("-DPH_COL_XCD_PAIR_ALL=1", "-DPH_COL_TM_MAX=2"), tag="pair")
    rng = np.random.default_rng(3)
    # kept x columns 22 (FP64, row pitch 24) and 43 (FP32, 24 packed pairs)
    for dtype, nx in ((np.float64, 64), (np.float32, 128)):
        U = (taylor_green_3d(nx, 16, 8) + 1e-3 * rng.standard_normal((5, 8, 16, nx))).astype(dtype)
        a, b = Emu3D(emu, nx, 16, 8, dtype=dtype), Emu3D(paired, nx, 16, 8, dtype=dtype)
        np.testing.assert_array_equal(a.step(U, 1e-3), b.step(U, 1e-3))
        W = np.ascontiguousarray(np.stack([khi_2d_periodic(nx, 32, seed=s) for s in range(3)]).astype(dtype))
        a2, b2 = Emu2D(emu, nx, 32, batch=3, dtype=dtype), Emu2D(paired, nx, 32, batch=3, dtype=dtype)
        np.testing.assert_array_equal(a2.step(W, 1e-3), b2.step(W, 1e-3))
