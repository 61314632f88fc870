"""Ensemble server: a run's result is the same bits alone and in a mix, and
equals the one-run SpectralSolver2D.run loop (phrike/solver.py:706-812)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _server(caps, **kw):
    from phrike_amd.serving import EnsembleServer
    return EnsembleServer(caps, precision="double", steps_per_replay=6, keep_states=True, **kw)


def test_alone_equals_mixed_and_solver_run(dev):
    import torch
    from phrike_amd.serving import RunRequest
    from phrike_amd.initial_conditions import khi_2d_periodic
    reqs = [RunRequest(0, 64, 11, 0.021), RunRequest(1, 32, 12, 0.013),
            RunRequest(2, 64, 13, 0.017), RunRequest(3, 32, 14, 0.009), RunRequest(4, 64, 15, 0.005)]
    mixed = _server({32: 2, 64: 2}).serve(reqs, realtime=False)
    assert [r.rid for r in mixed] == [0, 1, 2, 3, 4]
    alone = _server({64: 1}).serve([reqs[0]], realtime=False)[0]
    print("steps", [r.steps for r in mixed], flush=True)
    np.testing.assert_array_equal(alone.state, mixed[0].state)
    assert alone.steps == mixed[0].steps and alone.t_final == mixed[0].t_final
    # the same run through the host run loop of SpectralSolver2D
    from phrike_amd.grid import Grid2D
    from phrike_amd.equations import EulerEquations2D
    from phrike_amd.solver import SpectralSolver2D
    g = Grid2D(64, 64, 1.0, 1.0, filter_params={"enabled": True, "p": 16, "alpha": 36.0, "interval": 1},
               backend="hip", precision="double")
    s = SpectralSolver2D(g, EulerEquations2D(1.4), cfl=0.4)
    U0 = torch.as_tensor(khi_2d_periodic(64, 64, seed=11), device=dev)
    s.run(U0, 0.0, 0.021, output_interval=1e9)
    assert abs(s.t - mixed[0].t_final) == 0.0
    np.testing.assert_array_equal(s.U.cpu().numpy(), mixed[0].state)


def test_eager_equals_graph(dev):
    from phrike_amd.serving import RunRequest
    reqs = [RunRequest(0, 32, 5, 0.011), RunRequest(1, 32, 6, 0.007)]
    a = _server({32: 2}).serve(reqs, realtime=False)
    b = _server({32: 2}, use_graph=False).serve(reqs, realtime=False)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.state, y.state)


def test_pool_mode_resizes_and_matches_fixed(dev):
    """Memory-managed buckets (hbm_budget): buckets grow from zero slots as
    requests queue, carry their active runs across a resize and shrink when
    idle; every result equals the fixed-capacity server's bit for bit."""
    from phrike_amd.serving import EnsembleServer, RunRequest, slot_bytes
    reqs = [RunRequest(i, 32 if i % 2 else 64, 20 + i, 0.004 + 0.002 * i, arrival=0.03 * i) for i in range(8)]
    fixed = _server({32: 4, 64: 4}).serve(reqs, realtime=False)
    budget = 3 * slot_bytes(64, "double") + 2 * slot_bytes(32, "double")
    srv = EnsembleServer(hbm_budget=budget, sizes=(32, 64), precision="double", steps_per_replay=6,
                         keep_states=True, max_slots=4)
    pool = srv.serve(reqs, realtime=True)
    assert [r.rid for r in pool] == list(range(8))
    assert srv.peak_bytes <= budget
    assert sum(b.resizes for b in srv.buckets.values()) >= 2
    for a, b in zip(fixed, pool):
        assert a.steps == b.steps and a.t_final == b.t_final
        np.testing.assert_array_equal(a.state, b.state)
