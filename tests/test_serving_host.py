"""Host logic of the ensemble server: request streams and the HBM split."""
import numpy as np

from phrike_amd.serving import _pow2_ceil, capacities_for_budget, poisson_requests, slot_bytes


def test_poisson_stream_is_seeded_and_weighted():
    sizes = {128: 0.4, 256: 0.3, 512: 0.2, 1024: 0.1}
    a = poisson_requests(50.0, 40.0, sizes, seed=3)
    b = poisson_requests(50.0, 40.0, sizes, seed=3)
    assert [(r.n, r.seed, r.t_end, r.arrival) for r in a] == [(r.n, r.seed, r.t_end, r.arrival) for r in b]
    assert abs(len(a) - 2000) < 4 * np.sqrt(2000)
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0) and arr[-1] <= 40.0
    frac = {n: np.mean([r.n == n for r in a]) for n in sizes}
    for n, w in sizes.items():
        assert abs(frac[n] - w) < 0.04
    assert all(0.5 <= r.t_end <= 2.0 for r in a)


def test_capacities_follow_work_share():
    sizes = {128: 0.4, 256: 0.3, 512: 0.2, 1024: 0.1}
    budget = 8 << 30
    caps = capacities_for_budget(sizes, budget, "single")
    assert sum(caps[n] * slot_bytes(n, "single") for n in sizes) <= budget
    assert all(1 <= c <= 64 for c in caps.values())
    assert slot_bytes(1024, "single") > 60 * slot_bytes(128, "single")
    assert slot_bytes(256, "double") > slot_bytes(256, "single")


def test_pow2_ceil():
    assert [_pow2_ceil(k) for k in (0, 1, 2, 3, 4, 5, 63, 64, 65)] == [1, 1, 2, 4, 4, 8, 64, 64, 128]
