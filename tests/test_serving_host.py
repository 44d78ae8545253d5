"""Host logic of the serving workload (no GPU)."""

import numpy as np

from astronomix_amd.serving import initial_state, mixed_requests
from astronomix_amd.serving.engine import initial_state_shape, latency_summary
from astronomix_amd.serving.requests import GRIDS


def test_request_stream_is_seeded_and_in_range():
    a = mixed_requests(200, rate=50.0, seed=0)
    b = mixed_requests(200, rate=50.0, seed=0)
    assert a == b
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    # Poisson rate 50/s: mean gap 20 ms
    assert 0.015 < np.diff(arr).mean() < 0.025
    assert all(1.4 <= r.gamma <= 5 / 3 and 0.1 <= r.t_end <= 0.5 and 5 <= r.p_in <= 50 for r in a)
    assert {r.grid_class for r in a} == set(GRIDS)


def test_initial_states_have_the_reference_layout():
    for r in mixed_requests(30, seed=1, grids=(("orszag_tang_2d", 16), ("mhd_blast_3d", 16))):
        s = initial_state(r)
        assert s.shape == initial_state_shape(r) and s.dtype == np.float32
        if r.kind == "mhd_blast_3d":
            assert s[4].max() == np.float32(r.p_in)


def test_latency_summary():
    res = {i: dict(arrival=0.1 * i, finished=0.1 * i + 0.5 + 0.01 * i) for i in range(10)}
    s = latency_summary(res)
    assert s["requests"] == 10
    assert abs(s["makespan_s"] - (0.9 + 0.5 + 0.09)) < 1e-9
    assert abs(s["p50_latency_s"] - np.percentile([0.5 + 0.01 * i for i in range(10)], 50)) < 1e-4
