"""Host logic of the reverse-mode rollout on the CPU: the adaptive step-size
loop of ``astronomix_amd.adjoint.step_sizes`` with a stub context (a fixed
CFL step, a no-op step), against the float32 loop of the device
(k_step_begin / k_step_end: run while t < t_end in float32, dt clamped to
t_end - t when exact_end_time). No GPU calls."""

import numpy as np
import pytest

from astronomix_amd.adjoint import step_sizes


class StubCtx:
    """cfl_dt writes a fixed step; evolve_state counts calls and refuses to
    run away (the loop under test must end by itself)."""

    def __init__(self, dt):
        self.dt, self.calls = np.float32(dt), 0

    def cfl_dt(self, state, out):
        out.fill_(float(self.dt))

    def evolve_state(self, state, dt):
        self.calls += 1
        assert self.calls < 10_000, "step_sizes does not terminate"


def device_loop(dt, t_end, exact_end):
    te, t, out = np.float32(t_end), np.float32(0.0), []
    while t < te:
        d = np.float32(dt)
        if exact_end:
            d = min(d, np.float32(te - t))
        out.append(d)
        t = np.float32(t + d)
    return out


@pytest.mark.parametrize("t_end", [0.7, 0.9, 0.35, 0.45, 0.5, 1.0])
@pytest.mark.parametrize("dt", [0.1, 0.03, 0.0123])
@pytest.mark.parametrize("exact_end", [True, False])
def test_adaptive_step_sizes_match_the_device_loop(t_end, dt, exact_end):
    import torch
    # float32(t_end) < t_end for 0.7, 0.9, 0.35, 0.45: the clamp lands t on
    # float32(t_end), where a comparison against the double t_end never ends
    ctx = StubCtx(dt)
    got = step_sizes(torch.zeros(4, 8, 8), ctx, t_end, None, exact_end)
    want = device_loop(dt, t_end, exact_end)
    assert [np.float32(float(x)) for x in got] == want
    assert ctx.calls == len(want)
    assert all(float(x) > 0 for x in got)


def test_fixed_step_sizes_sum_to_t_end():
    import torch
    got = step_sizes(torch.zeros(4, 8, 8), None, 0.7, 7, True)
    assert len(got) == 7
    t = np.float32(0)
    for x in got:
        t = np.float32(t + np.float32(float(x)))
    assert abs(float(t) - float(np.float32(0.7))) <= 2e-7
