"""The differentiable torch restatement (oracle/hydro_torch.py) against the
C++ restatement's forward step, and its autograd against central finite
differences -- so that it can serve as the gradient reference."""

import numpy as np
import pytest
import torch

from astronomix_amd import workloads as W


@pytest.fixture(scope="module")
def ht():
    from oracle import hydro_torch
    return hydro_torch


@pytest.mark.parametrize("riemann", [0, 1])
def test_torch_step_equals_cpp_oracle(oracle, ht, riemann):
    s = W.kelvin_helmholtz_2d(32, seed=2).astype(np.float64)
    oc = oracle.make_config(s.shape, 2, mhd=False, riemann=riemann, gamma=5 / 3, t_end=0.02, fixed_timestep=True,
                            num_timesteps=5)
    ref = oracle.time_integration(s, oc)[0]
    out = ht.rollout(torch.from_numpy(s), 5, 2, 5 / 3, 1.0 / 32, 0.02 / 5, riemann).numpy()
    assert np.abs(out - ref).max() < 1e-12


def test_torch_step_3d_equals_cpp_oracle(oracle, ht):
    s = W.random_smooth((12, 12, 12), False, seed=4).astype(np.float64)
    oc = oracle.make_config(s.shape, 3, mhd=False, riemann=1, gamma=1.4, t_end=0.004, fixed_timestep=True,
                            num_timesteps=2)
    ref = oracle.time_integration(s, oc)[0]
    out = ht.rollout(torch.from_numpy(s), 2, 3, 1.4, 1.0 / 12, 0.002, 1).numpy()
    assert np.abs(out - ref).max() < 1e-12


def generic_state(n, seed):
    """KH plus a small random field on every variable: no two neighbouring
    values are equal, so minmod is differentiable along every path (at exact
    ties -- KH's flat layers -- it is not, and autograd's zero derivative
    there is not a finite-difference limit)."""
    rng = np.random.default_rng(seed)
    s = W.kelvin_helmholtz_2d(n, seed=seed).astype(np.float64)
    s += 1e-2 * rng.standard_normal(s.shape) * np.array([1, 1, 1, 1])[:, None, None]
    return s


def test_torch_gradient_matches_finite_differences(ht):
    torch.manual_seed(0)
    s = torch.from_numpy(generic_state(16, 1))
    target = torch.rand_like(s[0])

    def loss(w):
        return ((ht.rollout(w, 4, 2, 5 / 3, 1 / 16, 0.004, 1)[0] - target) ** 2).sum()

    w = s.clone().requires_grad_(True)
    loss(w).backward()
    g = w.grad
    for k in range(4):
        d = torch.randn_like(s)
        eps = 1e-7  # minmod switches within eps |d| of the state bias larger steps
        fd = (loss(s + eps * d) - loss(s - eps * d)) / (2 * eps)
        assert abs(float(fd) - float((g * d).sum())) <= 1e-6 * max(1.0, abs(float(fd)))
