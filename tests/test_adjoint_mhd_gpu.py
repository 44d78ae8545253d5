"""Reverse mode of the MHD step (amx_fv_adjoint_step with mhd = 1, and the
checkpointed rollout) against torch float64 autograd of the restatement
(oracle/mhd_torch.py) with the GPU forward's own number of fixed-point
applications, and against finite differences of the GPU forward.

As for the hydro step, a minmod scheme's gradient jumps where a limiter
switches branch and float32 / float64 trajectories can sit on different
sides at a few cells: the checks are directional derivatives <g, d> for
random normal d within 1e-3 of their typical size |g|_2, and at least
99.5 % of entries within 1e-3 of max |g|."""

import numpy as np
import pytest

from astronomix_amd import _native
from astronomix_amd import workloads as W
from astronomix_amd._finite_volume import native_config
from astronomix_amd.option_classes.simulation_config import HLL, HLLC, IMPLICIT_EULER, IMPLICIT_MIDPOINT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mt():
    from oracle import mhd_torch
    return mhd_torch


def mhd_state(shape, seed):
    rng = np.random.default_rng(seed)
    s = W.random_smooth(shape, True, seed=seed).astype(np.float64)
    s += 1e-2 * rng.standard_normal(s.shape)
    return s.astype(np.float32)


def compare(g, ref, tag):
    """<g, d> for random normal d within 1e-3 of its typical size |ref|_2
    (or of |<ref, d>| where that is larger), and at least 99.5 % of entries
    within 1e-3 of max |ref|."""
    g = np.asarray(g, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert np.all(np.isfinite(g)), tag
    scale = float(np.sqrt((ref * ref).sum()))
    rng = np.random.default_rng(0)
    for _ in range(4):
        d = rng.standard_normal(ref.shape)
        a, b = float((g * d).sum()), float((ref * d).sum())
        assert abs(a - b) <= 1e-3 * max(abs(b), scale), f"{tag}: <g,d> {a} vs {b} (|ref|_2 {scale})"
    bad = np.abs(g - ref) > 1e-3 * np.abs(ref).max()
    assert bad.mean() <= 5e-3, f"{tag}: {bad.mean():.2%} entries off"


def mhd_ctx(state, riemann, integrator=IMPLICIT_MIDPOINT):
    dim = state.ndim - 1
    cfg = W.fv_config(dim, state.shape[1], mhd=True, riemann=riemann)._replace(fv_magnetic_integrator=integrator)
    nc = native_config(state.shape, cfg, 5 / 3, 0.4)
    return _native.FvContext(_native.AmxFvConfig.from_buffer_copy(nc))


CASES = [
    ("2d_hll_midpoint", (32, 32), HLL, IMPLICIT_MIDPOINT, 0.004),
    ("2d_hllc_midpoint", (32, 32), HLLC, IMPLICIT_MIDPOINT, 0.004),
    ("2d_hll_euler", (32, 32), HLL, IMPLICIT_EULER, 0.004),
    ("3d_hllc_midpoint", (12, 12, 12), HLLC, IMPLICIT_MIDPOINT, 0.004),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_one_mhd_step_vjp_matches_torch(gpu, mt, case):
    import torch
    name, shape, riemann, integrator, dt = case
    print(f"[gpu] MHD adjoint one step {name}", flush=True)
    dim = len(shape)
    s = mhd_state(shape, 3)
    lam_out = np.random.default_rng(4).standard_normal(s.shape).astype(np.float32)
    ctx = mhd_ctx(s, riemann, integrator)
    S = torch.from_numpy(s).cuda()
    dtt = torch.tensor(dt, dtype=torch.float32, device="cuda")
    # the forward's fixed-point count (final launch index K: K + 1 applications)
    K = ctx.evolve_state(S.clone(), dtt)
    assert K >= 1
    L = torch.from_numpy(lam_out).cuda()
    Lin = torch.empty_like(L)
    ctx.adjoint_step(S, dtt, L, Lin)
    torch.cuda.synchronize()
    assert torch.equal(S.cpu(), torch.from_numpy(s))  # the input state is left as it was
    w = torch.from_numpy(s.astype(np.float64)).requires_grad_(True)
    out = mt.mhd_step(w, dim, 5 / 3, 1.0 / shape[0], float(np.float32(dt)), riemann, K + 1,
                      integrator == IMPLICIT_EULER)
    (ref,) = torch.autograd.grad(out, w, torch.from_numpy(lam_out.astype(np.float64)))
    compare(Lin.cpu().numpy(), ref.numpy(), name)


def _mhd_rollout_grad(s, steps, t_end, every):
    import torch
    from astronomix_amd.adjoint import rollout
    from astronomix_amd.option_classes.simulation_params import SimulationParams
    cfg = W.fv_config(2, s.shape[1], mhd=True, riemann=HLL)._replace(fixed_timestep=True, num_timesteps=steps)
    params = SimulationParams(C_cfl=0.4, t_end=t_end, gamma=5 / 3)
    target = torch.from_numpy(np.random.default_rng(9).random(s.shape[1:]).astype(np.float32)).cuda()
    w = torch.from_numpy(s).cuda().requires_grad_(True)
    out = rollout(w, cfg, params, checkpoint_every=every)
    loss = ((out[0] - target) ** 2).sum() + (out[-1] ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    return w.grad.cpu().numpy(), target.cpu().numpy(), out.detach().cpu().numpy()


def test_mhd_rollout_gradient_matches_torch_and_checkpointing_is_exact(gpu, mt):
    """Eight fixed steps of a 2D MHD state through the checkpointed rollout
    (every 3 and every step), against torch autograd of the restatement
    stepped with the GPU's fixed-point counts of each step; the loss reads
    the density and B_z."""
    import torch
    print("[gpu] MHD rollout gradient", flush=True)
    s = mhd_state((24, 24), 7)
    steps, t_end = 8, 0.024
    g3, target, out = _mhd_rollout_grad(s, steps, t_end, 3)
    g1, _, _ = _mhd_rollout_grad(s, steps, t_end, 1)
    np.testing.assert_array_equal(g1, g3)
    # the forward's step sizes (t_end / steps, the last one clamped to the
    # float32 time left) and per-step counts, for the restatement
    from astronomix_amd.adjoint import step_sizes
    ctx = mhd_ctx(s, HLL)
    S = torch.from_numpy(s).cuda()
    dts = step_sizes(S, ctx, t_end, steps, True)
    ks = [ctx.evolve_state(S, dt) for dt in dts]
    torch.cuda.synchronize()
    np.testing.assert_array_equal(S.cpu().numpy(), out)
    w = torch.from_numpy(s.astype(np.float64)).requires_grad_(True)
    x = w
    for dt, k in zip(dts, ks):
        x = mt.mhd_step(x, 2, 5 / 3, 1.0 / 24, float(dt.item()), HLL, k + 1)
    loss = ((x[0] - torch.from_numpy(target.astype(np.float64))) ** 2).sum() + (x[-1] ** 2).sum()
    (ref,) = torch.autograd.grad(loss, w)
    compare(g3, ref.numpy(), "MHD rollout")


def test_mhd_gradient_against_gpu_finite_differences(gpu):
    """Directional derivatives of a 4-step 2D MHD loss by the reverse mode
    against central differences of the GPU forward (float32, step 1e-2
    along smooth directions, where the forward's fixed-point counts and
    limiter branches stay put): within 2 % of the derivative."""
    import torch
    from astronomix_amd.adjoint import rollout
    from astronomix_amd.option_classes.simulation_params import SimulationParams
    print("[gpu] MHD adjoint vs finite differences", flush=True)
    s = W.random_smooth((32, 32), True, seed=11).astype(np.float32)
    cfg = W.fv_config(2, 32, mhd=True, riemann=HLL)._replace(fixed_timestep=True, num_timesteps=4)
    params = SimulationParams(C_cfl=0.4, t_end=0.01, gamma=5 / 3)
    tgt = torch.from_numpy(np.random.default_rng(2).random((32, 32)).astype(np.float32)).cuda()

    def loss_of(x):
        out = rollout(x, cfg, params, checkpoint_every=2)
        return ((out[0] - tgt) ** 2).sum() + (out[4] * out[5]).sum()

    w = torch.from_numpy(s).cuda().requires_grad_(True)
    loss_of(w).backward()
    g = w.grad.double().cpu()
    for k in range(3):
        # smooth directions: differences of two smooth states
        d = torch.from_numpy(np.stack([W.random_smooth((32, 32), True, seed=20 + k)[v] - s[v] for v in range(s.shape[0])])
                             .astype(np.float32))
        h = 1e-2
        with torch.no_grad():
            lp = float(loss_of((torch.from_numpy(s) + h * d).cuda()).double())
            lm = float(loss_of((torch.from_numpy(s) - h * d).cuda()).double())
        fd = (lp - lm) / (2 * h)
        ad = float((g * d.double()).sum())
        assert abs(fd - ad) <= 2e-2 * max(abs(ad), 1e-3), (k, fd, ad)


def test_mhd_adaptive_rollout_steps_equal_the_device_loop(gpu):
    """Adaptive dt on an MHD state: the rollout's recorded steps are the
    device loop's (same count, bit-identical final state), and the public
    time_integration differentiates an MHD state through them."""
    import torch
    from astronomix_amd import get_registered_variables
    from astronomix_amd.adjoint import rollout, step_sizes
    from astronomix_amd.option_classes.simulation_params import SimulationParams
    from astronomix_amd.time_stepping.time_integration import time_integration
    print("[gpu] MHD adaptive rollout vs device loop", flush=True)
    s = W.orszag_tang_2d(32)
    cfg = W.fv_config(2, 32, mhd=True, riemann=HLL)
    params = SimulationParams(C_cfl=0.4, t_end=0.05, gamma=5 / 3)
    ctx = _native.fv_context(native_config(s.shape, cfg, params.gamma, params.C_cfl))
    S = torch.from_numpy(s).cuda()
    st = ctx.time_integration(S, _native.AmxFvRun(t_end=0.05, fixed_timestep=0, num_timesteps=0,
                                                  exact_end_time=1, max_steps=-1))
    w = torch.from_numpy(s).cuda().requires_grad_(True)
    out = rollout(w, cfg, params, checkpoint_every=7)
    torch.cuda.synchronize()
    n = len(step_sizes(torch.from_numpy(s).cuda(), ctx, 0.05))
    assert n == st.steps and n > 5
    np.testing.assert_array_equal(out.detach().cpu().numpy(), S.cpu().numpy())
    w2 = torch.from_numpy(s).cuda().requires_grad_(True)
    out2 = time_integration(w2, cfg, params, get_registered_variables(cfg))
    out2[0].sum().backward()
    torch.cuda.synchronize()
    assert torch.isfinite(w2.grad).all() and float(w2.grad.abs().max()) > 0


def test_ot_64_mhd_20_steps_gradient_against_torch_float64(gpu, mt):
    """Orszag-Tang 64^2 with a 1e-2 random field added to every variable, 20
    fixed steps through the checkpointed rollout, against torch float64
    autograd with the GPU forward's per-step fixed-point counts. (Without
    the noise, the uniform density and pressure of the Orszag-Tang start put
    minmod at its ties over the whole grid, where the derivative is
    one-sided and float32 and float64 trajectories choose different sides:
    the directional derivatives then differ by 20 %.)"""
    import torch
    from astronomix_amd.adjoint import rollout, step_sizes
    from astronomix_amd.option_classes.simulation_params import SimulationParams
    print("[gpu] MHD OT 64^2 20 steps gradient", flush=True)
    n, steps, t_end = 64, 20, 0.02
    s = (W.orszag_tang_2d(n).astype(np.float64)
         + 1e-2 * np.random.default_rng(8).standard_normal((7, n, n))).astype(np.float32)
    cfg = W.fv_config(2, n, mhd=True, riemann=HLL)._replace(fixed_timestep=True, num_timesteps=steps)
    params = SimulationParams(C_cfl=0.4, t_end=t_end, gamma=5 / 3)
    target = torch.from_numpy(np.random.default_rng(3).random((n, n)).astype(np.float32)).cuda()
    w = torch.from_numpy(s).cuda().requires_grad_(True)
    out = rollout(w, cfg, params, checkpoint_every=5)
    (((out[0] - target) ** 2).sum() + (out[5] ** 2).sum()).backward()
    torch.cuda.synchronize()
    ctx = mhd_ctx(s, HLL)
    S = torch.from_numpy(s).cuda()
    dts = step_sizes(S, ctx, t_end, steps, True)
    ks = [ctx.evolve_state(S, dt) for dt in dts]
    torch.cuda.synchronize()
    x = w64 = torch.from_numpy(s.astype(np.float64)).requires_grad_(True)
    for dt, k in zip(dts, ks):
        x = mt.mhd_step(x, 2, 5 / 3, 1.0 / n, float(dt.item()), HLL, k + 1)
    loss = ((x[0] - torch.from_numpy(target.cpu().numpy().astype(np.float64))) ** 2).sum() + (x[5] ** 2).sum()
    (ref,) = torch.autograd.grad(loss, w64)
    compare(w.grad.cpu().numpy(), ref.numpy(), "OT 64 MHD rollout")
