"""Parity of the HIP finite-volume path against the CPU restatement (oracle).

Every case runs the same seeded float32 input through the GPU (via the C ABI,
``libamx.so``) and through ``oracle/fv_oracle.cpp``. Bar: relative L1 per
variable <= 1e-5 against the float32 oracle for runs of up to ~70 steps
(the float32-vs-float64 oracle gap on these cases is 1e-7..3e-6). Longer runs
are compared against the float64 oracle and must stay within 3x the
float32 oracle's own distance to it.
"""

import numpy as np
import pytest

from astronomix_amd import workloads as W
from astronomix_amd import _native
from astronomix_amd._finite_volume import native_config
from astronomix_amd.option_classes.simulation_config import HLL, HLLC, LAX_FRIEDRICHS, MINMOD, VAN_ALBADA_PP
from astronomix_amd.option_classes.simulation_params import SimulationParams

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel_l1(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.array([np.abs(a[v] - b[v]).sum() / max(np.abs(b[v]).sum(), 1e-30) for v in range(a.shape[0])])


def gpu_run(state, cfg, params, *, fixed=False, num_timesteps=0, max_steps=-1):
    import torch
    s = torch.from_numpy(np.ascontiguousarray(state, dtype=np.float32)).cuda()
    ctx = _native.fv_context(native_config(s.shape, cfg, params.gamma, params.C_cfl))
    run = _native.AmxFvRun(t_end=float(params.t_end), fixed_timestep=int(fixed), num_timesteps=int(num_timesteps),
                           exact_end_time=1, max_steps=int(max_steps))
    st = ctx.time_integration(s, run)
    torch.cuda.synchronize()
    return s.cpu().numpy(), st


def oracle_run(oracle, state, dim, params, *, mhd=True, riemann=0, fixed=False, num_timesteps=1000, dtype=np.float32):
    oc = oracle.make_config(state.shape, dim, mhd=mhd, riemann=riemann, gamma=params.gamma, C_cfl=params.C_cfl,
                            t_end=params.t_end, fixed_timestep=fixed, num_timesteps=num_timesteps)
    return oracle.time_integration(state.astype(dtype), oc)


CASES = [
    # name, state builder, dim, mhd, riemann, C_cfl, t_end
    ("blast_3d_32_hll", lambda: W.blast_3d(32), 3, True, HLL, 0.3, 0.005),
    ("blast_3d_32_hllc", lambda: W.blast_3d(32), 3, True, HLLC, 0.3, 0.005),
    ("random_3d_24_hll", lambda: W.random_smooth((24, 24, 24), True, seed=1), 3, True, HLL, 0.4, 0.01),
    ("random_3d_20_hllc", lambda: W.random_smooth((20, 20, 20), True, seed=2), 3, True, HLLC, 0.4, 0.01),
    ("random_3d_72_hll", lambda: W.random_smooth((72, 72, 72), True, seed=3), 3, True, HLL, 0.4, 0.004),
    # 21 planes: x chunks of 8, 8 and 5 (the stage kernels' lagged update
    # ends a chunk of odd length on the other register and buffer parity)
    ("random_3d_21_hllc", lambda: W.random_smooth((21, 21, 21), True, seed=6), 3, True, HLLC, 0.4, 0.01),
    ("hydro_3d_24_hll", lambda: W.random_smooth((24, 24, 24), False, seed=4), 3, False, HLL, 0.4, 0.02),
    ("ot_2d_64_hll", lambda: W.orszag_tang_2d(64), 2, True, HLL, 0.4, 0.1),
    ("ot_2d_100_hllc", lambda: W.orszag_tang_2d(100), 2, True, HLLC, 0.4, 0.05),
    ("random_2d_130_hll", lambda: W.random_smooth((130, 130), True, seed=5), 2, True, HLL, 0.4, 0.02),
    ("kh_2d_64_hllc", lambda: W.kelvin_helmholtz_2d(64), 2, False, HLLC, 0.4, 0.05),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_time_integration_matches_oracle(gpu, oracle, case):
    name, build, dim, mhd, riemann, C, t_end = case
    print(f"[gpu] parity case {name}", flush=True)
    state = build()
    n = state.shape[-1]
    cfg = W.fv_config(dim, n, mhd=mhd, riemann=riemann)
    params = SimulationParams(C_cfl=C, t_end=t_end, gamma=5.0 / 3.0)
    out, st = gpu_run(state, cfg, params)
    ref, steps, t, iters = oracle_run(oracle, state, dim, params, mhd=mhd, riemann=riemann)
    assert st.steps == steps
    assert abs(st.t - t) <= 1e-6 * max(1.0, abs(t))
    if mhd:
        assert st.fixed_point_iterations == iters
    err = rel_l1(out, ref)
    assert np.all(np.isfinite(out))
    if err.max() > TOL:
        # Variables whose own float32 round-off floor is above the bar (the
        # seeded 1e-3 noise in KH's vy, amplified by the shear instability):
        # the GPU must then be as close to the float64 oracle as the float32
        # oracle is (within 3x), and every other variable still meets 1e-5.
        r64, _, _, _ = oracle_run(oracle, state, dim, params, mhd=mhd, riemann=riemann, dtype=np.float64)
        floor = rel_l1(ref, r64)
        err64 = rel_l1(out, r64)
        for v in range(len(err)):
            if err[v] > TOL:
                assert floor[v] > TOL / 3, f"{name}: var {v} rel L1 {err[v]:.3e} (float32 floor {floor[v]:.3e})"
                assert err64[v] <= 3 * floor[v], f"{name}: var {v} {err64[v]:.3e} vs floor {floor[v]:.3e}"


def test_orszag_tang_256_full_run_tracks_float64(gpu, oracle):
    """The 2D Orszag-Tang configuration in full (256^2, HLL, CFL 0.4, to
    t = 0.5: 1426 steps): the GPU must stay as close to the float64 oracle as
    the float32 oracle does (within 3x)."""
    print("[gpu] OT 256^2 full run", flush=True)
    state = W.orszag_tang_2d(256)
    cfg = W.fv_config(2, 256)
    params = SimulationParams(C_cfl=0.4, t_end=0.5, gamma=5.0 / 3.0)
    out, st = gpu_run(state, cfg, params)
    r64, steps, _, _ = oracle_run(oracle, state, 2, params, dtype=np.float64)
    r32, _, _, _ = oracle_run(oracle, state, 2, params, dtype=np.float32)
    assert st.steps == steps
    floor = rel_l1(r32, r64)
    err = rel_l1(out, r64)
    for v in range(6):
        assert err[v] <= 3 * floor[v] + 1e-6, f"var {v}: gpu {err[v]:.3e} vs float32 oracle {floor[v]:.3e}"


def test_fixed_timestep_mode(gpu, oracle):
    print("[gpu] fixed timestep", flush=True)
    state = W.blast_3d(24)
    cfg = W.fv_config(3, 24)._replace(fixed_timestep=True, num_timesteps=4)
    params = SimulationParams(C_cfl=0.3, t_end=0.004, gamma=5.0 / 3.0)
    out, st = gpu_run(state, cfg, params, fixed=True, num_timesteps=4)
    ref, steps, _, _ = oracle_run(oracle, state, 3, params, fixed=True, num_timesteps=4)
    assert st.steps == steps == 4
    assert rel_l1(out, ref).max() <= TOL


def test_public_time_integration_numpy_roundtrip(gpu, oracle):
    import astronomix_amd as amx
    print("[gpu] public time_integration", flush=True)
    state, cfg, params = W.workload("orszag_tang_2d", 32)
    params = params._replace(t_end=0.05)
    rv = amx.get_registered_variables(cfg)
    out = amx.time_integration(state, cfg, params, rv)
    assert isinstance(out, np.ndarray) and out.shape == state.shape
    ref, _, _, _ = oracle_run(oracle, state, 2, params)
    assert rel_l1(out, ref).max() <= TOL


def test_t_end_zero_is_identity(gpu):
    import astronomix_amd as amx
    print("[gpu] t_end = 0", flush=True)
    state, cfg, params = W.workload("orszag_tang_2d", 32)
    out = amx.time_integration(state, cfg, params._replace(t_end=0.0), amx.get_registered_variables(cfg))
    np.testing.assert_array_equal(out, state)


def test_cfl_time_step_operator(gpu, oracle):
    import torch
    from astronomix_amd._finite_volume import _cfl_time_step
    print("[gpu] cfl operator", flush=True)
    for dim, state in ((3, W.random_smooth((16, 16, 16), True, seed=7)), (2, W.orszag_tang_2d(48))):
        n = state.shape[-1]
        cfg = W.fv_config(dim, n)
        s = torch.from_numpy(state).cuda()
        dt = float(_cfl_time_step(s, cfg.grid_spacing, 0.001, 5.0 / 3.0, cfg, None, 0.4).item())
        oc = oracle.make_config(state.shape, dim, C_cfl=0.4)
        ref = oracle.cfl_dt(state[: dim + 2], oc)
        assert abs(dt - ref) <= 2e-7 * ref


def test_evolve_state_operator_and_iteration_count(gpu, oracle):
    import torch
    from astronomix_amd._finite_volume import _evolve_state_fv
    print("[gpu] evolve_state operator", flush=True)
    state = W.random_smooth((16, 16, 16), True, seed=8)
    cfg = W.fv_config(3, 16)
    params = SimulationParams(C_cfl=0.4, gamma=5.0 / 3.0)
    dt = 2e-3
    s = torch.from_numpy(state).cuda()
    out = _evolve_state_fv(s, dt, params.gamma, 1.0, cfg, params, None,
                           __import__("astronomix_amd").get_registered_variables(cfg))
    oc = oracle.make_config(state.shape, 3, C_cfl=0.4)
    ref, k = oracle.evolve_state_fv(state, oc, dt)
    assert rel_l1(out.cpu().numpy(), ref).max() <= TOL
    # the input stays untouched (functional operator)
    np.testing.assert_array_equal(s.cpu().numpy(), state)


def _central_div(B, dim):
    d = np.zeros(B.shape[1:], dtype=np.float64)
    for a in range(dim):
        f = B[a].astype(np.float64)
        d += np.roll(f, -1, axis=a) - np.roll(f, 1, axis=a)
    return d


def test_divergence_and_mass_invariants_at_scale(gpu):
    """Size-independent properties at 128^3: the magnetic sub-step is a
    discrete curl, so the central-difference div B of every cell is unchanged
    to float32 round-off; total mass is conserved by the flux form."""
    print("[gpu] invariants 128^3", flush=True)
    state = W.random_smooth((128, 128, 128), True, seed=9)
    cfg = W.fv_config(3, 128)
    params = SimulationParams(C_cfl=0.4, t_end=1e9, gamma=5.0 / 3.0)
    out, st = gpu_run(state, cfg, params, max_steps=3)
    assert st.steps == 3
    d0 = _central_div(state[5:8], 3)
    d1 = _central_div(out[5:8], 3)
    scale = np.abs(state[5:8]).max()
    assert np.abs(d1 - d0).max() <= 1e-5 * scale
    m0 = state[0].astype(np.float64).sum()
    m1 = out[0].astype(np.float64).sum()
    assert abs(m1 - m0) <= 1e-5 * m0


def test_blast_256_invariants(gpu):
    """The 3D blast configuration at its full size (256^3), 10 steps:
    central div B unchanged to float32 round-off, mass conserved, and the
    x <-> y mirror symmetry of the problem (B = (1,1,0)/sqrt 2, spherical
    pressure) kept to round-off although the kernels treat x (marching) and
    y (in-plane) differently."""
    print("[gpu] blast 256^3 invariants", flush=True)
    state, cfg, params = W.workload("mhd_blast_3d", 256)
    out, st = gpu_run(state, cfg, params, max_steps=10)
    assert st.steps == 10
    d0 = _central_div(state[5:8], 3)
    d1 = _central_div(out[5:8], 3)
    assert np.abs(d1 - d0).max() <= 1e-5
    m0 = state[0].astype(np.float64).sum()
    assert abs(out[0].astype(np.float64).sum() - m0) <= 1e-6 * m0
    mirror = out[[0, 2, 1, 3, 4, 6, 5, 7]].transpose(0, 2, 1, 3)
    err = rel_l1(mirror, out)
    assert err.max() <= 1e-5, err
    assert np.abs(out[4] - state[4]).max() > 1.0  # the blast has moved


@pytest.mark.parametrize("case", ["blast3d", "ot2d"])
def test_loop_keeping_b_in_the_iterate_sets_equals_stepwise_copies(gpu, case):
    """The device-resident loop keeps each step's final B in the fixed
    point's iterate set that the next step reads as B0, and writes B into the
    state once after the last step; amx_fv_evolve_state copies B into the
    state every step. The same fixed-dt steps give the same bits."""
    import torch
    print(f"[gpu] loop vs stepwise B {case}", flush=True)
    if case == "blast3d":
        state, dim, n, C = W.blast_3d(32), 3, 32, 0.3
    else:
        state, dim, n, C = W.orszag_tang_2d(96), 2, 96, 0.4
    cfg = W.fv_config(dim, n, mhd=True, riemann=HLL)
    t_end, steps = 0.006, 7
    ctx = _native.fv_context(native_config(state.shape, cfg, 5.0 / 3.0, C))
    s_loop = torch.from_numpy(np.ascontiguousarray(state)).cuda()
    run = _native.AmxFvRun(t_end=t_end, fixed_timestep=1, num_timesteps=steps, exact_end_time=1, max_steps=-1)
    st = ctx.time_integration(s_loop, run)
    assert st.steps == steps and st.fixed_point_iterations > steps
    s_step = torch.from_numpy(np.ascontiguousarray(state)).cuda()
    dt = torch.tensor(np.float32(t_end) / np.float32(steps), dtype=torch.float32, device="cuda")
    iters = sum(ctx.evolve_state(s_step, dt) for _ in range(steps))
    torch.cuda.synchronize()
    assert iters == st.fixed_point_iterations
    assert torch.equal(s_loop, s_step)
    # and a second loop on the same context starts again from the state's B
    s_again = torch.from_numpy(np.ascontiguousarray(state)).cuda()
    ctx.time_integration(s_again, run)
    torch.cuda.synchronize()
    assert torch.equal(s_again, s_loop)


@pytest.mark.parametrize("riemann,limiter", [(HLL, MINMOD), (LAX_FRIEDRICHS, MINMOD), (HLL, VAN_ALBADA_PP)],
                         ids=["hll", "lf", "hll_pp"])
def test_fixed_point_past_the_queued_launches(gpu, oracle, riemann, limiter):
    """At fixed-point tolerance 1e-6 the first step of a fresh context ends at
    launch 6, past the launches (0..5) it queues before the host has read the
    residuals. The second half of the step, queued behind them, is held on the
    device and runs after the further launches; with Lax-Friedrichs and
    VAN_ALBADA_PP its pre-pass kernels (grid-wide alpha, positivity factors)
    are held too and run again from the final iterate. The loop and single
    steps both take that route and equal each other bit for bit, and the
    oracle to the bar with its iteration counts."""
    import torch
    print(f"[gpu] fixed point past the queued launches rs={riemann} lim={limiter}", flush=True)
    state = W.random_smooth((24, 24, 24), True, seed=1, amp=1.0)
    cfg = W.fv_config(3, 24, mhd=True, riemann=riemann)._replace(limiter=limiter)
    nc = native_config(state.shape, cfg, 5.0 / 3.0, 0.8)
    nc.fixed_point_tol = 1e-6
    steps, t_end = 12, 0.02
    loop = _native.FvContext(_native.AmxFvConfig.from_buffer_copy(nc))
    s_loop = torch.from_numpy(np.ascontiguousarray(state)).cuda()
    # no end-time clamp: every step of the loop takes the same dt as the single steps
    run = _native.AmxFvRun(t_end=t_end, fixed_timestep=1, num_timesteps=steps, exact_end_time=0, max_steps=-1)
    st = loop.time_integration(s_loop, run)
    single = _native.FvContext(_native.AmxFvConfig.from_buffer_copy(nc))
    s_step = torch.from_numpy(np.ascontiguousarray(state)).cuda()
    dt = torch.tensor(np.float32(t_end) / np.float32(steps), dtype=torch.float32, device="cuda")
    ks = [single.evolve_state(s_step, dt) for _ in range(steps)]
    torch.cuda.synchronize()
    assert ks[0] > 5, ks
    assert st.steps == steps and st.fixed_point_iterations == sum(ks)
    assert torch.equal(s_loop, s_step)
    oc = oracle.make_config(state.shape, 3, mhd=True, riemann=riemann, limiter=limiter, C_cfl=0.8, t_end=t_end,
                            tol=1e-6, fixed_timestep=True, num_timesteps=steps, exact_end_time=False)
    ref, osteps, _, iters = oracle.time_integration(state.astype(np.float32), oc)
    assert osteps == steps and iters == sum(ks)
    err = rel_l1(s_loop.cpu().numpy(), ref)
    assert err.max() < TOL, err


@pytest.mark.parametrize("max_steps", [-1, 3, 21])
def test_hydro_loop_batches_equal_single_steps(gpu, max_steps):
    """Hydro loops without snapshot rows or a step hook queue their steps in
    batches of 16 and read t once per batch; steps queued past the end do
    nothing. A fixed-dt loop of 21 steps (one full batch and part of one),
    alone or cut at 3 or 21 steps by max_steps, equals single steps bit for
    bit and reports the steps it took."""
    import torch
    print(f"[gpu] hydro loop batches, max_steps {max_steps}", flush=True)
    state = W.kelvin_helmholtz_2d(64)
    cfg = W.fv_config(2, 64, mhd=False, riemann=HLLC)
    ctx = _native.fv_context(native_config(state.shape, cfg, 5.0 / 3.0, 0.4))
    steps, t_end = 21, 0.021
    run = _native.AmxFvRun(t_end=t_end, fixed_timestep=1, num_timesteps=steps, exact_end_time=0,
                           max_steps=max_steps)
    s_loop = torch.from_numpy(np.ascontiguousarray(state)).cuda()
    st = ctx.time_integration(s_loop, run)
    n = steps if max_steps < 0 else min(steps, max_steps)
    assert st.steps == n
    s_step = torch.from_numpy(np.ascontiguousarray(state)).cuda()
    dt = torch.tensor(np.float32(t_end) / np.float32(steps), dtype=torch.float32, device="cuda")
    for _ in range(n):
        ctx.evolve_state(s_step, dt)
    torch.cuda.synchronize()
    assert torch.equal(s_loop, s_step)
    t = np.float32(0.0)
    for _ in range(n):
        t = np.float32(t + np.float32(t_end) / np.float32(steps))
    assert st.t == float(t)


@pytest.mark.parametrize("shape", [(16, 16, 16), (64, 64)], ids=["3d", "2d"])
def test_final_first_iterate_regenerated(gpu, oracle, shape):
    """A fixed-point pair stores its first iterate only where the previous
    step's final launch index says it will be the final one; when a step
    ends on the first launch of a pair that did not store it, a further
    launch recomputes that iterate on the device. Steps of changing dt on one
    context drive the final index up and down (5, 2, 4, 0, 6, 2, 3 in 3D)
    so that both happen, including the final launch 0, whose input is the
    state's own B; every step must match the oracle's step from the same
    input, with the same final launch index."""
    import torch
    print(f"[gpu] regenerated first iterate {shape}", flush=True)
    dim = len(shape)
    state = W.random_smooth(shape, True, seed=8)
    cfg = W.fv_config(dim, shape[-1], mhd=True, riemann=HLL)
    nc = native_config(state.shape, cfg, 5.0 / 3.0, 0.4)
    ctx = _native.FvContext(_native.AmxFvConfig.from_buffer_copy(nc))
    oc = oracle.make_config(state.shape, dim, C_cfl=0.4)
    s = torch.from_numpy(np.ascontiguousarray(state)).cuda()
    prev, regenerated = 4, []  # a fresh context queues launches 0..4 first
    for dt in [6e-3, 3e-4, 4e-3, 1e-7, 8e-3, 3e-4, 1e-3]:
        x = s.cpu().numpy()
        k = ctx.evolve_state(s, torch.tensor(dt, dtype=torch.float32, device="cuda"))
        torch.cuda.synchronize()
        ref, kref = oracle.evolve_state_fv(x, oc, np.float32(dt))
        assert k == kref, (dt, k, kref)
        err = rel_l1(s.cpu().numpy(), ref)
        assert err.max() <= TOL, (dt, k, err)
        if k % 2 == 0 and k != prev and k <= max(1, prev):
            regenerated.append(k)
        prev = k
    assert 0 in regenerated and 2 in regenerated, regenerated


def test_loop_without_regeneration_launches(gpu):
    """The device loop queues regeneration launches only in batches after
    one whose steps needed them; without them, a step that ends on a first
    iterate its pair did not store is held and finished by the host (the
    regeneration, then its second half). Loop calls of fixed dt on one
    context drive the final launch index 3, 3, 3 | 2, 2, 2 | 3, 3 | 4, 4 |
    2, 2: the first steps of the 2-calls take that route. The loops equal
    the same steps one at a time (which always queue the regeneration) bit
    for bit, with the same iteration counts."""
    import torch
    print("[gpu] loop without regeneration launches", flush=True)
    state = W.random_smooth((16, 16, 16), True, seed=8)
    cfg = W.fv_config(3, 16, mhd=True, riemann=HLL)
    nc = native_config(state.shape, cfg, 5.0 / 3.0, 0.4)
    loop = _native.FvContext(_native.AmxFvConfig.from_buffer_copy(nc))
    single = _native.FvContext(_native.AmxFvConfig.from_buffer_copy(nc))
    s_loop = torch.from_numpy(np.ascontiguousarray(state)).cuda()
    s_step = s_loop.clone()
    ks = []
    for dt, n in [(1e-3, 3), (3e-4, 3), (1e-3, 2), (4e-3, 2), (3e-4, 2)]:
        t_end = float(np.float32(dt) * np.float32(n))
        run = _native.AmxFvRun(t_end=t_end, fixed_timestep=1, num_timesteps=n, exact_end_time=0, max_steps=-1)
        st = loop.time_integration(s_loop, run)
        d = torch.tensor(np.float32(t_end) / np.float32(n), dtype=torch.float32, device="cuda")
        kk = [single.evolve_state(s_step, d) for _ in range(n)]
        torch.cuda.synchronize()
        assert st.steps == n and st.fixed_point_iterations == sum(kk), (dt, st.fixed_point_iterations, kk)
        assert torch.equal(s_loop, s_step), (dt, kk)
        ks += kk
    assert ks == [3, 3, 3, 2, 2, 2, 3, 3, 4, 4, 2, 2], ks
