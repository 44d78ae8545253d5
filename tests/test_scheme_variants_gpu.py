"""Scheme variants on the GPU through the C ABI against the CPU restatement
(oracle/fv_oracle.cpp): the HLLC-LM, hybrid / AM HLLC and Lax-Friedrichs
fluxes and the Osher,
double-minmod, superbee, van Albada and positivity-preserving van Albada
limiters, in 2D and 3D, hydro and
MHD. Same bar as the default scheme: equal step and fixed-point iteration
counts, relative L1 per variable <= 1e-5 against the float32 oracle, or, for
a variable whose own float32 floor is above that, within 3x of that floor
against the float64 oracle."""

import numpy as np
import pytest

from astronomix_amd import _native
from astronomix_amd import workloads as W
from astronomix_amd._finite_volume import native_config
from astronomix_amd.option_classes import simulation_config as sc
from astronomix_amd.option_classes.simulation_params import SimulationParams

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel_l1(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.array([np.abs(a[v] - b[v]).sum() / max(np.abs(b[v]).sum(), 1e-30) for v in range(a.shape[0])])


def fast_flow(shape, mhd, vscale, seed):
    """Smooth density and pressure with velocities scaled up: the
    VAN_ALBADA_PP velocity factor beta drops below 1 while its density and
    pressure factors stay 1 (where those act, a float32 face lands on 0 and
    the Riemann solver returns NaN, in the reference as here)."""
    s = W.random_smooth(shape, mhd, seed=seed, amp=0.2).astype(np.float64)
    s[1:len(shape) + 1] *= vscale
    return s.astype(np.float32)


CASES = [
    # name, builder, dim, mhd, riemann, limiter, C_cfl, t_end
    ("ot64_vapp", lambda: W.orszag_tang_2d(64), 2, True, sc.HLL, sc.VAN_ALBADA_PP, 0.4, 0.05),
    ("fast2d_vapp", lambda: fast_flow((48, 80), False, 6.0, 2), 2, False, sc.HLL, sc.VAN_ALBADA_PP, 0.4, 0.01),
    ("fast2d_mhd_vapp_hllc", lambda: fast_flow((40, 70), True, 6.0, 2), 2, True, sc.HLLC, sc.VAN_ALBADA_PP, 0.4,
     0.01),
    ("fast3d_vapp_lf", lambda: fast_flow((12, 10, 72), False, 6.0, 2), 3, False, sc.LAX_FRIEDRICHS,
     sc.VAN_ALBADA_PP, 0.4, 0.005),
    ("fast3d_mhd_vapp", lambda: fast_flow((12, 10, 72), True, 2.5, 4), 3, True, sc.HLL, sc.VAN_ALBADA_PP, 0.3,
     0.004),
    ("ot64_hllclm", lambda: W.orszag_tang_2d(64), 2, True, sc.HLLC_LM, sc.MINMOD, 0.4, 0.05),
    ("ot64_lf", lambda: W.orszag_tang_2d(64), 2, True, sc.LAX_FRIEDRICHS, sc.MINMOD, 0.4, 0.05),
    ("ot64_osher", lambda: W.orszag_tang_2d(64), 2, True, sc.HLL, sc.OSHER, 0.4, 0.05),
    ("ot64_mc", lambda: W.orszag_tang_2d(64), 2, True, sc.HLL, sc.DOUBLE_MINMOD, 0.4, 0.05),
    ("ot64_superbee", lambda: W.orszag_tang_2d(64), 2, True, sc.HLLC, sc.SUPERBEE, 0.4, 0.05),
    ("ot64_vanalbada", lambda: W.orszag_tang_2d(64), 2, True, sc.HLL, sc.VAN_ALBADA, 0.4, 0.05),
    ("kh64_hllclm_mc", lambda: W.kelvin_helmholtz_2d(64), 2, False, sc.HLLC_LM, sc.DOUBLE_MINMOD, 0.4, 0.05),
    ("kh64_lf_osher", lambda: W.kelvin_helmholtz_2d(64), 2, False, sc.LAX_FRIEDRICHS, sc.OSHER, 0.4, 0.05),
    ("blast3d24_hllclm_superbee", lambda: W.blast_3d(24), 3, True, sc.HLLC_LM, sc.SUPERBEE, 0.3, 0.004),
    ("blast3d24_lf_vanalbada", lambda: W.blast_3d(24), 3, True, sc.LAX_FRIEDRICHS, sc.VAN_ALBADA, 0.3, 0.004),
    ("hydro3d20_lf", lambda: W.random_smooth((20, 20, 20), False, seed=4), 3, False, sc.LAX_FRIEDRICHS,
     sc.MINMOD, 0.4, 0.02),
    ("mhd3d_ragged_hllclm_osher", lambda: W.random_smooth((12, 10, 72), True, seed=6), 3, True, sc.HLLC_LM,
     sc.OSHER, 0.4, 0.01),
    ("blast3d32_am", lambda: W.blast_3d(32), 3, True, sc.AM_HLLC, sc.MINMOD, 0.3, 0.005),
    ("blast3d24_hybrid_mc", lambda: W.blast_3d(24), 3, True, sc.HYBRID_HLLC, sc.DOUBLE_MINMOD, 0.3, 0.004),
    ("ot64_am_superbee", lambda: W.orszag_tang_2d(64), 2, True, sc.AM_HLLC, sc.SUPERBEE, 0.4, 0.05),
    ("kh64_hybrid", lambda: W.kelvin_helmholtz_2d(64), 2, False, sc.HYBRID_HLLC, sc.MINMOD, 0.4, 0.05),
    ("hydro3d_ragged_am", lambda: W.random_smooth((12, 10, 72), False, seed=9, amp=0.6), 3, False, sc.AM_HLLC,
     sc.MINMOD, 0.4, 0.01),
]


def gpu_run(state, cfg, params):
    import torch
    s = torch.from_numpy(np.ascontiguousarray(state, dtype=np.float32)).cuda()
    ctx = _native.fv_context(native_config(s.shape, cfg, params.gamma, params.C_cfl))
    run = _native.AmxFvRun(t_end=float(params.t_end), fixed_timestep=0, num_timesteps=0, exact_end_time=1,
                           max_steps=-1)
    st = ctx.time_integration(s, run)
    torch.cuda.synchronize()
    return s.cpu().numpy(), st


def oracle_run(oracle, state, cfg, params, dtype):
    oc = oracle.make_config(state.shape, cfg.dimensionality, mhd=cfg.mhd, riemann=cfg.riemann_solver,
                            limiter=cfg.limiter, gamma=params.gamma, C_cfl=params.C_cfl, t_end=params.t_end,
                            box_size=cfg.box_size)
    return oracle.time_integration(state.astype(dtype), oc)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_variant_matches_oracle(gpu, oracle, case):
    name, build, dim, mhd, rs, lim, C, t_end = case
    print(f"[gpu] scheme variant {name}", flush=True)
    state = build()
    shape = state.shape[1:]
    # ragged grids take their sizes from the state; the spacing is box / (last axis), as in the oracle
    cfg = W.fv_config(dim, shape[-1], mhd=mhd, riemann=rs)._replace(limiter=lim)
    params = SimulationParams(C_cfl=C, t_end=t_end, gamma=5.0 / 3.0)
    out, st = gpu_run(state, cfg, params)
    ref, steps, t, iters = oracle_run(oracle, state, cfg, params, np.float32)
    assert st.steps == steps
    assert abs(st.t - t) <= 1e-6 * max(1.0, abs(t))
    if mhd:
        assert st.fixed_point_iterations == iters
    assert np.all(np.isfinite(out))
    err = rel_l1(out, ref)
    if err.max() > TOL:
        r64, _, _, _ = oracle_run(oracle, state, cfg, params, np.float64)
        floor = rel_l1(ref, r64)
        err64 = rel_l1(out, r64)
        for v in range(len(err)):
            if err[v] > TOL:
                assert floor[v] > TOL / 3, f"{name}: var {v} rel L1 {err[v]:.3e} (float32 floor {floor[v]:.3e})"
                assert err64[v] <= 3 * floor[v], f"{name}: var {v} {err64[v]:.3e} vs floor {floor[v]:.3e}"


def test_variants_differ_from_the_default(gpu):
    """The variant switches reach the kernels: each scheme gives a different
    result from HLL + minmod on the same input."""
    state = W.orszag_tang_2d(64)
    params = SimulationParams(C_cfl=0.4, t_end=0.02, gamma=5.0 / 3.0)
    base, _ = gpu_run(state, W.fv_config(2, 64, mhd=True, riemann=sc.HLL), params)
    for rs, lim in ((sc.HLLC_LM, sc.MINMOD), (sc.LAX_FRIEDRICHS, sc.MINMOD), (sc.HLL, sc.SUPERBEE),
                    (sc.HLL, sc.OSHER), (sc.HLL, sc.DOUBLE_MINMOD), (sc.HLL, sc.VAN_ALBADA),
                    (sc.HYBRID_HLLC, sc.MINMOD), (sc.AM_HLLC, sc.MINMOD)):
        out, _ = gpu_run(state, W.fv_config(2, 64, mhd=True, riemann=rs)._replace(limiter=lim), params)
        assert np.abs(out - base).max() > 1e-6, (rs, lim)


def test_pp_scaling_reaches_the_kernels(gpu, oracle):
    """On a flow whose velocity factor acts, VAN_ALBADA_PP differs from plain
    van Albada on the GPU by what the oracle says it should."""
    state = fast_flow((48, 80), False, 6.0, 2)
    params = SimulationParams(C_cfl=0.4, t_end=0.01, gamma=5.0 / 3.0)
    cfg = W.fv_config(2, 80, mhd=False, riemann=sc.HLL)
    pp, _ = gpu_run(state, cfg._replace(limiter=sc.VAN_ALBADA_PP), params)
    va, _ = gpu_run(state, cfg._replace(limiter=sc.VAN_ALBADA), params)
    got = rel_l1(pp, va).max()
    ref_pp, *_ = oracle_run(oracle, state, cfg._replace(limiter=sc.VAN_ALBADA_PP), params, np.float32)
    ref_va, *_ = oracle_run(oracle, state, cfg._replace(limiter=sc.VAN_ALBADA), params, np.float32)
    want = rel_l1(ref_pp, ref_va).max()
    assert want > 1e-3 and abs(got - want) <= 0.05 * want


def test_adjoint_refuses_variants(gpu):
    import torch
    from astronomix_amd.adjoint import rollout
    state, cfg, params = W.workload("kelvin_helmholtz_2d", 32)
    s = torch.from_numpy(state).cuda().requires_grad_(True)
    with pytest.raises(NotImplementedError):
        rollout(s, cfg._replace(limiter=sc.SUPERBEE), params._replace(t_end=0.01))


def test_am_switch_fires_in_the_blast(gpu, oracle):
    """The blast's expanding shell compresses the gas ahead of it: the
    compression switch is set in some cells there, so the AM-HLLC run
    differs from plain HLLC with the same dissipation control off."""
    state = W.blast_3d(32)
    oc = oracle.make_config(state.shape, 3, mhd=True, riemann=sc.AM_HLLC)
    params = SimulationParams(C_cfl=0.3, t_end=0.005, gamma=5.0 / 3.0)
    out, _ = gpu_run(state, W.fv_config(3, 32, mhd=True, riemann=sc.HLLC), params)
    sw = oracle.am_switch(np.ascontiguousarray(out[:5]), oc)
    assert 0 < sw.sum() < sw.size


@pytest.mark.parametrize("rs,lim", [(sc.LAX_FRIEDRICHS, sc.OSHER), (sc.AM_HLLC, sc.MINMOD),
                                    (sc.HLL, sc.VAN_ALBADA_PP)])
def test_pool_slots_equal_single_runs_with_variants(gpu, rs, lim):
    """The slot-batched pool, which replays captured HIP graphs of its step
    halves, runs the variants' pre-pass kernels (Lax-Friedrichs maxima,
    compression switch, positivity factors) per slot: every slot equals its
    own single-simulation run bit for bit, over tens of replays."""
    import torch
    print(f"[gpu] pool variants rs={rs} lim={lim}", flush=True)
    states = [fast_flow((40, 70), True, 3.0, 2), W.random_smooth((40, 70), True, seed=5, amp=0.3)]
    cfg = W.fv_config(2, 70, mhd=True, riemann=rs)._replace(limiter=lim)
    params = SimulationParams(C_cfl=0.4, t_end=0.01, gamma=5.0 / 3.0)
    singles = [gpu_run(s, cfg, params) for s in states]
    pool = _native.FvPool(native_config(states[0].shape, cfg, params.gamma, params.C_cfl), 2)
    try:
        for i, s in enumerate(states):
            pool.load(i, s, params.gamma, params.t_end)
        done = {}
        while len(done) < 2:
            pool.step_launch()
            pool.step_finish()
            for i in range(2):
                if i not in done and pool.status(i).finished:
                    done[i] = pool.status(i)
        for i, (ref, st) in enumerate(singles):
            out = pool.read(i, torch.empty(states[i].shape, dtype=torch.float32, device="cuda")).cpu().numpy()
            assert done[i].steps == st.steps > 5
            np.testing.assert_array_equal(out, ref)
    finally:
        pool.close()
