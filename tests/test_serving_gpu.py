"""Continuous batching (amx_fv_pool_* and the serving engine) on the GPU.

Every slot of a pool runs the same kernels as the single-simulation path
with its own state, gamma, t_end, dt and fixed-point count, so each result
must equal the single-simulation result for the same request bit for bit,
whether it ran alone or in a mix, admitted at the start or mid-run.
"""

import numpy as np
import pytest

from astronomix_amd import _native
from astronomix_amd import workloads as W
from astronomix_amd.serving import Request, ServingEngine, initial_state, mixed_requests
from astronomix_amd.serving.engine import _pool_config, run_one_at_a_time

pytestmark = pytest.mark.gpu


def solo(req, state):
    import torch
    ctx = _native.fv_context(_pool_config(req.kind, req.n, req.gamma))
    s = torch.from_numpy(np.ascontiguousarray(state, dtype=np.float32)).cuda()
    run = _native.AmxFvRun(t_end=float(req.t_end), fixed_timestep=0, num_timesteps=0, exact_end_time=1,
                           max_steps=-1)
    st = ctx.time_integration(s, run)
    torch.cuda.synchronize()
    return s.cpu().numpy(), st


def step_until_done(pool, slots):
    done = {}
    while len(done) < len(slots):
        pool.step_launch()
        pool.step_finish()
        for slot in slots:
            if slot not in done and pool.status(slot).finished:
                done[slot] = pool.status(slot)
    return done


def test_pool_slots_equal_single_path_2d(gpu):
    import torch
    print("[gpu] pool 2D OT 64^2 x 3", flush=True)
    reqs = [Request(0, "orszag_tang_2d", 64, 1.4, 0.05, 0, 0), Request(1, "orszag_tang_2d", 64, 5 / 3, 0.08, 0, 0),
            Request(2, "orszag_tang_2d", 64, 1.55, 0.03, 0, 0)]
    pool = _native.FvPool(_pool_config("orszag_tang_2d", 64), 4)
    state = W.orszag_tang_2d(64)
    for i, r in enumerate(reqs):
        pool.load(i, state, r.gamma, r.t_end)
    st = step_until_done(pool, [0, 1, 2])
    for i, r in enumerate(reqs):
        out = pool.read(i, torch.empty(state.shape, dtype=torch.float32, device="cuda")).cpu().numpy()
        ref, rst = solo(r, state)
        assert st[i].steps == rst.steps
        assert abs(st[i].t - r.t_end) < 1e-6
        np.testing.assert_array_equal(out, ref)
    pool.close()


def test_pool_mid_run_admission_3d(gpu):
    """Capacity 2, three blasts: the third is loaded into the slot the first
    frees while the second keeps running."""
    import torch
    print("[gpu] pool 3D blast 24^3 mid-run admission", flush=True)
    reqs = [Request(0, "mhd_blast_3d", 24, 1.6, 0.002, 20.0, 0), Request(1, "mhd_blast_3d", 24, 1.45, 0.006, 8.0, 0),
            Request(2, "mhd_blast_3d", 24, 5 / 3, 0.004, 40.0, 0)]
    states = {r.rid: initial_state(r) for r in reqs}
    pool = _native.FvPool(_pool_config("mhd_blast_3d", 24), 2)
    pool.load(0, states[0], reqs[0].gamma, reqs[0].t_end)
    pool.load(1, states[1], reqs[1].gamma, reqs[1].t_end)
    owner = {0: 0, 1: 1}
    results = {}
    steps_taken = 0
    while len(results) < 3:
        pool.step_launch()
        pool.step_finish()
        steps_taken += 1
        for slot in (0, 1):
            rid = owner.get(slot)
            if rid is None:
                continue
            st = pool.status(slot)
            if st.finished:
                results[rid] = pool.read(slot, torch.empty(states[rid].shape, dtype=torch.float32,
                                                           device="cuda")).cpu().numpy()
                pool.release(slot)
                owner[slot] = None
                if 2 not in owner.values() and 2 not in results:
                    pool.load(slot, states[2], reqs[2].gamma, reqs[2].t_end)
                    owner[slot] = 2
        assert steps_taken < 10000
    for r in reqs:
        ref, _ = solo(r, states[r.rid])
        np.testing.assert_array_equal(results[r.rid], ref)
    pool.close()


def test_engine_mix_equals_alone(gpu):
    """A small mixed stream through the engine (2D and 3D pools overlapping
    on their own streams) against each request served alone."""
    print("[gpu] engine mixed stream", flush=True)
    grids = (("orszag_tang_2d", 64), ("orszag_tang_2d", 96), ("mhd_blast_3d", 24), ("mhd_blast_3d", 32))
    reqs = mixed_requests(10, rate=1000.0, seed=3, grids=grids, t_end_range=(0.002, 0.02))
    states = {r.rid: initial_state(r) for r in reqs}
    eng = ServingEngine(hbm_budget=2 << 30, slots_per_pool=3)
    res = eng.serve(reqs, states, realtime=False)
    eng.close()
    assert sorted(res) == [r.rid for r in reqs]
    alone = run_one_at_a_time(reqs, states, realtime=False)
    for r in reqs:
        np.testing.assert_array_equal(res[r.rid]["state"].cpu().numpy(), alone[r.rid]["state"].cpu().numpy())
        assert res[r.rid]["steps"] == alone[r.rid]["steps"]


def test_engine_admission_respects_hbm_budget(gpu):
    """A budget of one 32^3 slot: requests queue, run one at a time, all finish."""
    print("[gpu] engine budget", flush=True)
    cfg = _pool_config("mhd_blast_3d", 32)
    slot = _native.pool_slot_bytes(cfg)
    reqs = mixed_requests(3, rate=1000.0, seed=5, grids=(("mhd_blast_3d", 32),), t_end_range=(0.002, 0.004))
    eng = ServingEngine(hbm_budget=slot + slot // 2, slots_per_pool=8)
    peak = 0
    orig = eng.step

    def step():
        nonlocal peak
        peak = max(peak, eng.running())
        assert eng.used <= eng.budget
        return orig()

    eng.step = step
    res = eng.serve(reqs, realtime=False)
    eng.close()
    assert len(res) == 3 and peak == 1


def test_graph_replays_equal_direct_launches(gpu):
    """The captured step graphs (replayed many times, with slots loaded and
    released in between) give the same states as direct launches."""
    import torch
    print("[gpu] pool graphs vs direct", flush=True)
    reqs = [Request(0, "orszag_tang_2d", 48, 1.5, 0.03, 0, 0), Request(1, "orszag_tang_2d", 48, 1.62, 0.05, 0, 0),
            Request(2, "orszag_tang_2d", 48, 1.41, 0.04, 0, 0)]
    state = W.orszag_tang_2d(48)
    outs = []
    for graphs in (True, False):
        pool = _native.FvPool(_pool_config("orszag_tang_2d", 48), 2)
        pool.set_graphs(graphs)
        owner = {0: 0, 1: 1}
        pool.load(0, state, reqs[0].gamma, reqs[0].t_end)
        pool.load(1, state, reqs[1].gamma, reqs[1].t_end)
        res = {}
        while len(res) < 3:
            pool.step_launch()
            pool.step_finish()
            for slot in (0, 1):
                rid = owner.get(slot)
                if rid is not None and pool.status(slot).finished:
                    res[rid] = pool.read(slot, torch.empty(state.shape, dtype=torch.float32,
                                                           device="cuda")).cpu().numpy()
                    pool.release(slot)
                    owner[slot] = None
                    if 2 not in res and 2 not in owner.values():
                        pool.load(slot, state, reqs[2].gamma, reqs[2].t_end)
                        owner[slot] = 2
        pool.close()
        outs.append(res)
    for r in reqs:
        np.testing.assert_array_equal(outs[0][r.rid], outs[1][r.rid])
        ref, _ = solo(r, state)
        np.testing.assert_array_equal(outs[0][r.rid], ref)


def test_engine_serves_c4_grid_classes_at_the_oracle_bar(gpu, oracle):
    """A mixed stream drawn from the serving workload's own grid classes
    (Orszag-Tang 64^2 and 128^2, blast 64^3 with its p_in and gamma draws),
    served with continuous batching: every served result is checked against
    the CPU restatement at the parity bar (relative L1 <= 1e-5 per variable
    against the float32 oracle, or within 3x of the float32 oracle's own
    distance to float64 where that distance exceeds the bar), with the
    oracle's step and fixed-point counts. End times are short so that the
    oracle finishes in seconds."""
    print("[gpu] engine C4 classes vs oracle", flush=True)
    grids = (("orszag_tang_2d", 64), ("orszag_tang_2d", 128), ("mhd_blast_3d", 64))
    reqs = mixed_requests(6, rate=1000.0, seed=11, grids=grids, t_end_range=(0.004, 0.012))
    assert {r.grid_class for r in reqs} == set(grids)
    states = {r.rid: initial_state(r) for r in reqs}
    eng = ServingEngine(hbm_budget=4 << 30, slots_per_pool=4)
    res = eng.serve(reqs, states, realtime=False)
    eng.close()
    from astronomix_amd.serving.requests import C_CFL

    def rel(a, b):
        return np.array([np.abs(a[v].astype(np.float64) - b[v]).sum() / max(np.abs(b[v]).sum(), 1e-30)
                         for v in range(a.shape[0])])

    for r in reqs:
        dim = 2 if r.kind == "orszag_tang_2d" else 3
        s = states[r.rid]
        oc = oracle.make_config(s.shape, dim, mhd=True, gamma=r.gamma, C_cfl=C_CFL[r.kind], t_end=r.t_end)
        ref, steps, t, iters = oracle.time_integration(s.astype(np.float32), oc)
        out = res[r.rid]["state"].cpu().numpy()
        assert res[r.rid]["steps"] == steps, (r, res[r.rid]["steps"], steps)
        err = rel(out, ref)
        assert np.all(np.isfinite(out))
        if err.max() > 1e-5:
            r64, _, _, _ = oracle.time_integration(s.astype(np.float64), oc)
            floor, err64 = rel(ref, r64), rel(out, r64)
            for v in np.nonzero(err > 1e-5)[0]:
                assert floor[v] > 1e-5 / 3 and err64[v] <= 3 * floor[v], (r, v, err[v], floor[v], err64[v])


@pytest.mark.parametrize("batch", [5, 16])
def test_batched_pool_steps_equal_single_path(gpu, batch):
    """Pools that queue several whole steps per launch (the device decides
    each step's fixed point, one readback per batch): every slot equals the
    single-simulation path bit for bit, with slots that finish inside a batch
    and a batch size that does not divide the step counts."""
    import torch
    print(f"[gpu] batched pool 2D OT 64^2, {batch} steps per launch", flush=True)
    reqs = [Request(0, "orszag_tang_2d", 64, 1.4, 0.05, 0, 0), Request(1, "orszag_tang_2d", 64, 5 / 3, 0.08, 0, 0),
            Request(2, "orszag_tang_2d", 64, 1.55, 0.03, 0, 0)]
    pool = _native.FvPool(_pool_config("orszag_tang_2d", 64), 4)
    pool.set_batch(batch)
    state = W.orszag_tang_2d(64)
    for i, r in enumerate(reqs):
        pool.load(i, state, r.gamma, r.t_end)
    st = step_until_done(pool, [0, 1, 2])
    for i, r in enumerate(reqs):
        out = pool.read(i, torch.empty(state.shape, dtype=torch.float32, device="cuda")).cpu().numpy()
        ref, rst = solo(r, state)
        assert st[i].steps == rst.steps and st[i].fixed_point_iterations == rst.fixed_point_iterations
        np.testing.assert_array_equal(out, ref)
    pool.close()


def test_batched_pool_completes_stalled_steps(gpu):
    """At fixed-point tolerance 1e-6 the first steps need more launches than
    the batch queues: those slots stall, the rest of the batch does nothing
    for them, and step_finish completes their steps. Two 3D slots with
    different gamma and t_end equal their single-simulation loops (which
    stall and complete the same way) bit for bit."""
    import torch
    print("[gpu] batched pool with stalled fixed points", flush=True)
    state = W.random_smooth((24, 24, 24), True, seed=1, amp=1.0)
    cfg = _pool_config("mhd_blast_3d", 24)
    cfg.fixed_point_tol = 1e-6
    cfg.C_cfl = 0.8
    pool = _native.FvPool(cfg, 2)
    pool.set_batch(6)
    runs = [(5 / 3, 0.012), (1.5, 0.02)]
    for i, (g, te) in enumerate(runs):
        pool.load(i, state, g, te)
    st = step_until_done(pool, [0, 1])
    for i, (g, te) in enumerate(runs):
        out = pool.read(i, torch.empty(state.shape, dtype=torch.float32, device="cuda")).cpu().numpy()
        c1 = _native.AmxFvConfig.from_buffer_copy(cfg)
        c1.gamma = g
        ctx = _native.FvContext(c1)
        s = torch.from_numpy(np.ascontiguousarray(state)).cuda()
        rst = ctx.time_integration(s, _native.AmxFvRun(t_end=te, fixed_timestep=0, num_timesteps=0, exact_end_time=1,
                                                       max_steps=-1))
        torch.cuda.synchronize()
        assert st[i].steps == rst.steps and st[i].fixed_point_iterations == rst.fixed_point_iterations
        assert rst.max_fixed_point_iterations > 5
        np.testing.assert_array_equal(out, s.cpu().numpy())
    pool.close()
