"""Continuous-batching engine over amx_fv_pool (see the package docstring)."""

from __future__ import annotations

import time
from collections import deque

import numpy as np

from astronomix_amd import _native
from astronomix_amd._finite_volume import native_config
from astronomix_amd.serving.requests import Request, initial_state, request_config


def _pool_config(kind, n, gamma=5.0 / 3.0):
    cfg, C = request_config(kind, n)
    shape = (4 + 3 if kind == "orszag_tang_2d" else 8,) + ((n, n) if kind == "orszag_tang_2d" else (n, n, n))
    return native_config(shape, cfg, gamma, C)


class ServingEngine:
    """Admits requests into per-grid pools as slots and HBM allow, steps all
    running slots, hands back finished states as GPU tensors.

    hbm_budget: device bytes the pools may hold (slots are allocated with
    their pool; a pool of a grid class is created on its first request with
    up to `slots_per_pool` slots that fit the remaining budget, and destroyed
    when it is empty and another class is waiting for memory).
    """

    def __init__(self, hbm_budget: int = 24 << 30, slots_per_pool: int = 16, steps_per_launch: int = 16):
        _native.require_gpu()
        self.budget = int(hbm_budget)
        self.slots_per_pool = int(slots_per_pool)
        # whole steps queued per pool launch (one host readback per batch);
        # slots that finish inside a batch wait for its end
        self.steps_per_launch = int(steps_per_launch)
        self.pools = {}      # grid class -> FvPool
        self.slot_req = {}   # grid class -> [Request | None] per slot
        self.pool_bytes = {}
        self.used = 0

    # ---- memory / admission
    def _slot_bytes(self, cls):
        return _native.pool_slot_bytes(_pool_config(*cls))

    def _make_pool(self, cls):
        sb = self._slot_bytes(cls)
        fit = (self.budget - self.used) // sb
        cap = int(min(self.slots_per_pool, fit))
        if cap < 1:
            return None
        pool = _native.FvPool(_pool_config(*cls), cap)
        if self.steps_per_launch > 1:
            pool.set_batch(self.steps_per_launch)
        self.pools[cls] = pool
        self.slot_req[cls] = [None] * cap
        self.pool_bytes[cls] = cap * sb
        self.used += cap * sb
        return pool

    def _drop_empty_pools(self, keep):
        for cls in list(self.pools):
            if cls != keep and all(r is None for r in self.slot_req[cls]):
                self.pools[cls].close()
                self.used -= self.pool_bytes.pop(cls)
                del self.pools[cls], self.slot_req[cls]

    def reserve(self, classes):
        """Creates the pools of these grid classes ahead of the stream."""
        for cls in classes:
            if cls not in self.pools:
                self._make_pool(cls)

    def try_admit(self, req: Request, state) -> bool:
        cls = req.grid_class
        pool = self.pools.get(cls)
        if pool is None:
            pool = self._make_pool(cls)
            if pool is None:
                self._drop_empty_pools(keep=cls)
                pool = self._make_pool(cls)
                if pool is None:
                    return False
        slots = self.slot_req[cls]
        try:
            slot = slots.index(None)
        except ValueError:
            return False
        pool.load(slot, state, req.gamma, req.t_end)
        slots[slot] = req
        return True

    def running(self) -> int:
        return sum(r is not None for slots in self.slot_req.values() for r in slots)

    # ---- stepping
    def step(self):
        """One launch (steps_per_launch steps) of every pool with running
        slots; returns the finished (request, state tensor) pairs and frees
        their slots."""
        import torch
        active = [cls for cls, slots in self.slot_req.items() if any(r is not None for r in slots)]
        for cls in active:
            self.pools[cls].step_launch()
        for cls in active:
            self.pools[cls].step_finish()
        done = []
        for cls in active:
            pool, slots = self.pools[cls], self.slot_req[cls]
            for slot, req in enumerate(slots):
                if req is None:
                    continue
                st = pool.status(slot)
                if st.finished:
                    shape = initial_state_shape(req)
                    out = torch.empty(shape, dtype=torch.float32, device="cuda")
                    pool.read(slot, out)
                    pool.release(slot)
                    slots[slot] = None
                    done.append((req, out, st))
        return done

    def close(self):
        for pool in self.pools.values():
            pool.close()
        self.pools.clear()
        self.slot_req.clear()
        self.used = 0

    # ---- a whole request stream
    def serve(self, requests, states=None, realtime=True):
        """Serves `requests` (arrival times in seconds from now). Returns a
        dict rid -> (result tensor, arrival, admitted, finished, steps)."""
        import torch
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting = deque()
        results = {}
        t0 = time.perf_counter()
        admitted_at = {}
        while pending or waiting or self.running():
            now = time.perf_counter() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                waiting.append(pending.popleft())
            # admission in arrival order; a request that does not fit waits
            still = deque()
            while waiting:
                req = waiting.popleft()
                state = states[req.rid] if states is not None else initial_state(req)
                if self.try_admit(req, state):
                    admitted_at[req.rid] = time.perf_counter() - t0
                else:
                    still.append(req)
            waiting = still
            if self.running():
                for req, out, st in self.step():
                    torch.cuda.synchronize()
                    results[req.rid] = dict(state=out, arrival=req.arrival, admitted=admitted_at[req.rid],
                                            finished=time.perf_counter() - t0, steps=st.steps, t=st.t)
            elif pending:
                time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.01)))
        return results


def initial_state_shape(req: Request):
    if req.kind == "orszag_tang_2d":
        return (7, req.n, req.n)
    return (8, req.n, req.n, req.n)


def run_one_at_a_time(requests, states=None, realtime=True):
    """The same stream served one simulation at a time through the
    single-simulation path (amx_fv_time_integration), in arrival order."""
    import torch
    results = {}
    t0 = time.perf_counter()
    for req in sorted(requests, key=lambda r: r.arrival):
        if realtime:
            wait = req.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        start = time.perf_counter() - t0
        cfg = _pool_config(req.kind, req.n, req.gamma)
        ctx = _native.fv_context(cfg)
        state = states[req.rid] if states is not None else initial_state(req)
        s = torch.from_numpy(np.ascontiguousarray(state, dtype=np.float32)).cuda()
        run = _native.AmxFvRun(t_end=float(req.t_end), fixed_timestep=0, num_timesteps=0, exact_end_time=1,
                               max_steps=-1)
        st = ctx.time_integration(s, run)
        torch.cuda.synchronize()
        results[req.rid] = dict(state=s, arrival=req.arrival, admitted=start, finished=time.perf_counter() - t0,
                                steps=st.steps, t=st.t)
    return results


def latency_summary(results):
    lat = np.array([r["finished"] - r["arrival"] for r in results.values()])
    makespan = max(r["finished"] for r in results.values())
    return {
        "requests": len(results),
        "makespan_s": round(float(makespan), 4),
        "solves_per_s": round(len(results) / float(makespan), 3),
        "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
        "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
        "mean_latency_s": round(float(lat.mean()), 4),
    }
