"""Benchmark: 3D FV MHD step, 256^3 blast (HLL, minmod, SSP-RK2 half steps
around the implicit-midpoint magnetic sub-step, CFL 0.3, periodic, float32)
on one MI355X.

A "step" is one full _evolve_state_fv step including its CFL dt reduction.
The state stays resident in HBM for the whole timed region. Prints one JSON
line (rank 0 only; this build runs on one GPU).

    python bench.py [--steps 200] [--warmup 10] [--n 256] [--no-cpu-baseline] [--no-serving]
    python bench.py --workload mhd_turbulence_3d_fd [--steps 100]   # FD WENO-5 + CT
    python bench.py --workload serving_mix                          # the served mix alone
    python bench.py --workload adjoint_ot                           # MHD reverse mode, OT 256^2

The default line also carries the metric's second half, "served": solves/s and
p50 / p99 latency of the mixed Orszag-Tang / blast request stream served with
continuous batching (32 requests), beside the same stream one at a time.
Per-launch kernel times (the roofline's) come from HIP events in a separate
pass of steps after the timed region.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E datasheet peak
# Algorithmic bytes per cell of one launch (float32 fields read + written once):
#   k_mag   : one launch makes two fixed-point iterates; the first pair reads
#             B0, v0, rho (7 fields: the input iterate is B0, v0), later pairs
#             also the input iterate (6). A pair writes its second iterate (6)
#             and its first only where the previous step's final launch index
#             says that one is final (never in the blast's steady K = 3) ->
#             13 fields for the first pair, 19 for later ones
#   k_stage : mean over the four launches of a step
#             A1 (R5 W5) + A2 (R10 W5) + A1' (R8 W6: + the fixed pressure)
#             + A2' (R10 W5: stencil 5, Heun centre 5; the loop keeps B in the
#             fixed point's iterate sets, no per-step copy) = 54 fields -> 54 B
MAG_FIELDS_FIRST, MAG_FIELDS_LATER = 13.0, 19.0
STAGE_BYTES_PER_CELL = 54.0
STEP_FIELDS_STAGES = 54


def cpu_baseline(n: int, steps: int, start=None, first_step: int = 0):
    """Time the CPU restatement (oracle/fv_oracle.cpp, float32, OpenMP on all
    host cores of this process) on a bounded sample of the same workload:
    `steps` steps from `start` (the GPU run's state after its warm-up, so the
    sample is the first timed steps) or from the initial state."""
    import numpy as np
    from oracle import fv
    from astronomix_amd import workloads as W
    fv.build()
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    state = W.blast_3d(n) if start is None else np.ascontiguousarray(start, dtype=np.float32)
    oc = fv.make_config(state.shape, 3, C_cfl=0.3, t_end=1e9)
    t0 = time.perf_counter()
    _, done, _, iters = fv.time_integration(state, oc, max_steps=steps)
    el = time.perf_counter() - t0
    where = (f"steps {first_step + 1}-{first_step + done} of the benchmarked run (the first timed steps, "
             f"from the GPU state after the warm-up)" if start is not None else f"{done} steps from the initial state")
    return {
        "value": round(done * n ** 3 / el / 1e6, 4),
        "unit": "Mzone-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"mhd_blast_3d {n}^3, {where}, fixed-point iterations per step {iters / max(done, 1):.2f}; "
                  f"float32 C++/OpenMP restatement (repo CPU restatement, not reference JAX); "
                  f"cpu={platform.processor() or platform.machine()}",
        "seconds": round(el, 3),
    }


# FD (WENO-5 + CT + SSPRK(5,4)), per cell of one k_weno launch: reads
# rho, m (4) + E, B after the centre update (4), writes the 8 interface fluxes
# and the 2 CT modified fluxes -> 18 fields = 72 B.
WENO_BYTES_PER_CELL = 72.0
# Whole FD step, per cell (fields read + written once per pass):
#   init R11 W22; per stage: prep R12(+2 floors) W4(+2), 3 x weno 18, ct_edge R6 W3,
#   ct_avg R3 W3, face R9 W6 (final R6 W3), qupd R48 W16 (final R32 W8); finish R11 W11
FD_STEP_FIELDS = 33 + 4 * (20 + 54 + 9 + 6 + 15 + 64) + (16 + 54 + 9 + 6 + 9 + 40) + 22


def cpu_baseline_fd(n: int, steps: int):
    """Time the FD CPU restatement (oracle/fd_oracle.cpp, float32, OpenMP)
    on a bounded sample of the turbulence workload."""
    from oracle import fd
    from astronomix_amd import workloads as W
    fd._load()
    cores = len(os.sched_getaffinity(0))
    state = W.turbulence_3d(n, seed=0)
    oc = fd.make_config(state.shape, C_cfl=0.4, t_end=1e9, dt_max=0.001)
    t0 = time.perf_counter()
    _, done, _ = fd.time_integration(state, oc, max_steps=steps)
    el = time.perf_counter() - t0
    return {
        "value": round(done * n ** 3 / el / 1e6, 4),
        "unit": "Mzone-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"mhd_turbulence_3d_fd {n}^3, {done} steps from the initial state, float32 C++/OpenMP "
                  f"restatement (repo CPU restatement, not reference JAX); cpu={platform.processor() or platform.machine()}",
        "seconds": round(el, 3),
    }


def bench_fd(args):
    """FD WENO-5 + CT + SSPRK(5,4) on the 3D turbulence start (rho=1, p=1,
    Bx=0.1, solenoidal Mach-1 velocity, seed 0), CFL 0.4, dt_max 0.001."""
    import torch
    from astronomix_amd import _native
    from astronomix_amd import workloads as W
    from astronomix_amd._finite_difference import native_config

    n = args.n
    state_np, cfg, params = W.workload("mhd_turbulence_3d_fd", n)
    state = torch.from_numpy(state_np).cuda()
    ctx = _native.fd_context(native_config(state.shape, cfg, params))
    if args.warmup > 0:
        ctx.time_integration(state, _native.AmxFvRun(t_end=float(params.t_end), fixed_timestep=0, num_timesteps=0,
                                                     exact_end_time=1, max_steps=args.warmup))
    torch.cuda.synchronize()
    run = _native.AmxFvRun(t_end=float(params.t_end), fixed_timestep=0, num_timesteps=0, exact_end_time=1,
                           max_steps=args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = ctx.time_integration(state, run)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # per-launch WENO times from a separate event-timed pass after the timed region
    weno_ms, weno_n = 0.0, 0
    if not args.no_kernel_timing:
        ctx.set_timing(True)
        ctx.time_integration(state, _native.AmxFvRun(t_end=float(params.t_end), fixed_timestep=0, num_timesteps=0,
                                                     exact_end_time=1, max_steps=min(args.timing_steps, 10)))
        torch.cuda.synchronize()
        weno_ms, weno_n = ctx.get_timing()
        ctx.set_timing(False)
    if not bool(torch.isfinite(state).all()):
        print("non-finite state after the timed steps", file=sys.stderr)
        sys.exit(1)
    cells = n ** 3
    steps = st.steps
    avg_ms = weno_ms / max(weno_n, 1)
    achieved = WENO_BYTES_PER_CELL * cells / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    step_bytes = FD_STEP_FIELDS * 4.0
    traffic = None
    tpath = os.path.join(HERE, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(f"k_weno_{n}")
        except Exception:
            traffic = None
    out = {
        "metric": "Mzone-updates/s, 3D FD MHD (WENO-5 + CT) 256^3 fp32",
        "value": round(steps * cells / el / 1e6, 3),
        "unit": "Mzone-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / max(steps, 1) * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"mhd_turbulence_3d_fd_{n}: FD WENO-5 characteristic fluxes + HOW-MHD constrained transport, "
                        f"SSPRK(5,4), CFL 0.4, dt_max 0.001, periodic, rho=1, p=1, Bx=0.1, solenoidal Mach-1 "
                        f"velocity in shells |k| in [1,3], seed 0",
            "grid": [n, n, n],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_weno",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_cell_per_launch": WENO_BYTES_PER_CELL,
            "avg_launch_ms": round(avg_ms, 5),
        },
        "step_model": {
            "bytes_per_zone_update": step_bytes,
            "effective_GBs": round(step_bytes * cells * steps / el / 1e9, 1),
            "k_weno_avg_ms": round(avg_ms, 5),
            "k_weno_launches": weno_n,
            "k_weno_share_of_step": round(weno_ms / (el * 1e3), 4),
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_fd(min(n, 128), args.cpu_steps or 6)
    print(json.dumps(out), flush=True)


def bench_serving(args):
    """Continuous batching of the mixed request stream (Poisson 50/s, seed 0:
    Orszag-Tang 64^2..512^2 and blast 64^3..128^3, gamma ~ U[1.4, 5/3], blast
    p_in ~ U[5, 50], t_end ~ U[0.1, 0.5]) against the same stream served one
    simulation at a time; both keep states on the GPU and hand back GPU
    tensors. Pools and contexts are created before the clock starts."""
    res = serving_run(args)
    s_served, s_alone = res["serving"], res["one_at_a_time"]
    reqs = res["requests"]
    out = {
        "metric": "served solves/s (continuous batching) and p99 latency, mixed OT/blast stream",
        "value": s_served["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": res["steps"],
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"serving_mix: {len(reqs)} requests, Poisson 50/s seed 0; Orszag-Tang 64^2-512^2 (HLL, CFL 0.4) "
                        f"and MHD blast 64^3-128^3 (HLL, CFL 0.3), gamma~U[1.4,5/3], p_in~U[5,50], t_end~U[0.1,0.5]",
            "slots_per_pool": args.slots,
            "steps_per_launch": args.steps_per_launch,
            "hbm_budget_gib": args.hbm_budget_gib,
            "parallelism": "single GPU",
        },
        "serving": s_served,
        "one_at_a_time": s_alone,
        "speedup_solves_per_s": res["speedup_solves_per_s"],
        "p99_ratio_alone_over_served": res["p99_ratio_alone_over_served"],
        "max_abs_diff_served_vs_alone": res["max_abs_diff_served_vs_alone"],
        "zone_updates_per_s_served": res["zone_updates_per_s_served"],
        "pool_bytes": res["pool_bytes"],
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def serving_run(args):
    """The serving stream (args.requests requests) served with continuous
    batching and one simulation at a time; latency summaries of both."""
    import numpy as np
    import torch
    from astronomix_amd import _native
    from astronomix_amd.serving import ServingEngine, initial_state, latency_summary, mixed_requests
    from astronomix_amd.serving.engine import _pool_config, run_one_at_a_time
    from astronomix_amd.serving.requests import GRIDS

    reqs = mixed_requests(args.requests, rate=50.0, seed=0)
    states = {r.rid: initial_state(r) for r in reqs}
    eng = ServingEngine(hbm_budget=int(args.hbm_budget_gib * (1 << 30)), slots_per_pool=args.slots,
                        steps_per_launch=args.steps_per_launch)
    eng.reserve(sorted({r.grid_class for r in reqs}))
    for kind, n in GRIDS:
        _native.fv_context(_pool_config(kind, n))
    torch.cuda.synchronize()
    served = eng.serve(reqs, states, realtime=True)
    torch.cuda.synchronize()
    used = eng.used
    eng.close()
    alone = run_one_at_a_time(reqs, states, realtime=True)
    diff = max(float((served[r.rid]["state"] - alone[r.rid]["state"]).abs().max()) for r in reqs)
    s_served, s_alone = latency_summary(served), latency_summary(alone)
    zone_updates = sum(served[r.rid]["steps"] * int(np.prod(states[r.rid].shape[1:])) for r in reqs)
    steps = int(sum(served[r.rid]["steps"] for r in reqs))
    del served, alone
    _native.clear_context_cache()
    torch.cuda.empty_cache()
    return {
        "requests": reqs,
        "steps": steps,
        "serving": s_served,
        "one_at_a_time": s_alone,
        "speedup_solves_per_s": round(s_served["solves_per_s"] / s_alone["solves_per_s"], 3),
        "p99_ratio_alone_over_served": round(s_alone["p99_latency_s"] / s_served["p99_latency_s"], 3),
        "max_abs_diff_served_vs_alone": diff,
        "zone_updates_per_s_served": round(zone_updates / s_served["makespan_s"] / 1e6, 2),
        "pool_bytes": used,
    }


def bench_adjoint(args):
    """Gradient of L = sum (rho_final - target)^2 with respect to the initial
    v_y field of the 2D Kelvin-Helmholtz start (512^2, HLLC, minmod, SSP-RK2,
    1000 fixed steps to t = 0.2, checkpoint every 50 steps): forward run +
    checkpointed reverse sweep through the adjoint kernels."""
    import numpy as np
    import torch
    from astronomix_amd import workloads as W
    from astronomix_amd.adjoint import rollout
    from astronomix_amd.option_classes.simulation_config import HLLC
    from astronomix_amd.option_classes.simulation_params import SimulationParams

    n, steps = args.n if args.n != 256 else 512, args.steps if args.steps != 200 else 1000
    cfg = W.fv_config(2, n, mhd=False, riemann=HLLC)._replace(fixed_timestep=True, num_timesteps=steps)
    params = SimulationParams(C_cfl=0.4, t_end=0.2 * steps / 1000, gamma=5.0 / 3.0)
    s0 = torch.from_numpy(W.kelvin_helmholtz_2d(n, seed=0)).cuda()
    target = torch.from_numpy(np.random.default_rng(1).random((n, n)).astype(np.float32)).cuda()

    def grad_vy():
        vy = s0[2].clone().requires_grad_(True)
        w = torch.cat([s0[:2], vy[None], s0[3:]])
        out = rollout(w, cfg, params, checkpoint_every=50)
        loss = ((out[0].double() - target.double()) ** 2).sum()
        loss.backward()
        return loss.detach(), vy.grad

    if args.warmup > 0:
        grad_vy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss, g = grad_vy()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # the forward alone, for the cost split
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    with torch.no_grad():
        rollout(s0.clone(), cfg, params, checkpoint_every=50)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t1
    out = {
        "metric": "adjoint: zone-updates/s through forward + checkpointed reverse sweep, 2D KH 512^2 HLLC",
        "value": round(steps * n * n / el / 1e6, 3),
        "unit": "Mzone-updates/s (forward+reverse)",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"adjoint_kh_{n}: grad of sum((rho_T - target)^2) wrt initial v_y; KH rho=2/1, vx=+-0.5, p=2.5, "
                        f"vy=0.01 sin(4 pi x) + N(0,1e-3) seed 0; HLLC + minmod, SSP-RK2, {steps} fixed steps, "
                        f"checkpoint every 50",
            "grid": [n, n],
            "parallelism": "single GPU",
        },
        "gradient_seconds": round(el, 4),
        "forward_seconds": round(fwd, 4),
        "reverse_over_forward": round((el - fwd) / fwd, 3),
        "loss": float(loss),
        "grad_norm": float(g.double().norm()),
        "grad_finite": bool(torch.isfinite(g).all()),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_adjoint_mhd(args):
    """Gradient of L = sum (rho_final - target)^2 with respect to the whole
    initial state of the 2D Orszag-Tang vortex (256^2, HLL, minmod, split MHD
    step with the implicit-midpoint magnetic fixed point, 200 fixed steps to
    t = 0.1, checkpoint every 50 steps): forward run + checkpointed reverse
    sweep through the adjoint kernels, the fixed point's applications
    included."""
    import numpy as np
    import torch
    from astronomix_amd import workloads as W
    from astronomix_amd.adjoint import rollout
    from astronomix_amd.option_classes.simulation_config import HLL
    from astronomix_amd.option_classes.simulation_params import SimulationParams

    n, steps = args.n, args.steps
    cfg = W.fv_config(2, n, mhd=True, riemann=HLL)._replace(fixed_timestep=True, num_timesteps=steps)
    params = SimulationParams(C_cfl=0.4, t_end=0.1 * steps / 200, gamma=5.0 / 3.0)
    s0 = torch.from_numpy(W.orszag_tang_2d(n)).cuda()
    target = torch.from_numpy(np.random.default_rng(1).random((n, n)).astype(np.float32)).cuda()

    def grad_state():
        w = s0.clone().requires_grad_(True)
        out = rollout(w, cfg, params, checkpoint_every=50)
        loss = ((out[0].double() - target.double()) ** 2).sum()
        loss.backward()
        return loss.detach(), w.grad

    if args.warmup > 0:
        grad_state()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss, g = grad_state()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    with torch.no_grad():
        rollout(s0.clone(), cfg, params, checkpoint_every=50)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t1
    out = {
        "metric": "adjoint: zone-updates/s through forward + checkpointed reverse sweep, 2D Orszag-Tang MHD HLL",
        "value": round(steps * n * n / el / 1e6, 3),
        "unit": "Mzone-updates/s (forward+reverse)",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"adjoint_ot_{n}: grad of sum((rho_T - target)^2) wrt the initial state; Orszag-Tang; HLL + "
                        f"minmod, split MHD step with the implicit-midpoint fixed point, {steps} fixed steps to "
                        f"t = {0.1 * steps / 200:g}, checkpoint every 50",
            "grid": [n, n],
            "parallelism": "single GPU",
        },
        "gradient_seconds": round(el, 4),
        "forward_seconds": round(fwd, 4),
        "reverse_over_forward": round((el - fwd) / fwd, 3),
        "loss": float(loss),
        "grad_norm": float(g.double().norm()),
        "grad_finite": bool(torch.isfinite(g).all()),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def stream_copy_gbs(gib: float = 1.0, reps: int = 10):
    """HBM ceiling on this box: the library's streaming copy kernel on a
    buffer of `gib` GiB (read + write counted), timed with HIP events."""
    import torch
    from astronomix_amd import _native
    n = int(gib * (1 << 30) // 4)
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    _native.copy_f32(dst, src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        _native.copy_f32(dst, src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return {"GBs": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1), "bytes": 2 * n * 4, "avg_launch_ms": round(ms, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="CPU baseline sample length (default: 8 FV steps at 256^3, 6 FD steps at 128^3)")
    ap.add_argument("--copy-gib", type=float, default=1.0)
    ap.add_argument("--workload", default="mhd_blast_3d",
                    choices=["mhd_blast_3d", "mhd_turbulence_3d_fd", "serving_mix", "adjoint_kh", "adjoint_ot"])
    ap.add_argument("--requests", type=int, default=32, help="serving_mix: requests in the stream")
    ap.add_argument("--slots", type=int, default=16, help="serving_mix: slots per grid-class pool")
    ap.add_argument("--hbm-budget-gib", type=float, default=32.0, help="serving_mix: pool memory budget")
    ap.add_argument("--steps-per-launch", type=int, default=16,
                    help="serving_mix: whole steps a pool queues per launch (one readback per batch)")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="time the steps without per-launch HIP events (no roofline figures)")
    ap.add_argument("--timing-steps", type=int, default=20,
                    help="mhd_blast_3d: steps of the event-timed pass after the timed region")
    ap.add_argument("--no-serving", action="store_true",
                    help="mhd_blast_3d: leave out the served-mix figures of the line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "mhd_turbulence_3d_fd":
        return bench_fd(args)
    if args.workload == "serving_mix":
        return bench_serving(args)
    if args.workload == "adjoint_kh":
        return bench_adjoint(args)
    if args.workload == "adjoint_ot":
        return bench_adjoint_mhd(args)

    import torch
    from astronomix_amd import _native
    from astronomix_amd import workloads as W
    from astronomix_amd._finite_volume import native_config

    n = args.n
    state_np, cfg, params = W.workload("mhd_blast_3d", n)
    state = torch.from_numpy(state_np).cuda()
    ctx = _native.fv_context(native_config(state.shape, cfg, params.gamma, params.C_cfl))
    run_w = _native.AmxFvRun(t_end=float(params.t_end), fixed_timestep=0, num_timesteps=0, exact_end_time=1,
                             max_steps=args.warmup)
    if args.warmup > 0:
        ctx.time_integration(state, run_w)
    torch.cuda.synchronize()
    # the CPU baseline starts from the state the timed steps start from: kept
    # on the device here (a copy to the host just before the timed region
    # would leave the GPU idle for ~0.1 s and the timed steps would start at
    # lowered clocks) and copied to the host after the timed passes
    start_dev = state.clone() if not args.no_cpu_baseline else None
    run = _native.AmxFvRun(t_end=float(params.t_end), fixed_timestep=0, num_timesteps=0, exact_end_time=1,
                           max_steps=args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = ctx.time_integration(state, run)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    if not bool(torch.isfinite(state).all()):
        print("non-finite state after the timed steps", file=sys.stderr)
        sys.exit(1)
    # per-launch kernel times: HIP events around every launch of a separate
    # pass of steps that continues the run after the timed region (the timed
    # region carries no events)
    every = 0 if args.no_kernel_timing else 1
    tsteps = 0
    if every:
        ctx.set_timing(every)
        run_t = _native.AmxFvRun(t_end=float(params.t_end), fixed_timestep=0, num_timesteps=0, exact_end_time=1,
                                 max_steps=args.timing_steps)
        tst = ctx.time_integration(state, run_t)
        torch.cuda.synchronize()
        tsteps = tst.steps
    tm = ctx.get_timing()
    ctx.set_timing(0)
    start_np = None
    if start_dev is not None:
        start_np = start_dev.cpu().numpy()
        del start_dev

    cells = n ** 3
    steps = st.steps
    value = steps * cells / el / 1e6
    kbar = st.fixed_point_iterations / max(steps, 1)
    # executed pair launches per step (from the event records; without them,
    # the mean of floor(K / 2) + 1 is approximated by K / 2 + 1)
    pairs = tm.mag_launches / max(tsteps, 1) if tm.mag_launches else kbar / 2.0 + 1.0
    mag_fields = MAG_FIELDS_FIRST + MAG_FIELDS_LATER * max(pairs - 1.0, 0.0)
    step_bytes = (STEP_FIELDS_STAGES + mag_fields) * 4.0
    MAG_BYTES_PER_CELL = mag_fields * 4.0 / max(pairs, 1.0)
    mag_avg_ms = tm.mag_ms / max(tm.mag_launches, 1)
    stage_avg_ms = tm.stage_ms / max(tm.stage_launches, 1)
    if tm.mag_ms >= tm.stage_ms:
        dom, avg_ms, per_cell = "k_mag", mag_avg_ms, MAG_BYTES_PER_CELL
    else:
        dom, avg_ms, per_cell = "k_stage", stage_avg_ms, STAGE_BYTES_PER_CELL
    achieved = per_cell * cells / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = None
    tpath = os.path.join(HERE, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(f"{dom}_{n}")
        except Exception:
            traffic = None
    out = {
        "metric": "Mzone-updates/s, 3D FV MHD 256^3 fp32 (% HBM roofline)",
        "value": round(value, 3),
        "unit": "Mzone-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / max(steps, 1) * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"mhd_blast_3d_{n}: FV MHD, HLL + minmod, SSP-RK2 half steps around the implicit-midpoint "
                        f"magnetic fixed point, CFL 0.3, periodic, rho=1, p=10 (r<0.1) / 0.1, v=0, B=(1,1,0)/sqrt2",
            "grid": [n, n, n],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_cell_per_launch": per_cell,
            "avg_launch_ms": round(avg_ms, 5),
            "launch_timing": (f"HIP events around every launch of {tsteps} steps run right after the timed "
                              f"region (the timed region itself carries no events)" if every else None),
        },
        "step_model": {
            "fixed_point_iterations_per_step": round(kbar, 3),
            "fixed_point_pair_launches_per_step": round(pairs, 3),
            "bytes_per_zone_update": round(step_bytes, 1),
            "effective_GBs": round(step_bytes * cells * steps / el / 1e9, 1),
            "frac_of_peak": round(step_bytes * cells * steps / el / 1e9 / HBM_PEAK_GBS, 4),
            "k_mag_avg_ms": round(mag_avg_ms, 5),
            "k_mag_launches": tm.mag_launches,
            "k_stage_avg_ms": round(stage_avg_ms, 5),
            "k_stage_launches": tm.stage_launches,
            # the metric's own byte model: the reference dataflow's compulsory
            # traffic, one pass per RK stage or global sync, 264 + 76 K B per
            # zone-update in 3D (K fixed-point iterations per step)
            "metric_model": {
                "bytes_per_zone_update": round(264.0 + 76.0 * kbar, 1),
                "effective_GBs": round((264.0 + 76.0 * kbar) * cells * steps / el / 1e9, 1),
                "frac_of_peak": round((264.0 + 76.0 * kbar) * cells * steps / el / 1e9 / HBM_PEAK_GBS, 4),
            },
        },
        "stream_copy": None,
        "cpu_baseline": None,
    }
    if args.copy_gib > 0:
        out["stream_copy"] = stream_copy_gbs(args.copy_gib)
    if args.requests > 0 and not args.no_serving:
        # the metric's second half: the served mix's solves/s and latency
        del state
        _native.clear_context_cache()
        torch.cuda.empty_cache()
        res = serving_run(args)
        out["served"] = {
            "workload": f"serving_mix: {args.requests} requests, Poisson 50/s seed 0; Orszag-Tang 64^2-512^2 "
                        f"(HLL, CFL 0.4) and MHD blast 64^3-128^3 (HLL, CFL 0.3), gamma~U[1.4,5/3], "
                        f"p_in~U[5,50], t_end~U[0.1,0.5]; {args.slots} slots per grid class",
            "solves_per_s": res["serving"]["solves_per_s"],
            "p50_latency_s": res["serving"]["p50_latency_s"],
            "p99_latency_s": res["serving"]["p99_latency_s"],
            "one_at_a_time_solves_per_s": res["one_at_a_time"]["solves_per_s"],
            "one_at_a_time_p99_latency_s": res["one_at_a_time"]["p99_latency_s"],
            "speedup_solves_per_s": res["speedup_solves_per_s"],
            "max_abs_diff_served_vs_alone": res["max_abs_diff_served_vs_alone"],
        }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_steps or 8, start=start_np, first_step=args.warmup)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
