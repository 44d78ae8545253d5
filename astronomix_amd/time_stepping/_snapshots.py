"""Snapshot output of ``time_integration``: the ``SnapshotData`` return of
``config.return_snapshots`` and the user callback of
``config.activate_snapshot_callback`` (``time_stepping/time_integration.py:297-369``
array set-up, ``:404-610`` rows and schedule, ``:671-678`` dt clamps,
``:736-745`` end-time row, ``:811-833`` return).

The schedule runs inside the native step loop
(``amx_fv_time_integration_snapshots`` / ``amx_fd_time_integration_snapshots``);
each row (time point, state copy, totals, max |div B|) is written by the
device into preallocated arrays, so the loop never copies a state to the host.
``run_loop`` also installs the turbulent-forcing step hook when the
configuration asks for it (time_integration.py:694-704).
"""

from __future__ import annotations

import time

from astronomix_amd import _native
from astronomix_amd.data_classes.simulation_snapshot_data import SnapshotData


def wants_snapshots(config) -> bool:
    return bool(config.return_snapshots or config.activate_snapshot_callback)


def exact_end_time(config) -> bool:
    """The end-time clamp is off when specific snapshot time points are used
    (time_integration.py:677), whether or not snapshots are returned."""
    return bool(config.exact_end_time) and not config.use_specific_snapshot_timepoints


def _timepoints(params):
    tp = params.snapshot_timepoints
    if hasattr(tp, "detach"):
        tp = tp.detach().cpu().numpy()
    vals = [float(v) for v in (tp.reshape(-1) if hasattr(tp, "reshape") else tp)]
    if not vals:
        raise ValueError("params.snapshot_timepoints is empty")
    return vals


def allocate(config, state):
    """Device arrays of the requested SnapshotData fields (zeros, as the
    reference's jnp.zeros set-up, time_integration.py:297-343)."""
    import torch
    n = int(config.num_snapshots)
    if n < 1:
        raise ValueError("num_snapshots must be positive")
    ss = config.snapshot_settings

    def z(*shape):
        return torch.zeros(shape or (n,), dtype=torch.float32, device=state.device)

    arrays = {"time_points": z()}
    if ss.return_states:
        arrays["states"] = z(n, *state.shape)
    for flag, name in ((ss.return_total_mass, "total_mass"), (ss.return_total_energy, "total_energy"),
                       (ss.return_internal_energy, "internal_energy"), (ss.return_kinetic_energy, "kinetic_energy"),
                       (ss.return_radial_momentum, "radial_momentum")):
        if flag:
            arrays[name] = z()
    if ss.return_magnetic_divergence and config.mhd:
        arrays["magnetic_divergence"] = z()
    return arrays


def _forcing(config, params, state):
    """The turbulent-forcing hook of a run, or None (time_integration.py:694-704)."""
    if not config.turbulent_forcing_config.turbulent_forcing:
        return None
    from astronomix_amd._physics_modules.turbulent_forcing import TurbulentForcing
    shape = tuple(state.shape[1:])
    if len(shape) != 3 or len(set(shape)) != 1:
        raise NotImplementedError("turbulent forcing needs a cubic 3D grid (the reference's num_cells^3 box)")
    return TurbulentForcing(shape, float(config.box_size), float(config.grid_spacing),
                            float(params.turbulent_forcing_params.energy_injection_rate),
                            rho_row=0, vel_row=1, num_vars=int(state.shape[0]), device=state.device)


def run_loop(ctx, state, run, config, params, registered_variables, snapshot_callable=None, kind="fv"):
    """Runs the native loop of ctx on `state` in place. Returns
    (stats, SnapshotData with device tensors or None, seconds)."""
    forcing = _forcing(config, params, state)
    if forcing is None:
        return _run_loop(ctx, state, run, config, params, registered_variables, snapshot_callable)
    forcing.install(ctx, kind)
    try:
        out = _run_loop(ctx, state, run, config, params, registered_variables, snapshot_callable)
    except _native.AmxError:
        if forcing.error is not None:
            raise forcing.error
        raise
    finally:
        forcing.uninstall(ctx, kind)
    return out


def _run_loop(ctx, state, run, config, params, registered_variables, snapshot_callable=None):
    import torch
    recording = bool(config.return_snapshots)
    callback = None
    if not recording and config.activate_snapshot_callback:
        if snapshot_callable is None:
            raise TypeError("activate_snapshot_callback needs a snapshot_callable")

        def callback(t, checkpoint, _ptr):
            # the callable sees the live device state (read only), as the
            # reference passes the traced state (time_integration.py:595)
            snapshot_callable(t, state, registered_variables)

    t0 = time.perf_counter()
    if not wants_snapshots(config):
        stats = ctx.time_integration(state, run)
        torch.cuda.synchronize()
        return stats, None, time.perf_counter() - t0
    arrays = allocate(config, state) if recording else {}
    timepoints = _timepoints(params) if (recording and config.use_specific_snapshot_timepoints) else None
    srun = _native.SnapshotRun(config.num_snapshots, float(config.box_size), arrays, timepoints, callback)
    stats = ctx.time_integration_snapshots(state, run, srun)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    if not recording:
        return stats, None, seconds
    data = SnapshotData(
        time_points=arrays["time_points"],
        states=arrays.get("states"),
        final_state=state if config.snapshot_settings.return_final_state else None,
        total_mass=arrays.get("total_mass"),
        total_energy=arrays.get("total_energy"),
        internal_energy=arrays.get("internal_energy"),
        kinetic_energy=arrays.get("kinetic_energy"),
        gravitational_energy=None,
        radial_momentum=arrays.get("radial_momentum"),
        magnetic_divergence=arrays.get("magnetic_divergence"),
        runtime=seconds,
        num_iterations=srun.num_iterations,
        current_checkpoint=srun.current_checkpoint,
    )
    return stats, data, seconds


def to_host(data: SnapshotData) -> SnapshotData:
    """SnapshotData with numpy arrays in place of device tensors."""
    return data._replace(**{k: v.cpu().numpy() for k, v in data._asdict().items() if hasattr(v, "cpu")})
