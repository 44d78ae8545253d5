"""``time_integration`` -- the public entry point of the reference
(``astronomix/time_stepping/time_integration.py:61-236``) on the HIP path.

Accepts the reference's arguments and returns the final primitive state in
the same layout. The state may be a numpy array (copied to the GPU and back)
or a float32 torch tensor on the GPU (returned as a new tensor, or updated in
place when ``config.donate_state``). The step loop runs on the device: dt is
reduced, clamped to ``t_end - t`` and consumed on the GPU
(time_integration.py:629-801). With ``config.return_snapshots`` the return
is ``SnapshotData``; with ``config.activate_snapshot_callback`` the
``snapshot_callable(time, state, registered_variables)`` is called at the
snapshot times (time_integration.py:539-610).
"""

from __future__ import annotations

import numpy as np

from astronomix_amd import _native
from astronomix_amd._finite_volume import native_config, require_supported
from astronomix_amd.option_classes.simulation_config import FINITE_DIFFERENCE
from astronomix_amd.time_stepping._snapshots import exact_end_time, run_loop, to_host, wants_snapshots


def time_integration(primitive_state, config, params, registered_variables, snapshot_callable=None,
                     sharding=None):
    import torch

    if sharding is not None:
        raise NotImplementedError("sharded states are not supported: one GPU per simulation")
    if config.solver_mode == FINITE_DIFFERENCE:
        from astronomix_amd._finite_difference import time_integration_fd
        return time_integration_fd(primitive_state, config, params, registered_variables, snapshot_callable)
    require_supported(config, registered_variables)
    if isinstance(primitive_state, torch.Tensor) and primitive_state.requires_grad:
        if wants_snapshots(config):
            raise NotImplementedError("snapshots are not implemented for the differentiable loop")
        # differentiation through the loop (differentiation_mode=BACKWARDS in
        # the reference): reverse-mode kernels with checkpointed recomputation
        from astronomix_amd.adjoint import rollout
        every = max(1, -(-int(config.num_timesteps) // max(1, int(config.num_checkpoints)))) \
            if config.fixed_timestep else 50
        return rollout(primitive_state, config, params, registered_variables, checkpoint_every=every)

    struct = None
    if config.state_struct:
        struct = primitive_state
        primitive_state = struct.primitive_state

    _native.require_gpu()
    host_numpy = not isinstance(primitive_state, torch.Tensor)
    if host_numpy:
        arr = np.ascontiguousarray(np.asarray(primitive_state, dtype=np.float32))
        state = torch.from_numpy(arr).to("cuda")
    else:
        if primitive_state.device.type != "cuda" or primitive_state.dtype != torch.float32:
            raise TypeError("a tensor state must be float32 on the GPU")
        state = primitive_state if (config.donate_state and primitive_state.is_contiguous()) \
            else primitive_state.contiguous().clone()

    ctx = _native.fv_context(native_config(state.shape, config, params.gamma, params.C_cfl))
    run = _native.AmxFvRun(t_end=float(params.t_end), fixed_timestep=int(bool(config.fixed_timestep)),
                           num_timesteps=int(config.num_timesteps), exact_end_time=int(exact_end_time(config)),
                           max_steps=-1)
    if config.memory_analysis:
        import ctypes
        nbytes = ctypes.c_size_t(0)
        ctx.lib.amx_fv_workspace_bytes(ctypes.byref(ctx.cfg), ctypes.byref(nbytes))
        print("=== device memory PER DEVICE ===")
        print(f"State size: {state.numel() * 4 / 1024**2:.2f} MB")
        print(f"Workspace size: {nbytes.value / 1024**2:.2f} MB")
    if config.print_elapsed_time:
        print("Starting simulation...")
    stats, snaps, seconds = run_loop(ctx, state, run, config, params, registered_variables, snapshot_callable)
    if config.print_elapsed_time:
        print("Simulation finished!")
        print(f"Time elapsed: {seconds:.2f} seconds")
        print(f"Number of iterations: {stats.steps}")
    if config.runtime_debugging:
        gas = state[: registered_variables.pressure_index + 1]
        if not torch.isfinite(state).all():
            raise FloatingPointError("non-finite values in the final state")
        if (gas[registered_variables.pressure_index] < 0).any() or (gas[registered_variables.density_index] < 0).any():
            raise FloatingPointError("negative density or pressure in the final state")
    if snaps is not None:
        return to_host(snaps) if host_numpy else snaps
    out = state.cpu().numpy() if host_numpy else state
    if struct is not None:
        return struct._replace(primitive_state=out)
    return out

