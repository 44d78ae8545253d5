"""Finite-difference operators of the reference, backed by the HIP kernels.

``_cfl_time_step_fd``, ``_evolve_state_fd``, ``initialize_interface_fields``
and the FD branch of ``time_integration`` keep the reference's names and
argument meaning:

* ``_cfl_time_step_fd``  -- ``_finite_difference/_timestep_estimation/_timestep_estimator.py:26-91``
* ``_evolve_state_fd``   -- ``_finite_difference/_state_evolution/_evolve_state.py:25-72``
* ``initialize_interface_fields`` -- ``_finite_difference/_magnetic_update/_constrained_transport.py:201-213``
* ``time_integration_fd`` -- ``time_stepping/time_integration.py:629-801`` (FD branch)

States are float32 torch tensors on the GPU, (11, nx, ny, nz) =
[rho, vx, vy, vz, p, Bx, By, Bz, bx, by, bz] with the face fields last
(``variable_registry/registered_variables.py:188-198``).
"""

from __future__ import annotations


import numpy as np

from astronomix_amd import _native
from astronomix_amd.option_classes.simulation_config import (
    ALL_PERIODIC,
    BoundarySettings,
    CARTESIAN,
    DOUBLE_PRECISION,
    FINITE_DIFFERENCE,
)


def unsupported_reason(config, registered_variables=None):
    """None when ``config`` lies on the implemented FD path, else why not.

    The reference's FD mode is itself restricted to 3D periodic MHD with
    SSPRK(5,4) (simulation_config.py finalize_config); on top of that this
    build runs float32 without physics modules.
    """
    c = config
    if c.solver_mode != FINITE_DIFFERENCE:
        return "solver_mode must be FINITE_DIFFERENCE for this operator"
    if not c.mhd:
        return "finite difference mode is MHD only"
    if c.dimensionality != 3:
        return "finite difference mode is 3D only"
    if c.numerical_precision == DOUBLE_PRECISION:
        return "only SINGLE_PRECISION (float32) is implemented"
    if c.geometry != CARTESIAN:
        return "only CARTESIAN geometry is implemented"
    bs = c.boundary_settings
    periodic = bs == ALL_PERIODIC or (isinstance(bs, BoundarySettings) and all(
        (getattr(bs, ax).left_boundary, getattr(bs, ax).right_boundary) == (2, 2) for ax in ("x", "y", "z")))
    if not periodic:
        return "finite difference mode needs periodic boundaries on every axis"
    if c.self_gravity:
        return "self-gravity is not implemented"
    if c.turbulent_forcing_config.turbulent_forcing and c.dimensionality != 3:
        return "turbulent forcing is 3D only (a cubic box)"
    for name, flag in (("stellar wind", c.wind_config.stellar_wind),
                       ("cosmic rays", c.cosmic_ray_config.cosmic_rays),
                       ("cooling", c.cooling_config.cooling),
                       ("neural net force", c.neural_net_force_config.neural_net_force),
                       ("CNN MHD corrector", c.cnn_mhd_corrector_config.cnn_mhd_corrector)):
        if flag:
            return f"physics module '{name}' is not implemented"
    if registered_variables is not None and registered_variables.num_vars != 11:
        return "the FD state has exactly 11 registered variables (no tracers)"
    return None


def require_supported(config, registered_variables=None):
    why = unsupported_reason(config, registered_variables)
    if why is not None:
        raise NotImplementedError(why)


def native_config(state_shape, config, params, gamma=None, C_cfl=None, dt_max=None,
                  grid_spacing=None) -> _native.AmxFdConfig:
    if len(state_shape) != 4 or state_shape[0] != 11:
        raise ValueError(f"FD state must be (11, nx, ny, nz), got {tuple(state_shape)}")
    nx, ny, nz = (int(v) for v in state_shape[1:])
    return _native.AmxFdConfig(
        nx=nx, ny=ny, nz=nz, enforce_positivity=int(bool(config.enforce_positivity)),
        gamma=float(params.gamma if gamma is None else gamma),
        C_cfl=float(params.C_cfl if C_cfl is None else C_cfl),
        dt_max=float(params.dt_max if dt_max is None else dt_max),
        minimum_density=float(params.minimum_density), minimum_pressure=float(params.minimum_pressure),
        pad_=0.0, grid_spacing=float(config.grid_spacing if grid_spacing is None else grid_spacing))


def _as_device_state(state):
    import torch
    _native.require_gpu()
    if not isinstance(state, torch.Tensor):
        raise TypeError("state must be a torch tensor on the GPU")
    if state.dtype != torch.float32 or state.device.type != "cuda":
        raise TypeError("state must be a float32 tensor on the GPU")
    return state.contiguous()


def _cfl_time_step_fd(primitive_state, grid_spacing, dt_max, gamma, config, params, registered_variables,
                      C_CFL=0.8):
    """min(C dx / (max|l_x| + max|l_y| + max|l_z|), dt_max) as a 0-d device tensor."""
    import torch
    require_supported(config)
    s = _as_device_state(primitive_state)
    ctx = _native.fd_context(native_config(s.shape, config, params, gamma=gamma, C_cfl=C_CFL, dt_max=dt_max,
                                           grid_spacing=grid_spacing))
    out = torch.empty((), dtype=torch.float32, device=s.device)
    ctx.cfl_dt(s, out)
    return out


def _evolve_state_fd(primitive_state, dt, gamma, gravitational_constant, config, params, helper_data,
                     registered_variables):
    """One SSPRK(5,4) step with constrained transport; returns the new state
    (the input is left unchanged)."""
    import torch
    require_supported(config, registered_variables)
    s = _as_device_state(primitive_state).clone()
    ctx = _native.fd_context(native_config(s.shape, config, params, gamma=gamma))
    if not isinstance(dt, torch.Tensor):
        dt = torch.tensor(float(dt), dtype=torch.float32, device=s.device)
    dt = dt.to(device=s.device, dtype=torch.float32).reshape(())
    ctx.evolve_state(s, dt)
    return s


def initialize_interface_fields(magnetic_field_x, magnetic_field_y, magnetic_field_z):
    """Face fields (bx, by, bz) at i+1/2 from cell-centred B by the 4th-order
    centre-to-face interpolation along each field's own axis. Accepts numpy
    arrays (returns numpy) or float32 GPU tensors (returns tensors)."""
    import ctypes
    import torch
    _native.require_gpu()
    host = not isinstance(magnetic_field_x, torch.Tensor)
    comps = [torch.as_tensor(np.asarray(b, dtype=np.float32)) if host else b
             for b in (magnetic_field_x, magnetic_field_y, magnetic_field_z)]
    shape = tuple(comps[0].shape)
    if len(shape) != 3 or any(tuple(b.shape) != shape for b in comps):
        raise ValueError("three 3D fields of one shape are required")
    B = torch.stack([b.to(device="cuda", dtype=torch.float32) for b in comps]).contiguous()
    out = torch.empty_like(B)
    lib = _native.load_library()
    stream = torch.cuda.current_stream()
    _native.check(lib.amx_fd_interface_fields(ctypes.c_void_p(B.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                              shape[0], shape[1], shape[2], ctypes.c_void_p(stream.cuda_stream)),
                  "amx_fd_interface_fields")
    if host:
        o = out.cpu().numpy()
        return o[0], o[1], o[2]
    return out[0], out[1], out[2]


def time_integration_fd(primitive_state, config, params, registered_variables, snapshot_callable=None):
    """FD branch of time_integration: device-resident SSPRK(5,4) loop with
    the CFL step (clamped to params.dt_max) or t_end / num_timesteps, and
    the exact-end clamp; SnapshotData / snapshot callbacks as in the FV path
    (max |div B| over the face fields)."""
    import torch
    from astronomix_amd.time_stepping._snapshots import exact_end_time, run_loop, to_host
    require_supported(config, registered_variables)
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
    ctx = _native.fd_context(native_config(state.shape, config, params))
    run = _native.AmxFvRun(t_end=float(params.t_end), fixed_timestep=int(bool(config.fixed_timestep)),
                           num_timesteps=int(config.num_timesteps), exact_end_time=int(exact_end_time(config)),
                           max_steps=-1)
    if config.memory_analysis:
        import ctypes
        nbytes = ctypes.c_size_t(0)
        ctx.lib.amx_fd_workspace_bytes(ctypes.byref(ctx.cfg), ctypes.byref(nbytes))
        print("=== device memory PER DEVICE ===")
        print(f"State size: {state.numel() * 4 / 1024**2:.2f} MB")
        print(f"Workspace size: {nbytes.value / 1024**2:.2f} MB")
    if config.print_elapsed_time:
        print("Starting simulation...")
    stats, snaps, seconds = run_loop(ctx, state, run, config, params, registered_variables, snapshot_callable,
                                     kind="fd")
    if config.print_elapsed_time:
        print("Simulation finished!")
        print(f"Time elapsed: {seconds:.2f} seconds")
        print(f"Number of iterations: {stats.steps}")
    if config.runtime_debugging:
        if not torch.isfinite(state).all():
            raise FloatingPointError("non-finite values in the final state")
        if (state[4] < 0).any() or (state[0] < 0).any():
            raise FloatingPointError("negative density or pressure in the final state")
    if snaps is not None:
        return to_host(snaps) if host_numpy else snaps
    out = state.cpu().numpy() if host_numpy else state
    if struct is not None:
        return struct._replace(primitive_state=out)
    return out
