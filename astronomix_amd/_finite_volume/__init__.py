"""Finite-volume operators of the reference, backed by the HIP kernels.

``_cfl_time_step`` and ``_evolve_state_fv`` keep the reference's names and
argument meaning (``_finite_volume/_timestep_estimation/_timestep_estimator.py:94-103``,
``_finite_volume/_state_evolution/evolve_state.py:496-506``) and operate on
float32 torch tensors on the GPU in the reference's state layout.
"""

from __future__ import annotations

from astronomix_amd.option_classes.simulation_config import (
    ALL_PERIODIC,
    AM_HLLC,
    CARTESIAN,
    DOUBLE_MINMOD,
    DOUBLE_PRECISION,
    FINITE_VOLUME,
    HLL,
    HLLC,
    HLLC_LM,
    HYBRID_HLLC,
    IMPLICIT_EULER,
    IMPLICIT_MIDPOINT,
    LAX_FRIEDRICHS,
    MINMOD,
    OSHER,
    RK2_SSP,
    SPLIT,
    SUPERBEE,
    VAN_ALBADA,
    BoundarySettings,
)
from astronomix_amd import _native


def unsupported_reason(config, registered_variables=None):
    """None when ``config`` lies on the implemented path, else why not.

    The implemented path is the unsplit SSP-RK2 finite-volume scheme with the
    minmod, Osher, double-minmod, superbee or van Albada limiter, HLL, HLLC,
    HLLC-LM, hybrid / AM HLLC or Lax-Friedrichs fluxes, periodic boundaries on every axis, 2D
    or 3D, hydro or MHD (Pang-Wu magnetic sub-step), float32, no physics
    modules.
    """
    c = config
    if c.solver_mode != FINITE_VOLUME:
        return "solver_mode must be FINITE_VOLUME for this operator"
    if c.dimensionality not in (2, 3):
        return "only 2D and 3D grids are implemented"
    if c.numerical_precision == DOUBLE_PRECISION:
        return "only SINGLE_PRECISION (float32) is implemented"
    if c.geometry != CARTESIAN:
        return "only CARTESIAN geometry is implemented"
    if c.split == SPLIT:
        return "only the UNSPLIT scheme is implemented"
    if c.time_integrator != RK2_SSP:
        return "only the RK2_SSP integrator is implemented for the unsplit scheme"
    if c.limiter not in (MINMOD, OSHER, DOUBLE_MINMOD, SUPERBEE, VAN_ALBADA):
        return "the VAN_ALBADA_PP limiter is not implemented"
    if c.riemann_solver not in (HLL, HLLC, HLLC_LM, LAX_FRIEDRICHS, HYBRID_HLLC, AM_HLLC):
        return "unknown Riemann solver"
    if c.fv_magnetic_integrator not in (IMPLICIT_MIDPOINT, IMPLICIT_EULER):
        return "unknown fv_magnetic_integrator"
    if c.first_order_fallback:
        return "first_order_fallback is not implemented"
    bs = c.boundary_settings
    periodic = bs == ALL_PERIODIC or (isinstance(bs, BoundarySettings) and all(
        (getattr(bs, ax).left_boundary, getattr(bs, ax).right_boundary) == (2, 2)
        for ax in ("x", "y", "z")[: c.dimensionality]))
    if not periodic:
        return "only periodic boundaries on every axis are implemented"
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
    if c.source_term_aware_timestep:
        return "source_term_aware_timestep is not implemented"
    if registered_variables is not None:
        expected = c.dimensionality + 2 + (3 if c.mhd else 0)
        if registered_variables.num_vars != expected:
            return "extra registered variables (tracers) are not implemented"
    return None


def require_supported(config, registered_variables=None):
    why = unsupported_reason(config, registered_variables)
    if why is not None:
        raise NotImplementedError(why)


def native_config(state_shape, config, gamma, C_cfl, grid_spacing=None) -> _native.AmxFvConfig:
    dim = config.dimensionality
    sizes = list(state_shape[1:])
    nx, ny = sizes[0], sizes[1]
    nz = sizes[2] if dim == 3 else 1
    h = config.grid_spacing if grid_spacing is None else grid_spacing
    return _native.AmxFvConfig(
        dim=dim, nx=nx, ny=ny, nz=nz, mhd=int(bool(config.mhd)),
        riemann=int(config.riemann_solver), limiter=int(config.limiter),
        magnetic_integrator=int(config.fv_magnetic_integrator), max_fixed_point_iterations=1000,
        gamma=float(gamma), C_cfl=float(C_cfl), fixed_point_tol=1e-5, grid_spacing=float(h))


def _as_device_state(state):
    import torch
    _native.require_gpu()
    if not isinstance(state, torch.Tensor):
        raise TypeError("state must be a torch tensor on the GPU")
    if state.dtype != torch.float32:
        raise TypeError("state must be float32")
    if state.device.type != "cuda":
        raise TypeError("state must live on the GPU")
    return state.contiguous()


def _cfl_time_step(primitive_state, grid_spacing, dt_max, gamma, config, registered_variables, C_CFL=0.8):
    """CFL step of the unsplit scheme: C / sum_a(max(|u_a| + c) / dx) as a
    0-d device tensor. No dt_max clamp on the unsplit branch
    (_timestep_estimator.py:118-128)."""
    import torch
    require_supported(config)
    s = _as_device_state(primitive_state)
    ctx = _native.fv_context(native_config(s.shape, config, gamma, C_CFL, grid_spacing))
    out = torch.empty((), dtype=torch.float32, device=s.device)
    ctx.cfl_dt(s, out)
    return out


def _evolve_state_fv(primitive_state, dt, gamma, gravitational_constant, config, params, helper_data,
                     registered_variables):
    """One finite-volume step; returns the new state (input unchanged).
    MHD: S_A(dt/2) o S_B(dt) o S_A(dt/2); hydro: one SSP-RK2 step
    (evolve_state.py:496-600)."""
    import torch
    require_supported(config, registered_variables)
    s = _as_device_state(primitive_state).clone()
    ctx = _native.fv_context(native_config(s.shape, config, gamma, params.C_cfl))
    if not isinstance(dt, torch.Tensor):
        dt = torch.tensor(float(dt), dtype=torch.float32, device=s.device)
    dt = dt.to(device=s.device, dtype=torch.float32).reshape(())
    ctx.evolve_state(s, dt)
    return s
