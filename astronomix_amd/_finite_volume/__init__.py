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
    GHOST_CELLS,
    LAX_FRIEDRICHS,
    MINMOD,
    OPEN_BOUNDARY,
    OSHER,
    PERIODIC_BOUNDARY,
    REFLECTIVE_BOUNDARY,
    RK2_SSP,
    SPLIT,
    SUPERBEE,
    VAN_ALBADA,
    VAN_ALBADA_PP,
    BoundarySettings,
)
import ctypes

from astronomix_amd import _native


def unsupported_reason(config, registered_variables=None):
    """None when ``config`` lies on the implemented path, else why not.

    The implemented path is the unsplit SSP-RK2 finite-volume scheme with the
    minmod, Osher, double-minmod, superbee, van Albada or positivity-preserving
    van Albada limiter, HLL, HLLC,
    HLLC-LM, hybrid / AM HLLC or Lax-Friedrichs fluxes, periodic, open or
    reflective boundaries (both sides of an axis periodic or neither; the
    hybrid / AM solvers periodic only), 2D or 3D, hydro or MHD (Pang-Wu
    magnetic sub-step), float32, no physics modules.
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
    if c.limiter not in (MINMOD, OSHER, DOUBLE_MINMOD, SUPERBEE, VAN_ALBADA, VAN_ALBADA_PP):
        return "unknown limiter"
    if c.riemann_solver not in (HLL, HLLC, HLLC_LM, LAX_FRIEDRICHS, HYBRID_HLLC, AM_HLLC):
        return "unknown Riemann solver"
    if c.fv_magnetic_integrator not in (IMPLICIT_MIDPOINT, IMPLICIT_EULER):
        return "unknown fv_magnetic_integrator"
    if c.first_order_fallback:
        return "first_order_fallback is not implemented"
    try:
        codes = boundary_codes(c)
    except NotImplementedError as e:
        return str(e)
    if any(codes) and c.riemann_solver in (HYBRID_HLLC, AM_HLLC):
        return "HYBRID_HLLC and AM_HLLC are implemented with periodic boundaries only"
    if any(codes) and c.limiter == VAN_ALBADA_PP:
        return "VAN_ALBADA_PP is implemented with periodic boundaries only"
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


# reference boundary code (simulation_config.py:49-52) -> AMX_BC_* (amx.h)
_BC_CODE = {PERIODIC_BOUNDARY: _native.AMX_BC_PERIODIC, OPEN_BOUNDARY: _native.AMX_BC_OPEN,
            REFLECTIVE_BOUNDARY: _native.AMX_BC_REFLECTIVE}


def boundary_codes(config) -> tuple:
    """(x low, x high, y low, y high, z low, z high) AMX_BC_* codes of a
    finite-volume configuration. Ghost cells are used only with
    boundary_handling == GHOST_CELLS (time_integration.py:276-286,
    evolve_state.py:350-412); otherwise every stencil rolls, i.e. periodic."""
    bs = config.boundary_settings
    if config.boundary_handling != GHOST_CELLS or bs is None or bs == ALL_PERIODIC:
        return (0,) * 6
    if not isinstance(bs, BoundarySettings):
        raise NotImplementedError("2D / 3D runs need BoundarySettings (x, y, z)")
    codes = []
    for ax in ("x", "y", "z")[: config.dimensionality]:
        side = getattr(bs, ax)
        pair = (side.left_boundary, side.right_boundary)
        if any(b not in _BC_CODE for b in pair):
            raise NotImplementedError("only periodic, open and reflective boundaries are implemented "
                                      "(not the MHD jet boundary)")
        if (pair[0] == PERIODIC_BOUNDARY) != (pair[1] == PERIODIC_BOUNDARY):
            raise NotImplementedError(f"axis {ax}: periodic on one side only has no defined ghost cells "
                                      "(boundaries.py applies periodic ghosts when both sides are periodic)")
        codes += [_BC_CODE[pair[0]], _BC_CODE[pair[1]]]
    return tuple(codes + [0] * (6 - len(codes)))


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
        gamma=float(gamma), C_cfl=float(C_cfl), fixed_point_tol=1e-5, grid_spacing=float(h),
        boundary=(ctypes.c_int * 6)(*boundary_codes(config)))


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
