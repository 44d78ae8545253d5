"""Static simulation configuration, mirroring the reference's
``astronomix/option_classes/simulation_config.py``.

The integer option values are the reference's (``simulation_config.py:21-93``)
so that configurations written against the reference mean the same thing
here. ``SimulationConfig`` carries the same field names and defaults
(``simulation_config.py:160-342``); ``finalize_config`` reproduces the
reference's validation and defaulting (``simulation_config.py:345-470``).
"""

from __future__ import annotations

from typing import NamedTuple, Optional, Union

# solver modes
FINITE_VOLUME = 0
FINITE_DIFFERENCE = 1

# differentiation modes
FORWARDS = 0
BACKWARDS = 1

# limiters
MINMOD = 0
OSHER = 1
DOUBLE_MINMOD = 2
SUPERBEE = 3
VAN_ALBADA = 4
VAN_ALBADA_PP = 5

# splitting
UNSPLIT = 0
SPLIT = 1

# Riemann solvers
HLL = 0
HLLC = 1
HLLC_LM = 2
LAX_FRIEDRICHS = 3
HYBRID_HLLC = 4
AM_HLLC = 5

# time integrators
RK2_SSP = 0
MUSCL = 1
RK4_SSP = 2

# boundary conditions
OPEN_BOUNDARY = 0
REFLECTIVE_BOUNDARY = 1
PERIODIC_BOUNDARY = 2
MHD_JET_BOUNDARY = 3

GAS_STATE = 0
VELOCITY_ONLY = 1
MAGNETIC_FIELD_ONLY = 2

# geometry
CARTESIAN = 0
CYLINDRICAL = 1
SPHERICAL = 2

# axes of the state array
VARAXIS = 0
XAXIS = 1
YAXIS = 2
ZAXIS = 3

# boundary handling
GHOST_CELLS = 0
PERIODIC_ROLL = 1

# self-gravity variants
SIMPLE_SOURCE_TERM = 0
DONOR_ACCOUNTING = 1
RIEMANN_SPLIT = 2
RIEMANN_SPLIT_UNSTABLE = 3
HALF_SPLIT = 4

# magnetic sub-step integrators
IMPLICIT_MIDPOINT = 0
IMPLICIT_EULER = 1

# numerical precision
SINGLE_PRECISION = 0
DOUBLE_PRECISION = 1


class SnapshotSettings(NamedTuple):
    return_states: bool = True
    return_final_state: bool = False
    return_total_mass: bool = False
    return_total_energy: bool = False
    return_internal_energy: bool = False
    return_kinetic_energy: bool = False
    return_gravitational_energy: bool = False
    return_radial_momentum: bool = False
    return_magnetic_divergence: bool = False


class BoundarySettings1D(NamedTuple):
    left_boundary: int = OPEN_BOUNDARY
    right_boundary: int = OPEN_BOUNDARY


class BoundarySettings(NamedTuple):
    x: BoundarySettings1D = BoundarySettings1D()
    y: BoundarySettings1D = BoundarySettings1D()
    z: BoundarySettings1D = BoundarySettings1D()


# Physics-module switches. The modules themselves hook in between steps and
# are not part of this build; the switches exist so that configurations keep
# their field names, and time_integration rejects any module that is turned on.
class TurbulentForcingConfig(NamedTuple):
    turbulent_forcing: bool = False


class WindConfig(NamedTuple):
    stellar_wind: bool = False
    num_injection_cells: int = 10
    wind_injection_scheme: int = 1
    trace_wind_density: bool = False


class CosmicRayConfig(NamedTuple):
    cosmic_rays: bool = False
    diffusive_shock_acceleration: bool = False


class CoolingConfig(NamedTuple):
    cooling: bool = False
    cooling_method: int = 2
    cooling_curve_config: object = None


class NeuralNetForceConfig(NamedTuple):
    neural_net_force: bool = False
    network_static: object = None


class CNNMHDconfig(NamedTuple):
    cnn_mhd_corrector: bool = False
    network_static: object = None


class SimulationConfig(NamedTuple):
    solver_mode: int = FINITE_VOLUME
    numerical_precision: int = SINGLE_PRECISION
    runtime_debugging: bool = False
    donate_state: bool = False
    memory_analysis: bool = False
    print_elapsed_time: bool = False
    progress_bar: bool = False
    dimensionality: int = 1
    state_struct: bool = False
    geometry: int = CARTESIAN
    mhd: bool = False
    fv_magnetic_integrator: int = IMPLICIT_MIDPOINT
    enforce_positivity: bool = True
    self_gravity: bool = False
    self_gravity_version: int = DONOR_ACCOUNTING
    box_size: float = 1.0
    num_cells: int = 400
    reconstruction_order: int = 1
    limiter: int = MINMOD
    riemann_solver: int = HLL
    split: int = UNSPLIT
    time_integrator: int = RK2_SSP
    num_ghost_cells: int = 2
    grid_spacing: float = 1.0 / 400
    boundary_handling: int = GHOST_CELLS
    boundary_settings: Optional[Union[BoundarySettings1D, BoundarySettings]] = None
    fixed_timestep: bool = False
    exact_end_time: bool = True
    source_term_aware_timestep: bool = False
    num_timesteps: int = 1000
    use_max_adaptive_timestep: bool = True
    differentiation_mode: int = FORWARDS
    num_checkpoints: int = 100
    return_snapshots: bool = False
    snapshot_settings: SnapshotSettings = SnapshotSettings()
    activate_snapshot_callback: bool = False
    use_specific_snapshot_timepoints: bool = False
    num_snapshots: int = 10
    first_order_fallback: bool = False
    turbulent_forcing_config: TurbulentForcingConfig = TurbulentForcingConfig()
    wind_config: WindConfig = WindConfig()
    cosmic_ray_config: CosmicRayConfig = CosmicRayConfig()
    cooling_config: CoolingConfig = CoolingConfig()
    neural_net_force_config: NeuralNetForceConfig = NeuralNetForceConfig()
    cnn_mhd_corrector_config: CNNMHDconfig = CNNMHDconfig()


_PERIODIC_1D = BoundarySettings1D(PERIODIC_BOUNDARY, PERIODIC_BOUNDARY)
ALL_PERIODIC = BoundarySettings(_PERIODIC_1D, _PERIODIC_1D, _PERIODIC_1D)


def finalize_config(config: SimulationConfig, state_shape) -> SimulationConfig:
    """Set num_cells and grid_spacing from the state shape and apply the
    reference's consistency rules (simulation_config.py:345-470)."""
    n = int(state_shape[-1])
    config = config._replace(num_cells=n)
    dim = config.dimensionality
    if dim >= 2:
        sizes = tuple(int(s) for s in state_shape[-dim:])
        if any(s != sizes[0] for s in sizes):
            axes = "x and y" if dim == 2 else "x, y and z"
            raise ValueError(f"The number of cells in {axes} must be equal.")
    config = config._replace(grid_spacing=config.box_size / config.num_cells)

    if config.geometry == SPHERICAL:
        print("For spherical geometry, only HLL is currently supported. "
              "Also, only the unsplit mode has been tested.")
        if config.riemann_solver != HLL:
            print("Setting HLL Riemann solver for spherical geometry.")
            config = config._replace(riemann_solver=HLL)
        if config.split != SPLIT:
            print("Setting unsplit mode for spherical geometry")
            config = config._replace(split=SPLIT)
        if config.limiter in (VAN_ALBADA, VAN_ALBADA_PP):
            print("Setting minmod limiter for spherical geometry")
            config = config._replace(limiter=MINMOD)
        if config.time_integrator != MUSCL:
            print("Setting MUSCL time integrator for spherical geometry")
            config = config._replace(time_integrator=MUSCL)

    if config.self_gravity and config.limiter != MINMOD:
        print("Setting MINMOD limiter for self-gravity.")
        config = config._replace(limiter=MINMOD)

    if config.solver_mode == FINITE_DIFFERENCE:
        if not config.mhd:
            raise ValueError("Finite difference solver mode is currently only supported for MHD simulations.")
        if dim != 3:
            raise ValueError("Finite difference solver mode is currently only supported for 3D simulations.")
        if config.boundary_settings != ALL_PERIODIC:
            raise ValueError("Finite difference solver mode currently only supports periodic boundaries.")
        if config.time_integrator != RK4_SSP:
            print("Setting time integrator to RK4_SSP for finite difference solver mode.")
            config = config._replace(time_integrator=RK4_SSP)
        if config.boundary_handling != PERIODIC_ROLL:
            print("Setting boundary handling to PERIODIC_ROLL for finite difference solver mode.")
            config = config._replace(boundary_handling=PERIODIC_ROLL)

    if config.boundary_settings is None:
        if config.geometry == CARTESIAN:
            print("Automatically setting open boundaries for Cartesian geometry.")
            if dim == 1:
                config = config._replace(boundary_settings=BoundarySettings1D(OPEN_BOUNDARY, OPEN_BOUNDARY))
            else:
                config = config._replace(boundary_settings=BoundarySettings())
        elif config.geometry == SPHERICAL and dim == 1:
            print("Automatically setting reflective left and open right boundary for spherical geometry.")
            config = config._replace(
                boundary_settings=BoundarySettings1D(REFLECTIVE_BOUNDARY, OPEN_BOUNDARY))

    if config.wind_config.stellar_wind:
        print("For stellar wind simulations, we need source term aware timesteps, turning on.")
        config = config._replace(source_term_aware_timestep=True)

    return config
