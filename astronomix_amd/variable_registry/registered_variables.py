"""Where each field lives in the state array -- the layout contract of the
reference (``astronomix/variable_registry/registered_variables.py:54-238``).

Layouts produced by ``get_registered_variables`` (indices into axis 0):

=========  =======================================  ========
case       rows                                     num_vars
=========  =======================================  ========
1D hydro   rho, v, p                                3
2D hydro   rho, vx, vy, p                           4
2D MHD     rho, vx, vy, p, Bx, By, Bz               7
3D hydro   rho, vx, vy, vz, p                       5
3D MHD     rho, vx, vy, vz, p, Bx, By, Bz           8
FD MHD     3D MHD + face fields bx, by, bz          11
=========  =======================================  ========

Optional tracers (wind density, cosmic-ray n) are appended after these.
Finite-volume MHD relies on B being the last three rows, finite-difference
MHD on the face fields being the last three (evolve_state.py:510-517).
"""

from __future__ import annotations

from typing import NamedTuple, Tuple, Union

from astronomix_amd.option_classes.simulation_config import (
    FINITE_DIFFERENCE,
    XAXIS,
    YAXIS,
    ZAXIS,
    SimulationConfig,
)


class StaticIntVector(NamedTuple):
    x: int
    y: int
    z: int


class AxisInfo(NamedTuple):
    axis_in_array: int
    velocity_index: int
    magnetic_index: int


class RegisteredVariables(NamedTuple):
    axes_info: Tuple[AxisInfo, ...] = ()
    num_vars: int = 3
    density_index: int = 0
    velocity_index: Union[int, StaticIntVector] = 1
    momentum_index: Union[int, StaticIntVector] = 1
    magnetic_index: Union[int, StaticIntVector] = -1
    interface_magnetic_field_index: Union[int, StaticIntVector] = -1
    pressure_index: int = 2
    energy_index: int = 2
    wind_density_index: int = -1
    wind_density_active: bool = False
    cosmic_ray_n_index: int = -1
    cosmic_ray_n_active: bool = False


def get_registered_variables(config: SimulationConfig) -> RegisteredVariables:
    """Index map of the state array for ``config`` (see the module table)."""
    dim = config.dimensionality
    rv = RegisteredVariables()
    if dim == 2:
        velocity = StaticIntVector(1, 2, -1)
        if config.mhd:
            rv = rv._replace(num_vars=7, velocity_index=velocity, pressure_index=3,
                             magnetic_index=StaticIntVector(4, 5, 6))
        else:
            rv = rv._replace(num_vars=4, velocity_index=velocity, pressure_index=3)
    elif dim == 3:
        velocity = StaticIntVector(1, 2, 3)
        rv = rv._replace(num_vars=5, velocity_index=velocity, pressure_index=4)
        if config.mhd:
            rv = rv._replace(num_vars=8, magnetic_index=StaticIntVector(5, 6, 7))

    if config.mhd and config.solver_mode == FINITE_DIFFERENCE:
        n = rv.num_vars
        rv = rv._replace(interface_magnetic_field_index=StaticIntVector(n, n + 1, n + 2), num_vars=n + 3)

    if config.wind_config.trace_wind_density:
        rv = rv._replace(wind_density_index=rv.num_vars, num_vars=rv.num_vars + 1, wind_density_active=True)

    if config.cosmic_ray_config.cosmic_rays:
        rv = rv._replace(cosmic_ray_n_index=rv.num_vars, num_vars=rv.num_vars + 1, cosmic_ray_n_active=True)

    if dim == 3 and config.mhd:
        rv = rv._replace(axes_info=tuple(
            AxisInfo(axis_in_array=ax, velocity_index=rv.velocity_index[i], magnetic_index=rv.magnetic_index[i])
            for i, ax in enumerate((XAXIS, YAXIS, ZAXIS))))

    return rv._replace(momentum_index=rv.velocity_index, energy_index=rv.pressure_index)
