"""Stack primitive fields into a state array in the registry's layout,
mirroring ``astronomix/initial_condition_generation/construct_primitive_state.py:17-116``.

Fields may be numpy arrays or torch tensors; the result is a torch tensor on
the fields' device when any field is a tensor, else a numpy array. The state
dtype is float32 (the reference's single-precision default) unless ``dtype``
says otherwise.
"""

from __future__ import annotations

import numpy as np

from astronomix_amd.option_classes.simulation_config import FINITE_DIFFERENCE


def construct_primitive_state(config, registered_variables, density, velocity_x=None, velocity_y=None,
                              velocity_z=None, magnetic_field_x=None, magnetic_field_y=None,
                              magnetic_field_z=None, interface_magnetic_field_x=None,
                              interface_magnetic_field_y=None, interface_magnetic_field_z=None,
                              gas_pressure=None, cosmic_ray_pressure=None, sharding=None, dtype=None):
    rv = registered_variables
    fields = [density, velocity_x, velocity_y, velocity_z, magnetic_field_x, magnetic_field_y,
              magnetic_field_z, interface_magnetic_field_x, interface_magnetic_field_y,
              interface_magnetic_field_z, gas_pressure, cosmic_ray_pressure]
    torch_like = None
    for f in fields:
        if f is not None and type(f).__module__.startswith("torch"):
            torch_like = f
            break
    if torch_like is not None:
        import torch
        dt = dtype or torch.float32
        state = torch.zeros((rv.num_vars, *density.shape), dtype=dt, device=torch_like.device)

        def put(i, f):
            state[i] = f if isinstance(f, torch.Tensor) else torch.as_tensor(np.asarray(f), dtype=dt)
    else:
        dt = dtype or np.float32
        state = np.zeros((rv.num_vars, *np.shape(density)), dtype=dt)

        def put(i, f):
            state[i] = np.asarray(f)

    put(rv.density_index, density)
    dim = config.dimensionality
    if dim == 1:
        put(rv.velocity_index, velocity_x)
    else:
        put(rv.velocity_index.x, velocity_x)
        put(rv.velocity_index.y, velocity_y)
        if dim == 3:
            put(rv.velocity_index.z, velocity_z)
    if config.mhd and dim >= 2:
        for comp, f in zip("xyz", (magnetic_field_x, magnetic_field_y, magnetic_field_z)):
            if f is not None:
                put(getattr(rv.magnetic_index, comp), f)
        if config.solver_mode == FINITE_DIFFERENCE:
            for comp, f in zip("xyz", (interface_magnetic_field_x, interface_magnetic_field_y,
                                       interface_magnetic_field_z)):
                if f is not None:
                    put(getattr(rv.interface_magnetic_field_index, comp), f)
    put(rv.pressure_index, gas_pressure)
    if rv.cosmic_ray_n_active:
        put(rv.pressure_index, gas_pressure + cosmic_ray_pressure)
        put(rv.cosmic_ray_n_index, cosmic_ray_pressure ** (3 / 4))
    return state
