"""astronomix_amd -- the per-step MHD update of astronomix on MI355X.

Host-side mirror of the reference's public surface (``astronomix/__init__.py``)
for the finite-volume / finite-difference step path; the numerics run in the
hand-written gfx950 HIP kernels of ``libamx.so`` (``include/amx.h``).
"""

from astronomix_amd.option_classes.simulation_config import (
    BACKWARDS,
    CARTESIAN,
    CYLINDRICAL,
    FORWARDS,
    HLL,
    HLLC,
    HLLC_LM,
    MINMOD,
    OPEN_BOUNDARY,
    OSHER,
    PERIODIC_BOUNDARY,
    REFLECTIVE_BOUNDARY,
    SPHERICAL,
    SimulationConfig,
    finalize_config,
)
from astronomix_amd.option_classes.simulation_params import SimulationParams, WindParams
from astronomix_amd.variable_registry.registered_variables import get_registered_variables
from astronomix_amd.initial_condition_generation.construct_primitive_state import construct_primitive_state
from astronomix_amd._finite_difference import initialize_interface_fields
from astronomix_amd.time_stepping.time_integration import time_integration

__all__ = [
    "BACKWARDS", "CARTESIAN", "CYLINDRICAL", "FORWARDS", "HLL", "HLLC", "HLLC_LM", "MINMOD",
    "OPEN_BOUNDARY", "OSHER", "PERIODIC_BOUNDARY", "REFLECTIVE_BOUNDARY", "SPHERICAL",
    "SimulationConfig", "SimulationParams", "WindParams", "finalize_config", "get_registered_variables",
    "construct_primitive_state", "initialize_interface_fields", "time_integration",
]
