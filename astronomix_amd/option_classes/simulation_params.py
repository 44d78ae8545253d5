"""Run-time parameters, mirroring the reference's
``astronomix/option_classes/simulation_params.py:11-66`` (same field names
and defaults). Physics-module parameter blocks are kept as opaque fields."""

from __future__ import annotations

from typing import NamedTuple


class TurbulentForcingParams(NamedTuple):
    energy_injection_rate: float = 2.0


class WindParams(NamedTuple):
    wind_mass_loss_rate: float = 0.0
    wind_final_velocity: float = 0.0
    pressure_floor: float = 100000.0


class SimulationParams(NamedTuple):
    C_cfl: float = 0.4
    gravitational_constant: float = 1.0
    gamma: float = 5 / 3
    minimum_density: float = 1e-14
    minimum_pressure: float = 1e-14
    dt_max: float = 0.001
    t_end: float = 0.2
    snapshot_timepoints: tuple = (0.0,)
    turbulent_forcing_params: TurbulentForcingParams = TurbulentForcingParams()
    wind_params: WindParams = WindParams()
    cosmic_ray_params: object = None
    cooling_params: object = None
    neural_net_force_params: object = None
    cnn_mhd_corrector_params: object = None
