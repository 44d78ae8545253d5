"""``SnapshotData``: the return of ``time_integration`` with
``config.return_snapshots`` (``data_classes/simulation_snapshot_data.py:5-49``
of the reference, same field names, order and defaults).

Array fields hold numpy arrays when the initial state was a numpy array and
float32 torch tensors on the GPU when it was a GPU tensor. ``runtime`` holds
the wall-clock seconds of the step loop (the reference leaves its default).
"""

from __future__ import annotations

from typing import Any, NamedTuple


class SnapshotData(NamedTuple):
    #: times at which the snapshots were taken, (num_snapshots,)
    time_points: Any = None
    #: primitive states at those times, (num_snapshots, num_vars, nx, ny[, nz])
    states: Any = None
    #: final state (SnapshotSettings.return_final_state)
    final_state: Any = None
    #: sum(rho) * box_size**dim per snapshot
    total_mass: Any = None
    #: sum(p / (gamma - 1) + rho u^2 / 2) dx**dim per snapshot
    total_energy: Any = None
    #: sum(p / (gamma - 1)) dx**dim per snapshot
    internal_energy: Any = None
    #: sum(rho u^2 / 2) dx**dim per snapshot
    kinetic_energy: Any = None
    #: self-gravity energy (self-gravity is outside the implemented path: None)
    gravitational_energy: Any = None
    #: sum(rho u . r_hat) dx**dim per snapshot
    radial_momentum: Any = None
    #: max |div B| per snapshot
    magnetic_divergence: Any = None
    #: wall-clock seconds of the step loop
    runtime: float = 0.0
    #: steps taken
    num_iterations: int = 0
    #: snapshots taken
    current_checkpoint: int = 0
