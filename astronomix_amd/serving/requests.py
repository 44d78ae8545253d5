"""Synthetic request stream of the serving workload.

Poisson arrivals (rate 50/s, seeded) of Orszag-Tang (2D, grids 64^2..512^2)
and MHD blast (3D, grids 64^3..128^3) simulations with gamma ~ U[1.4, 5/3],
blast inner pressure ~ U[5, 50] and t_end ~ U[0.1, 0.5]; grid sizes drawn
uniformly from the seven.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from astronomix_amd import workloads as W
from astronomix_amd.option_classes.simulation_config import HLL

GRIDS = (("orszag_tang_2d", 64), ("orszag_tang_2d", 128), ("orszag_tang_2d", 256), ("orszag_tang_2d", 512),
         ("mhd_blast_3d", 64), ("mhd_blast_3d", 96), ("mhd_blast_3d", 128))
C_CFL = {"orszag_tang_2d": 0.4, "mhd_blast_3d": 0.3}


@dataclass(frozen=True)
class Request:
    rid: int
    kind: str
    n: int
    gamma: float
    t_end: float
    p_in: float
    arrival: float  # seconds after the start of the stream

    @property
    def grid_class(self):
        return (self.kind, self.n)


def mixed_requests(count: int, rate: float = 50.0, seed: int = 0, grids=GRIDS, t_end_range=(0.1, 0.5)):
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    out = []
    for i in range(count):
        kind, n = grids[int(rng.integers(len(grids)))]
        gamma = float(rng.uniform(1.4, 5.0 / 3.0))
        p_in = float(rng.uniform(5.0, 50.0))
        t_end = float(rng.uniform(*t_end_range))
        out.append(Request(i, kind, n, gamma, t_end, p_in, float(arrivals[i])))
    return out


def initial_state(req: Request) -> np.ndarray:
    if req.kind == "orszag_tang_2d":
        return W.orszag_tang_2d(req.n)
    if req.kind == "mhd_blast_3d":
        return W.blast_3d(req.n, p_in=req.p_in)
    raise KeyError(req.kind)


def request_config(kind: str, n: int):
    """(SimulationConfig, C_cfl) of a grid class."""
    dim = 2 if kind == "orszag_tang_2d" else 3
    return W.fv_config(dim, n, mhd=True, riemann=HLL), C_CFL[kind]
