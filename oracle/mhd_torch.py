"""Differentiable float64 restatement of the split finite-volume MHD step in
torch -- TEST INFRASTRUCTURE ONLY (used by tests/ to check the reverse-mode
kernels of the MHD step; the reference's jax.grad cannot run here).

Follows the same reference lines as ``oracle/fv_oracle.cpp``
(evolve_state_fv, magnetic_update): S_A(dt/2) S_B(dt) S_A(dt/2)
(evolve_state.py:496-600) with S_A the gas SSP-RK2 step of
``oracle/hydro_torch.py`` at dt/2 and S_B the Pang-Wu magnetic update
(_magnetic_field_update.py:36-242): phi = (curl(avgB x avgv),
(avgB x curl avgB) / rho) with central differences (_vector_maths.py:5-100),
B_{k+1} = B0 - dt phi1, v_{k+1} = v0 - dt phi2, the first application from
avg = (B0 + B0) / 2, the later ones from (B_k + B0) / 2 (IMPLICIT_MIDPOINT)
or B_k (IMPLICIT_EULER); then the energy fix with gamma 5/3 through |v| =
sqrt(sum v^2) (lines 208-222) and the write-back of v (2D: its first two
rows) and p (lines 224-240). Periodic boundaries via torch.roll.

The number of map applications is an argument (``n_apply``): the reverse
mode differentiates the applications the forward pass made, as jax.grad
through the reference's while loop does, and the float64 restatement would
stop at a different count than a float32 run. Like JAX, torch gives the
derivative of sqrt at 0 as inf, so the energy fix's gradient is NaN where
|v| = 0; the states of the tests have no such cells.
"""

from __future__ import annotations

import torch

from oracle import hydro_torch as H


def _d(f, ax, dx):
    return (0.5 * (torch.roll(f, -1, ax) - torch.roll(f, 1, ax))) / dx


def curl(F, dim, dx):
    """curl3D / curl2D of a 3-component field F (3, *grid)."""
    if dim == 3:
        cx = _d(F[2], 1, dx) - _d(F[1], 2, dx)
        cy = _d(F[0], 2, dx) - _d(F[2], 0, dx)
        cz = _d(F[1], 0, dx) - _d(F[0], 1, dx)
    else:
        cx = _d(F[2], 1, dx)
        cy = -_d(F[2], 0, dx)
        cz = _d(F[1], 0, dx) - _d(F[0], 1, dx)
    return torch.stack([cx, cy, cz])


def cross(a, b):
    return torch.stack([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def magnetic_update(gas, B, dim, dx, dt, n_apply, euler=False):
    """S_B(dt): returns (gas with v and p updated, B_n)."""
    P = dim + 1
    rho = gas[0]
    v0 = gas[1:dim + 1]
    if dim == 2:
        v0 = torch.cat([v0, torch.zeros_like(v0[:1])])
    B0 = B
    Bk, vk = B0, v0
    for it in range(n_apply):
        if it == 0:
            a, u = (B0 + B0) / 2, (v0 + v0) / 2
        elif euler:
            a, u = Bk, vk
        else:
            a, u = (Bk + B0) / 2, (vk + v0) / 2
        phi1 = curl(cross(a, u), dim, dx)
        phi2 = cross(a, curl(a, dim, dx)) / rho
        Bk, vk = B0 - dt * phi1, v0 - dt * phi2
    g53 = 5.0 / 3.0
    p = gas[P]
    n0 = torch.sqrt((v0[0] * v0[0] + v0[1] * v0[1]) + v0[2] * v0[2])
    nn = torch.sqrt((vk[0] * vk[0] + vk[1] * vk[1]) + vk[2] * vk[2])
    E = p / (g53 - 1) + (0.5 * rho) * (n0 * n0)
    E2 = (E - (0.5 * rho) * (n0 * n0)) + (0.5 * rho) * (nn * nn)
    e = E2 / rho - 0.5 * (nn * nn)
    p2 = ((g53 - 1) * rho) * e
    gas_out = torch.stack([rho] + [vk[k] for k in range(dim)] + [p2])
    return gas_out, Bk


def mhd_step(w, dim, gamma, dx, dt, riemann, n_apply, euler=False):
    """One MHD step of the state rows (gas rows, then B_x, B_y, B_z)."""
    nv = dim + 2
    gas, B = w[:nv], w[nv:]
    gas = H.rk2_step(gas, dim, gamma, dx, dt / 2, riemann)
    gas, B = magnetic_update(gas, B, dim, dx, dt, n_apply, euler)
    gas = H.rk2_step(gas, dim, gamma, dx, dt / 2, riemann)
    return torch.cat([gas, B])
