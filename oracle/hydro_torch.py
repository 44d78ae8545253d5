"""Differentiable float64 restatement of the unsplit SSP-RK2 hydro step in
torch -- TEST INFRASTRUCTURE ONLY (used by tests/ to check the reverse-mode
kernels; the reference's jax.grad cannot run here).

Follows the same reference lines as ``oracle/fv_oracle.cpp`` (gas_stage,
gas_rk2): minmod on primitive variables (limiters.py:13-15,
limited_gradients.py:50-60, reconstruction.py:409-429), HLL / HLLC
(hll.py:33-261) with the Euler flux (_fluxes.py:22-93), the conversions of
_equations.py:23-199 (2D |v| without the 1e-20, 3D with it; the inverse with
it), U += dt/dx (F_i - F_{i+1}) per axis (evolve_state.py:398-404), Heun
average (evolve_state.py:456-469). Periodic boundaries via torch.roll.
torch's autograd rules at ties (minimum / maximum split the gradient, sign
and |x| have zero derivative at 0) are JAX's.
"""

from __future__ import annotations

import torch

HLL, HLLC = 0, 1


def _absv(w, dim):
    s = w[1] * w[1] + w[2] * w[2]
    if dim == 3:
        return torch.sqrt(s + w[3] * w[3] + 1e-20)
    return torch.sqrt(s)


def cons(w, dim, gamma):
    P = dim + 1
    rho = w[0]
    av = _absv(w, dim)
    out = [rho] + [rho * w[a] for a in range(1, dim + 1)] + [w[P] / (gamma - 1) + (0.5 * rho) * (av * av)]
    return torch.stack(out)


def prim(u, dim, gamma):
    P = dim + 1
    rho = u[0]
    v = [u[a] / rho for a in range(1, dim + 1)]
    s = v[0] * v[0] + v[1] * v[1]
    if dim == 3:
        s = s + v[2] * v[2]
    uabs = torch.sqrt(s + 1e-20)
    e = u[P] / rho - 0.5 * (uabs * uabs)
    p = ((gamma - 1) * rho) * e
    return torch.stack([rho] + v + [p])


def euler_flux(w, dim, axis, gamma):
    P = dim + 1
    rho, p = w[0], w[P]
    E = p / (gamma - 1) + (0.5 * rho) * (_absv(w, dim) ** 2)
    ua = w[axis]
    f = [ua * rho] + [ua * (w[a] * rho) + (p if a == axis else 0.0) for a in range(1, dim + 1)] + [ua * (p + E)]
    return torch.stack(f)


def hll(wl, wr, dim, axis, gamma):
    P = dim + 1
    rl, ul, pl = wl[0], wl[axis], wl[P]
    rr, ur, pr = wr[0], wr[axis], wr[P]
    cl = torch.sqrt((gamma * pl) / rl)
    cr = torch.sqrt((gamma * pr) / rr)
    fl, fr = euler_flux(wl, dim, axis, gamma), euler_flux(wr, dim, axis, gamma)
    zero = torch.zeros_like(ul)
    sp = torch.maximum(torch.maximum(ul + cl, ur + cr), zero)
    sm = torch.minimum(torch.minimum(ul - cl, ur - cr), zero)
    UL, UR = cons(wl, dim, gamma), cons(wr, dim, gamma)
    return ((sp * fl - sm * fr) + (sm * sp) * (UR - UL)) / (sp - sm)


def hllc(wl, wr, dim, axis, gamma):
    P = dim + 1
    rl, ul, pl = wl[0], wl[axis], wl[P]
    rr, ur, pr = wr[0], wr[axis], wr[P]
    cl = torch.sqrt((gamma * pl) / rl)
    cr = torch.sqrt((gamma * pr) / rr)
    FL, FR = euler_flux(wl, dim, axis, gamma), euler_flux(wr, dim, axis, gamma)
    sql, sqr = torch.sqrt(rl), torch.sqrt(rr)
    ssum = sql + sqr
    u_hat = (sql * ul + sqr * ur) / ssum
    du = ur - ul
    c_hat = torch.sqrt(((cl * cl) * sql + (cr * cr) * sqr) / ssum + (0.5 * ((sql * sqr) / (ssum * ssum))) * (du * du))
    SL = torch.minimum(ul - cl, u_hat - c_hat)
    SR = torch.maximum(ur + cr, u_hat + c_hat)
    s_star = (((pr - pl) + (rl * ul) * (SL - ul)) - (rr * ur) * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))
    UL, UR = cons(wl, dim, gamma), cons(wr, dim, gamma)
    USL = list(UL)
    USR = list(UR)
    USL[axis] = rl * s_star
    USL[P] = USL[P] + (s_star - ul) * (rl * s_star + pl / (SL - ul))
    USR[axis] = rr * s_star
    USR[P] = USR[P] + (s_star - ur) * (rr * s_star + pr / (SR - ur))
    USL = torch.stack([(x * (SL - ul)) / (SL - s_star) for x in USL])
    USR = torch.stack([(x * (SR - ur)) / (SR - s_star) for x in USR])
    a = torch.abs(s_star)
    fs = 0.5 * (FL + FR) + 0.5 * ((SL * (USL - UL) + a * (USL - USR)) + SR * (USR - UR))
    out = torch.where(SL >= 0, FL, fs)
    return torch.where(SR <= 0, FR, out)


def minmod(a, b):
    return (0.5 * (torch.sign(a) + torch.sign(b))) * torch.minimum(torch.abs(a), torch.abs(b))


def stage(w, dim, gamma, dx, dt, riemann):
    U = cons(w, dim, gamma)
    for axis in range(1, dim + 1):
        d = axis  # array dim of this axis (row 0 is the variable index)
        wm2, wm1, wp1 = torch.roll(w, 2, d), torch.roll(w, 1, d), torch.roll(w, -1, d)
        g_m1 = minmod((wm1 - wm2) / dx, (w - wm1) / dx)
        g_0 = minmod((w - wm1) / dx, (wp1 - w) / dx)
        wl = wm1 + (g_m1 * dx) / 2
        wr = w - (g_0 * dx) / 2
        F = (hllc if riemann == HLLC else hll)(wl, wr, dim, axis, gamma)  # F[i] at i - 1/2
        U = U + ((1.0 / dx) * (F - torch.roll(F, -1, d))) * dt
    return prim(U, dim, gamma)


def rk2_step(w, dim, gamma, dx, dt, riemann):
    w1 = stage(w, dim, gamma, dx, dt, riemann)
    w2 = stage(w1, dim, gamma, dx, dt, riemann)
    return prim(0.5 * (cons(w, dim, gamma) + cons(w2, dim, gamma)), dim, gamma)


def rollout(w, steps, dim, gamma, dx, dt, riemann):
    for _ in range(steps):
        w = rk2_step(w, dim, gamma, dx, dt, riemann)
    return w
