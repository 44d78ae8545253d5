"""Pins the scheme variants of the CPU restatement (oracle/fv_oracle.cpp):
the Osher, double-minmod (monotonized central), superbee and van Albada
limiters of limited_gradients.py:49-88 / limiters.py:18-64 against their
textbook flux-limiter functions, HLLC-LM (hll.py:206-221) against Toro's
upwind HLLC star states with the low-Mach scaling of the published method,
the Lax-Friedrichs flux (_lax_friedrichs.py:21-78) against its closed
form and its grid-wide alpha, and the hybrid / AM HLLC solvers
(hll.py:264-326): their switch against the closed-form divergence of a
sinusoidal flow, and their fluxes against HLLC / HLLC-LM with the low-Mach
dissipation control."""

import math

import numpy as np
import pytest

from tests.test_oracle import euler_flux

MINMOD, OSHER, DOUBLE_MINMOD, SUPERBEE, VAN_ALBADA = 0, 1, 2, 3, 4
HLL, HLLC, HLLC_LM, LF, HYBRID, AM = 0, 1, 2, 3, 4, 5


def _pairs(n=4000, seed=0):
    rng = np.random.default_rng(seed)
    a = rng.normal(0, 1, n)
    b = rng.normal(0, 1, n)
    b[:50] = a[:50]          # r = 1
    b[50:100] = 0.0          # a flat side
    b[100:150] = -a[100:150]  # an extremum
    return a, b


def _phi(lim, r):
    """Sweby flux-limiter functions phi(r), r = forward / backward gradient."""
    if lim == MINMOD:
        return np.maximum(0, np.minimum(r, 1))
    if lim == SUPERBEE:
        return np.maximum(0, np.maximum(np.minimum(2 * r, 1), np.minimum(r, 2)))
    if lim == DOUBLE_MINMOD:  # monotonized central
        return np.maximum(0, np.minimum(np.minimum(2 * r, (1 + r) / 2), 2))
    if lim == OSHER:  # beta = 1.3
        return np.maximum(0, np.minimum(r, 1.3))
    raise ValueError(lim)


@pytest.mark.parametrize("lim", [MINMOD, SUPERBEE, DOUBLE_MINMOD, OSHER])
def test_tvd_limiters_are_their_phi_functions(oracle, lim):
    a, b = _pairs()
    got = oracle.limit(lim, a, b, 0.01)
    want = a * _phi(lim, b / a)
    # Osher adds 1e-11 to the backward gradient before dividing
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)


def test_symmetric_limiters_do_not_depend_on_the_side(oracle):
    a, b = _pairs(seed=1)
    for lim in (MINMOD, SUPERBEE, DOUBLE_MINMOD, VAN_ALBADA):
        np.testing.assert_allclose(oracle.limit(lim, a, b, 0.01), oracle.limit(lim, b, a, 0.01), rtol=1e-14)
        np.testing.assert_allclose(oracle.limit(lim, -a, -b, 0.01), -oracle.limit(lim, a, b, 0.01), rtol=1e-14)


def test_van_albada_closed_form(oracle):
    dx = 0.02
    eps = 3 * dx
    a, b = _pairs(seed=2)
    got = oracle.limit(VAN_ALBADA, a, b, dx)
    # equal gradients pass through unchanged; a flat side keeps eps a / (a^2 + 2 eps)
    np.testing.assert_allclose(got[:50], a[:50], rtol=1e-14)
    np.testing.assert_allclose(got[50:100], eps * a[50:100] / (a[50:100] ** 2 + 2 * eps), rtol=1e-13)
    # the harmonic-like average lies between the two gradients when they share a sign
    same = a * b > 0
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    assert np.all((got[same] >= lo[same] - 1e-15) & (got[same] <= hi[same] + 1e-15))


def _states(n, mach, dim=3, seed=0, gamma=1.4):
    rng = np.random.default_rng(seed)
    w = np.empty((2, dim + 2, n))
    for s in range(2):
        w[s, 0] = rng.uniform(0.5, 2.0, n)
        w[s, dim + 1] = rng.uniform(0.5, 2.0, n)
        c = np.sqrt(gamma * w[s, dim + 1] / w[s, 0])
        w[s, 1:dim + 1] = rng.uniform(-1, 1, (dim, n)) * mach * c
    return w[0], w[1]


def _toro_hllc_parts(wl, wr, dim, gamma):
    """Einfeldt-Roe wave speeds as in hll.py:164-184, the contact speed and
    Toro's star states U*_K = rho_K (S_K - u_K) / (S_K - S*) [1, S*, v_t,
    E_K / rho_K + (S* - u_K)(S* + p_K / (rho_K (S_K - u_K)))]."""
    P = dim + 1

    def cons(w):
        E = w[P] / (gamma - 1) + 0.5 * w[0] * (w[1:P] ** 2).sum(0)
        return np.concatenate([w[:1], w[:1] * w[1:P], E[None]])

    rl, ul, pl = wl[0], wl[1], wl[P]
    rr, ur, pr = wr[0], wr[1], wr[P]
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    uh = (sl * ul + sr * ur) / (sl + sr)
    ch = np.sqrt((cl ** 2 * sl + cr ** 2 * sr) / (sl + sr) + 0.5 * sl * sr / (sl + sr) ** 2 * (ur - ul) ** 2)
    SL = np.minimum(ul - cl, uh - ch)
    SR = np.maximum(ur + cr, uh + ch)
    Ss = (pr - pl + rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))
    UL, UR = cons(wl), cons(wr)

    def star(w, U, S):
        r, u, p = w[0], w[1], w[P]
        f = r * (S - u) / (S - Ss)
        out = np.empty_like(U)
        out[0] = f
        out[1] = f * Ss
        out[2:P] = f * w[2:P]
        out[P] = f * (U[P] / r + (Ss - u) * (Ss + p / (r * (S - u))))
        return out

    Ma = np.maximum(np.abs(ul / cl), np.abs(ur / cr))
    phi = np.sin(np.minimum(1.0, Ma / 0.1) * np.pi / 2)
    return SL, SR, Ss, UL, UR, star(wl, UL, SL), star(wr, UR, SR), phi


@pytest.mark.parametrize("mach", [0.02, 0.07, 0.5])
def test_hllc_lm_scales_the_outer_dissipation(oracle, mach):
    dim, gamma = 3, 1.4
    wl, wr = _states(500, mach, dim, seed=int(mach * 100))
    f = oracle.riemann(wl, wr, dim, 1, gamma, HLLC)
    flm = oracle.riemann(wl, wr, dim, 1, gamma, HLLC_LM)
    SL, SR, Ss, UL, UR, USL, USR, phi = _toro_hllc_parts(wl, wr, dim, gamma)
    sub = (SL < 0) & (SR > 0)
    # HLLC-LM = HLLC with S_L, S_R scaled by phi in the dissipation only
    want = f - 0.5 * (1 - phi) * (SL * (USL - UL) + SR * (USR - UR))
    np.testing.assert_allclose(flm[:, sub], want[:, sub], rtol=1e-9, atol=1e-11)
    # interfaces at Ma >= 0.1 have phi = 1: plain HLLC, bit for bit
    fast = phi == 1.0
    assert fast.any() == (mach > 0.1)
    np.testing.assert_array_equal(flm[:, fast], f[:, fast])


def test_hllc_lm_consistency_and_upwinding(oracle):
    dim, gamma = 2, 1.4
    wl, _ = _states(64, 0.05, dim)
    np.testing.assert_allclose(oracle.riemann(wl, wl, dim, 2, gamma, HLLC_LM), euler_flux(wl, dim, 2, gamma),
                               rtol=1e-12, atol=1e-12)
    sup = np.array([[1.0], [5.0], [0.3], [1.0]])
    sup2 = np.array([[0.5], [4.0], [0.1], [0.8]])
    np.testing.assert_allclose(oracle.riemann(sup, sup2, dim, 1, gamma, HLLC_LM), euler_flux(sup, dim, 1, gamma),
                               rtol=1e-13)


def test_lax_friedrichs_closed_form(oracle):
    dim, gamma, alpha = 3, 5.0 / 3.0, 2.75
    wl, wr = _states(200, 0.8, dim, seed=4, gamma=gamma)

    def cons(w):
        E = w[dim + 1] / (gamma - 1) + 0.5 * w[0] * (w[1:dim + 1] ** 2).sum(0)
        return np.concatenate([w[:1], w[:1] * w[1:dim + 1], E[None]])

    for axis in (1, 2, 3):
        got = oracle.riemann(wl, wr, dim, axis, gamma, LF, alpha=alpha)
        want = 0.5 * (euler_flux(wl, dim, axis, gamma) + euler_flux(wr, dim, axis, gamma)) \
            - 0.5 * alpha * (cons(wr) - cons(wl))
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_lax_friedrichs_alpha_is_grid_wide(oracle):
    """Raising |u| + c in one cell changes the LF update everywhere (alpha is
    the grid maximum); with HLL the far cells are untouched."""
    n = 32
    s = np.zeros((4, n, n))
    x = (np.arange(n) + 0.5) / n
    s[0] = 1.0 + 0.2 * np.sin(2 * np.pi * x)[:, None]
    s[1] = 0.3 * np.cos(2 * np.pi * x)[:, None]
    s[2] = 0.1
    s[3] = 1.0
    hot = s.copy()
    hot[1, 3, 3] = 4.0  # the fastest cell
    for solver, far_changes in ((LF, True), (HLL, False)):
        oc = oracle.make_config(s.shape, 2, mhd=False, riemann=solver)
        a, b = oracle.gas_stage(s, oc, 1e-3), oracle.gas_stage(hot, oc, 1e-3)
        far = np.abs(a - b)[:, 16:, 16:].max()
        assert (far > 1e-9) == far_changes, (solver, far)
    # at far cells the new density is affine in the x-flux alpha (the
    # reconstruction there does not change): the change caused by a hot cell
    # scales with alpha - alpha_cold, alpha = max over the grid of |u_x| + c
    g = 5.0 / 3.0
    oc = oracle.make_config(s.shape, 2, mhd=False, riemann=LF, gamma=g)
    c = np.sqrt(g * s[3] / s[0])
    a0 = np.max(np.abs(s[1]) + c)
    base = oracle.gas_stage(s, oc, 1e-3)[0, 16:, 16:]
    d = {}
    for h in (2.0, 4.0):
        hot = s.copy()
        hot[1, 3, 3] = h
        d[h] = oracle.gas_stage(hot, oc, 1e-3)[0, 16:, 16:] - base
        assert h + c[3, 3] > a0
    ratio = ((4.0 + c[3, 3]) - a0) / ((2.0 + c[3, 3]) - a0)
    big = np.abs(d[2.0]) > 1e-3 * np.abs(d[2.0]).max()
    np.testing.assert_allclose(d[4.0][big] / d[2.0][big], ratio, rtol=1e-6)


def test_hybrid_hllc_selects_hllc_or_hllc_lm(oracle):
    dim, gamma = 3, 1.4
    wl, wr = _states(300, 0.05, dim, seed=7)
    f = oracle.riemann(wl, wr, dim, 2, gamma, HLLC)
    flm = oracle.riemann(wl, wr, dim, 2, gamma, HLLC_LM)
    np.testing.assert_array_equal(oracle.riemann(wl, wr, dim, 2, gamma, HYBRID, alpha=0.0), f)
    np.testing.assert_array_equal(oracle.riemann(wl, wr, dim, 2, gamma, HYBRID, alpha=1.0), flm)


@pytest.mark.parametrize("switch", [0.0, 1.0])
def test_am_hllc_low_mach_dissipation_control(oracle, switch):
    """AM-HLLC = (HLLC or HLLC-LM) with the velocity rows of the dissipation
    scaled by min(1, max(|v_L| / c_L, |v_R| / c_R)); bulk = (F_L + F_R) / 2."""
    dim, gamma = 3, 1.4
    wl, wr = _states(300, 0.3, dim, seed=8)
    base = oracle.riemann(wl, wr, dim, 1, gamma, HYBRID, alpha=switch)
    am = oracle.riemann(wl, wr, dim, 1, gamma, AM, alpha=switch)
    bulk = 0.5 * (euler_flux(wl, dim, 1, gamma) + euler_flux(wr, dim, 1, gamma))
    cl, cr = np.sqrt(gamma * wl[dim + 1] / wl[0]), np.sqrt(gamma * wr[dim + 1] / wr[0])
    vl = np.sqrt((wl[1:dim + 1] ** 2).sum(0) + 1e-20)
    vr = np.sqrt((wr[1:dim + 1] ** 2).sum(0) + 1e-20)
    fmd = np.minimum(1.0, np.maximum(vl / cl, vr / cr))
    SL, SR = _toro_hllc_parts(wl, wr, dim, gamma)[:2]
    sub = (SL < 0) & (SR > 0)
    want = base.copy()
    want[1:dim + 1] = bulk[1:dim + 1] + fmd * (base[1:dim + 1] - bulk[1:dim + 1])
    np.testing.assert_allclose(am[:, sub], want[:, sub], rtol=1e-10, atol=1e-12)
    np.testing.assert_array_equal(am[[0, dim + 1]][:, sub], base[[0, dim + 1]][:, sub])


def test_am_switch_closed_form(oracle):
    """v_x = -A sin(2 pi x) on a uniform gas: the central divergence is
    -A cos(2 pi x) sin(2 pi dx) / dx, so the switch is set where
    A cos(2 pi x) sin(2 pi dx) > 0.05 c."""
    n, A, g = 64, 2.0, 1.4
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx
    s = np.zeros((5, n, 8, 8))
    s[0], s[4] = 1.0, 1.0
    s[1] = (-A * np.sin(2 * np.pi * x))[:, None, None]
    oc = oracle.make_config(s.shape, 3, mhd=False, riemann=AM, gamma=g, box_size=8 * dx)
    oc.dx = dx
    sw = oracle.am_switch(s, oc)
    c = np.sqrt(g)
    want = (A * np.cos(2 * np.pi * x) * np.sin(2 * np.pi * dx) > 0.05 * c)
    margin = np.abs(A * np.cos(2 * np.pi * x) * np.sin(2 * np.pi * dx) - 0.05 * c) > 1e-9
    assert want.any() and (~want).any()
    np.testing.assert_array_equal(sw[:, 0, 0][margin], want[margin])
    assert np.all(sw == sw[:, :1, :1])
