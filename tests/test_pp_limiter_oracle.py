"""Pins the VAN_ALBADA_PP limiter of the CPU restatement (oracle/fv_oracle.cpp
pp_factors, _finite_volume/_state_evolution/reconstruction.py:168-365):

* its factors against a numpy restatement written in the reference's array
  form (van Albada differences along every axis stacked into an
  (axis, variable) array, the weights broadcast as the reference broadcasts
  them), on states where every factor leaves 1 somewhere;
* on a smooth, weakly varying state every factor is 1 and the stage equals the
  plain van Albada stage bit for bit;
* the scaled differences keep the reconstructed density and pressure of
  every face non-negative where the plain van Albada faces go negative."""

import numpy as np
import pytest

VAN_ALBADA, VAN_ALBADA_PP = 4, 5
HLL, HLLC = 0, 1


def va_differences(w, dim, dx):
    """(dim, nv) + grid: van Albada gradient along each axis times dx / 2
    (limited_gradients.py:21-88, limiters.py:48-64, reconstruction.py:157)."""
    out = []
    eps = 3 * dx
    for a in range(dim):
        ax = 1 + a
        b = (w - np.roll(w, 1, axis=ax)) / dx
        f = (np.roll(w, -1, axis=ax) - w) / dx
        g = ((f ** 2 + eps) * b + (b ** 2 + eps) * f) / (f ** 2 + b ** 2 + 2 * eps)
        out.append(g * dx / 2)
    return np.stack(out)


def pp_factors_arrays(w, dim, dx, gamma, C_cfl, mhd):
    """The factor computation in the reference's array form (broadcasting
    included): returns (alpha (dim,...), kappa (dim,...), beta (...))."""
    P = dim + 1
    d = va_differences(w, dim, dx)
    rho, p = w[0], w[P]
    c = np.sqrt(gamma * p / rho)
    alpha_lax = np.array([np.max(np.abs(w[1 + a]) + c) for a in range(dim)])
    C = alpha_lax / np.sum(alpha_lax)
    q = 2 / C_cfl if mhd else 1 / C_cfl
    eps = 1e-14
    Cb = C.reshape((dim,) + (1,) * dim)
    dr, dp, dv = d[:, 0], d[:, P], d[:, 1:1 + dim]
    drp = np.where(np.abs(dr) > eps, dr, eps)
    dpp = np.where(np.abs(dp) > eps, dp, eps)
    alpha = np.where(np.abs(dr) > eps, np.minimum(rho / (np.abs(drp) * (1 + eps)), 1), 1)
    kappa = np.where(np.abs(dp) > eps, np.minimum(p / (np.abs(dpp) * (1 + eps)), 1), 1)
    A1 = np.sum(np.sum(Cb * alpha * dr * dv, axis=0) ** 2, axis=0)
    A2 = np.sum(Cb * np.sum(dv ** 2, axis=1), axis=0)
    vsum = np.sum(np.sum(dv ** 2, axis=1), axis=0)
    A1 = np.where(vsum > eps, A1, eps)
    A2 = np.where(vsum > eps, A2, eps)
    with np.errstate(invalid="ignore"):
        beta = np.where(vsum > eps, np.minimum(np.sqrt(((q - 2) ** 2 * rho * p) /
                                                       ((gamma - 1) * (2 * A1 + (q - 2) * rho ** 2 * A2))), 1), 1)
    return alpha, kappa, beta


def rough_state(shape, seed):
    """Steep density / pressure contrasts and strong shear, so that every
    factor drops below 1 somewhere."""
    rng = np.random.default_rng(seed)
    dim = len(shape)
    s = np.empty((dim + 2,) + shape)
    s[0] = np.exp(rng.normal(0, 1.5, shape))
    s[1:dim + 1] = rng.normal(0, 3.0, (dim,) + shape)
    s[dim + 1] = np.exp(rng.normal(0, 1.5, shape))
    return s


@pytest.mark.parametrize("dim,shape,mhd,C_cfl", [(2, (12, 10), False, 0.4), (2, (9, 14), True, 0.3),
                                                 (3, (8, 6, 10), False, 0.4), (3, (6, 7, 8), True, 0.25)])
def test_pp_factors_match_the_array_form(oracle, dim, shape, mhd, C_cfl):
    w = rough_state(shape, seed=dim + int(mhd))
    dx = 1.0 / shape[-1]
    oc = oracle.make_config((dim + 2 + (3 if mhd else 0),) + shape, dim, mhd=mhd, limiter=VAN_ALBADA_PP,
                            C_cfl=C_cfl)
    f = oracle.pp_factors(w, oc)
    alpha, kappa, beta = pp_factors_arrays(w, dim, dx, 5.0 / 3.0, C_cfl, mhd)
    np.testing.assert_allclose(f[:dim], alpha, rtol=1e-12, atol=0)
    np.testing.assert_allclose(f[dim:2 * dim], kappa, rtol=1e-12, atol=0)
    np.testing.assert_allclose(f[2 * dim], beta, rtol=1e-12, atol=0)
    # the state exercises every branch: factors below 1 for all three kinds
    assert (alpha < 1).any() and (kappa < 1).any() and (beta < 1).any()
    assert (alpha == 1).any() and (beta == 1).any()


def test_a1_follows_the_component_broadcast(oracle):
    """A1 weights component k by C[k] alpha[k] d_rho[k] (the broadcast), not
    by axis: with alpha_lax very different per axis the two readings give
    different betas, and the oracle matches the broadcast one."""
    dim, shape = 2, (10, 12)
    w = rough_state(shape, seed=7)
    w[1] *= 6.0  # alpha_lax[x] >> alpha_lax[y]
    dx = 1.0 / shape[-1]
    oc = oracle.make_config((4,) + shape, dim, mhd=False, limiter=VAN_ALBADA_PP, C_cfl=0.4)
    f = oracle.pp_factors(w, oc)
    _, _, beta = pp_factors_arrays(w, dim, dx, 5.0 / 3.0, 0.4, False)
    np.testing.assert_allclose(f[2 * dim], beta, rtol=1e-12)
    # the per-axis reading: sum_k (sum_a C[a] alpha[a] d_rho[a] d_v[a][k])^2
    d = va_differences(w, dim, dx)
    c = np.sqrt(5.0 / 3.0 * w[3] / w[0])
    al = np.array([np.max(np.abs(w[1 + a]) + c) for a in range(dim)])
    C = al / al.sum()
    alpha = f[:dim]
    A1_axis = sum(sum(C[a] * alpha[a] * d[a, 0] * d[a, 1 + k] for a in range(dim)) ** 2 for k in range(dim))
    A2 = sum(C[a] * sum(d[a, 1 + k] ** 2 for k in range(dim)) for a in range(dim))
    q = 1 / 0.4
    with np.errstate(invalid="ignore"):
        beta_axis = np.minimum(np.sqrt((q - 2) ** 2 * w[0] * w[3] / ((2 / 3) * (2 * A1_axis + (q - 2) * w[0] ** 2 * A2))), 1)
    assert np.abs(beta_axis - f[2 * dim]).max() > 1e-3


@pytest.mark.parametrize("dim,mhd", [(2, False), (3, True)])
def test_pp_equals_van_albada_where_no_factor_acts(oracle, dim, mhd):
    shape = (16, 12, 10)[:dim]
    grids = np.meshgrid(*[(np.arange(n) + 0.5) / n for n in shape], indexing="ij")
    nv = dim + 2
    w = np.zeros((nv,) + shape, dtype=np.float32)
    w[0] = 1.0 + 0.02 * np.sin(2 * np.pi * grids[0])
    for a in range(dim):
        w[1 + a] = 0.01 * np.cos(2 * np.pi * grids[-1 - a])
    w[dim + 1] = 1.0 + 0.02 * np.cos(2 * np.pi * grids[-1])
    full = (nv + (3 if mhd else 0),) + shape
    oc_pp = oracle.make_config(full, dim, mhd=mhd, limiter=VAN_ALBADA_PP, riemann=HLLC)
    oc_va = oracle.make_config(full, dim, mhd=mhd, limiter=VAN_ALBADA, riemann=HLLC)
    assert np.all(oracle.pp_factors(w, oc_pp) == 1)
    a = oracle.gas_stage(w, oc_pp, 1e-3)
    b = oracle.gas_stage(w, oc_va, 1e-3)
    np.testing.assert_array_equal(a, b)


def test_pp_faces_stay_non_negative(oracle):
    dim, shape = 2, (24, 20)
    w = rough_state(shape, seed=3)
    w[0][::3] *= 1e-3   # near-vacuum rows next to dense ones
    w[3][1::4] *= 1e-3
    dx = 1.0 / shape[-1]
    oc = oracle.make_config((4,) + shape, dim, mhd=False, limiter=VAN_ALBADA_PP, C_cfl=0.4)
    f = oracle.pp_factors(w, oc)
    d = va_differences(w, dim, dx)
    plain_min = min((w[v] - np.abs(d[a, v])).min() for a in range(dim) for v in (0, 3))
    assert plain_min < 0  # the unscaled van Albada faces go negative
    for a in range(dim):
        for v, row in ((0, a), (3, dim + a)):
            faces = w[v] - np.abs(d[a, v]) * f[row]
            assert faces.min() >= -1e-12 * w[v].max(), (a, v, faces.min())
