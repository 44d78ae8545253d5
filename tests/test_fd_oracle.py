"""Pins of the finite-difference CPU restatement (oracle/fd_oracle.cpp).

The reference holds no golden data for its FD path, so each check below is
fixed by the mathematics rather than by retyping a formula:

* eigensystem: L R = I and L A R = diag(lambda) with A the finite-difference
  Jacobian of the flux; the Jacobian's eigenvalues equal the closed-form
  magnetosonic / Alfven speeds
* WENO-5 interface fluxes: exact on uniform states, 5th-order convergence of
  the flux divergence on smooth data
* SSPRK(5,4): 4th-order temporal convergence on a linearised problem
* CT: the 6th-order face divergence is preserved to round-off, face sums are
  conserved, and the centre/face interpolants have their textbook orders
* whole step: uniform states are fixed points, mass/momentum are conserved,
  every permutation of the axes commutes with the step
* dt: closed form on uniform states and the dt_max clamp
"""

import numpy as np
import pytest

from astronomix_amd import workloads as W

GAMMA = 5.0 / 3.0


@pytest.fixture(scope="module")
def fd(oracle):
    from oracle import fd as _fd
    return _fd


def conserved(w, gamma=GAMMA):
    rho, v, p, B = w[0], w[1:4], w[4], w[5:8]
    E = p / (gamma - 1) + 0.5 * rho * (v ** 2).sum(0) + 0.5 * (B ** 2).sum(0)
    return np.ascontiguousarray(np.concatenate([rho[None], rho[None] * v, E[None], B]))


def random_cell(rng):
    return np.array([rng.uniform(0.5, 2.0), *rng.normal(0, 0.7, 3), rng.uniform(0.5, 2.0), *rng.normal(0, 1.0, 3)])


def fast_slow_alfven(w, gamma=GAMMA):
    rho, p, B = w[0], w[4], w[5:8]
    a2 = gamma * p / rho
    b2 = (B ** 2).sum() / rho
    bx2 = B[0] ** 2 / rho
    root = np.sqrt((a2 + b2) ** 2 - 4 * a2 * bx2)
    return np.sqrt(0.5 * (a2 + b2 + root)), np.sqrt(0.5 * (a2 + b2 - root)), np.sqrt(bx2)


def test_eigensystem_diagonalises_the_flux_jacobian(fd):
    c = fd.make_config((8, 8, 8))
    rng = np.random.default_rng(0)
    varying = [0, 1, 2, 3, 4, 6, 7]  # Bx is constant along x
    for _ in range(6):
        w = random_cell(rng)
        q = conserved(w[:, None])[:, 0].copy()
        L, R, lam = fd.eigensystem(q, q, c)
        assert np.abs(L @ R - np.eye(7)).max() < 1e-10
        A = np.zeros((8, 8))
        h = 1e-6
        for j in varying:
            dq = np.zeros(8)
            dq[j] = h
            A[:, j] = (fd.mhd_flux(q + dq, c) - fd.mhd_flux(q - dq, c)) / (2 * h)
        assert np.abs(L @ A @ R - np.diag(lam)).max() < 1e-6 * max(1.0, np.abs(lam).max())
        # the flux Jacobian's spectrum is the textbook MHD wave fan
        cf, cs, ca = fast_slow_alfven(w)
        vx = w[1]
        expect = np.array([vx - cf, vx - ca, vx - cs, vx, vx + cs, vx + ca, vx + cf])
        ev = np.sort(np.linalg.eigvals(A[np.ix_(varying, varying)]).real)
        assert np.allclose(ev, np.sort(expect), atol=1e-6)
        assert np.allclose(lam, expect, atol=1e-12)  # ordering: vx-cf, vx-ca, vx-cs, vx, ...


def test_weno_flux_exact_on_uniform_state(fd):
    w = np.zeros((8, 8, 8, 8))
    w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7] = 1.3, 0.4, -0.2, 0.1, 0.9, 0.3, -0.5, 0.2
    q = conserved(w)
    c = fd.make_config(q.shape)
    f = fd.mhd_flux(q[:, 0, 0, 0].copy(), c)
    for d in range(3):
        F = fd.weno_flux(q, c, d)
        perm = [[0, 1, 2, 3, 4, 5, 6, 7], [0, 2, 1, 3, 4, 6, 5, 7], [0, 3, 2, 1, 4, 7, 6, 5]][d]
        fd_ = fd.mhd_flux(q[perm, 0, 0, 0].copy(), c)
        expect = np.empty(8)
        expect[perm] = fd_
        assert np.abs(F - expect[:, None, None, None]).max() < 1e-13
    assert np.abs(fd.weno_flux(q, c, 0)[:, 0, 0, 0] - f).max() < 1e-13


def _smooth_prim(x):
    s = 2 * np.pi * x
    return np.array([1 + 0.2 * np.sin(s), 0.5 * np.sin(s + 1), 0.1 * np.cos(s), 0.05 * np.sin(2 * s),
                     1 + 0.1 * np.cos(s), 0.3 + 0 * x, 0.2 * np.sin(s + 0.3), 0.1 + 0.05 * np.cos(s)])


def test_weno_flux_divergence_is_fifth_order(fd):
    c0 = fd.make_config((8, 8, 8))
    N = 256  # the flux of the smooth state, sampled finely, as a Fourier series
    xf = (np.arange(N) + 0.5) / N
    fh = np.fft.fft(fd.mhd_flux(conserved(_smooth_prim(xf)), c0), axis=1) / N
    kk = np.fft.fftfreq(N, d=1.0 / N)

    def exact_derivative(x):
        return np.real((fh * (2j * np.pi * kk)) @ np.exp(2j * np.pi * np.outer(kk, x - 0.5 / N)))

    errs = []
    for n in (64, 128):
        x = (np.arange(n) + 0.5) / n
        q = np.ascontiguousarray(np.broadcast_to(conserved(_smooth_prim(x))[:, :, None, None], (8, n, 6, 6)))
        c = fd.make_config((n, 6, 6))
        c.dx = 1.0 / n
        F = fd.weno_flux(q, c, 0)
        div = ((F - np.roll(F, 1, axis=1)) / c.dx)[:, :, 0, 0]
        errs.append(np.abs(div - exact_derivative(x)).mean(axis=1))
    rows = [0, 1, 2, 3, 4, 6, 7]
    order = np.log2(errs[0][rows] / errs[1][rows])
    assert (order > 4.0).all(), order


def test_ssprk_is_fourth_order_in_time(fd):
    # small-amplitude hydro state: the update is linear to round-off, so only
    # the Runge-Kutta coefficients set the temporal order
    n = 12
    s = W.turbulence_3d(n, seed=3, bx=0.0, mach=1e-3, dtype=np.float64)
    s[0] += 1e-3 * np.sin(2 * np.pi * (np.arange(n) + 0.5) / n)[:, None, None]
    outs = {}
    for m in (4, 8, 16, 32):
        c = fd.make_config(s.shape, t_end=0.1, fixed_timestep=True, num_timesteps=m, dt_max=1.0)
        outs[m] = fd.time_integration(s, c)[0]
    e = [np.abs(outs[m] - outs[2 * m]).mean(axis=(1, 2, 3))[[0, 1, 4]] for m in (4, 8, 16)]
    o1, o2 = np.log2(e[0] / e[1]), np.log2(e[1] / e[2])
    assert ((o1 > 3.8) & (o1 < 4.5)).all(), o1
    assert ((o2 > 3.8) & (o2 < 4.3)).all(), o2


def _mhd_state(n, seed=5):
    s = W.turbulence_3d(n, seed=seed, bx=0.3, dtype=np.float64)
    s[5:8] += W.solenoidal_field(n, 1, 2, seed=seed + 2, rms=0.2)
    for a in range(3):
        s[8 + a] = W.center_to_face(s[5 + a], a)
    return s


def test_ct_preserves_face_divergence_and_face_sums(fd):
    n = 12
    s = _mhd_state(n)
    rng = np.random.default_rng(3)
    s[8:11] += 0.05 * rng.normal(size=(3, n, n, n))  # arbitrary, non-solenoidal faces
    c = fd.make_config(s.shape)
    dt = fd.cfl_dt(s, c)
    out = fd.evolve_state(s, c, dt)
    d0 = fd.face_divergence(s[8], s[9], s[10], c)
    d1 = fd.face_divergence(out[8], out[9], out[10], c)
    assert np.abs(d0).max() > 1e-2
    assert np.abs(d1 - d0).max() < 1e-13
    assert np.abs(out[8:11] - s[8:11]).max() > 1e-4  # the faces did move
    assert np.allclose(out[8:11].sum(axis=(1, 2, 3)), s[8:11].sum(axis=(1, 2, 3)), rtol=0, atol=1e-12)


def test_step_conserves_mass_and_momentum(fd):
    s = _mhd_state(12)
    c = fd.make_config(s.shape)
    out = fd.evolve_state(s, c, fd.cfl_dt(s, c))
    m0 = np.array([s[0].sum(), *(s[0] * s[1:4]).sum(axis=(1, 2, 3))])
    m1 = np.array([out[0].sum(), *(out[0] * out[1:4]).sum(axis=(1, 2, 3))])
    assert np.abs(m1 - m0).max() < 1e-11 * s[0].sum()
    assert np.abs(out[:8] - s[:8]).max() > 1e-4


def test_uniform_state_is_a_fixed_point(fd):
    s = np.zeros((11, 8, 8, 8))
    vals = [1.2, 0.3, -0.4, 0.25, 0.8, 0.35, -0.2, 0.6]
    for v in range(8):
        s[v] = vals[v]
    s[8:11] = s[5:8]
    c = fd.make_config(s.shape)
    out = fd.evolve_state(s, c, 0.01)
    assert np.abs(out - s).max() < 1e-13


@pytest.mark.parametrize("p", [(1, 0, 2), (0, 2, 1), (2, 1, 0), (1, 2, 0), (2, 0, 1)])
def test_axis_permutation_commutes_with_the_step(fd, p):
    def perm_state(w):
        rows = [0] + [1 + p[d] for d in range(3)] + [4] + [5 + p[d] for d in range(3)] + [8 + p[d] for d in range(3)]
        return np.ascontiguousarray(np.transpose(w[rows], (0,) + tuple(1 + p[d] for d in range(3))))

    s = _mhd_state(12)
    c = fd.make_config(s.shape)
    dt = fd.cfl_dt(s, c)
    a = fd.evolve_state(perm_state(s), c, dt)
    b = perm_state(fd.evolve_state(s, c, dt))
    assert np.abs(a - b).max() < 1e-13


def _order(errs, ns):
    return np.log2(errs[0] / errs[1]) / np.log2(ns[1] / ns[0])


def test_centre_face_interpolants_and_divergence_have_textbook_orders(fd):
    ns = (16, 32)
    e_c2f, e_f2c, e_d6 = [], [], []
    for n in ns:
        h = 1.0 / n
        xc = (np.arange(n) + 0.5) * h
        s = np.zeros((11, n, n, n))
        s[0], s[4] = 1.0, 1.0
        f = lambda x: np.sin(2 * np.pi * x + 0.4)
        X = xc[:, None, None] * np.ones((1, n, n))
        Y = np.transpose(X, (1, 0, 2))
        Z = np.transpose(X, (1, 2, 0))
        B = np.stack([f(X), f(Y), f(Z)])
        c = fd.make_config(s.shape)
        b = fd.interface_fields(np.ascontiguousarray(B), c)
        e_c2f.append(np.abs(b[0] - f(X + 0.5 * h)).max())
        # faces -> centres: evolve by dt = 0 rebuilds B at centres from the faces
        s[5:8] = 0.0
        s[8] = f(X + 0.5 * h)
        s[9] = f(Y + 0.5 * h)
        s[10] = f(Z + 0.5 * h)
        out = fd.evolve_state(s, c, 0.0)
        e_f2c.append(np.abs(out[5] - f(X)).max())
        div = fd.face_divergence(s[8], s[9], s[10], c) / h
        exact = 2 * np.pi * (np.cos(2 * np.pi * X + 0.4) + np.cos(2 * np.pi * Y + 0.4) + np.cos(2 * np.pi * Z + 0.4))
        e_d6.append(np.abs(div - exact).max())
    assert abs(_order(e_c2f, ns) - 4) < 0.2, e_c2f
    assert abs(_order(e_f2c, ns) - 6) < 0.3, e_f2c
    assert abs(_order(e_d6, ns) - 6) < 0.3, e_d6


def test_cfl_dt_closed_form_and_clamp(fd):
    n = 8
    s = np.zeros((11, n, n, n))
    s[0], s[1], s[4], s[5], s[8] = 1.0, 0.2, 1.0, 0.5, 0.5
    cs = np.sqrt(GAMMA)
    lx = 0.2 + max(cs, 0.5)  # fast speed along B equals max(c_s, c_A)
    lt = np.sqrt(cs ** 2 + 0.25)  # perpendicular to B: sqrt(c_s^2 + c_A^2)
    c = fd.make_config(s.shape, C_cfl=0.4, dt_max=1.0)
    dt = fd.cfl_dt(s, c)
    assert dt == pytest.approx(0.4 * (1.0 / n) / (lx + 2 * lt), rel=1e-12)
    c = fd.make_config(s.shape, C_cfl=0.4, dt_max=1e-3)
    assert fd.cfl_dt(s, c) == pytest.approx(1e-3, rel=1e-15)


def test_time_loop_hits_t_end_exactly(fd):
    s = _mhd_state(8)
    c = fd.make_config(s.shape, C_cfl=0.4, t_end=0.0123, dt_max=0.005)
    out, steps, t = fd.time_integration(s, c)
    assert t == pytest.approx(0.0123, rel=0, abs=1e-15)
    assert steps == 3  # dt_max binds: 0.005 + 0.005 + 0.0023
    c = fd.make_config(s.shape, t_end=0.01, fixed_timestep=True, num_timesteps=4)
    _, steps, t = fd.time_integration(s, c)
    assert steps == 4 and t == pytest.approx(0.01, abs=1e-15)


def test_float32_tracks_float64(fd):
    s = _mhd_state(12)
    c = fd.make_config(s.shape, t_end=0.004, dt_max=0.001)
    o64 = fd.time_integration(s, c)[0]
    o32 = fd.time_integration(s.astype(np.float32), c)[0]
    rel = np.abs(o32 - o64).sum(axis=(1, 2, 3)) / np.maximum(np.abs(o64).sum(axis=(1, 2, 3)), 1e-30)
    assert rel.max() < 2e-5, rel


def ct_scaled_error(fd):
    """rhs_bx / dt + V b' of B_x = b(z) advected by v = (0, 0, V) (B_z = B0),
    as a multiple of the h^2 term -V h^2 b_zzz / 24, at n = 32, 64, 128."""
    V, B0, A = 0.7, 0.5, 0.2
    ratios = []
    for n in (32, 64, 128):
        h = 1.0 / n
        z = (np.arange(n) + 0.5) * h
        k = 2 * np.pi
        b = A * np.sin(k * z)
        shape = (6, 6, n)
        q = np.zeros((8,) + shape)
        q[0] = 1.0
        q[3] = V
        q[5] = b
        q[7] = B0
        q[4] = 1.0 / (GAMMA - 1) + 0.5 * V * V + 0.5 * (b ** 2 + B0 ** 2)
        faces = np.zeros((3,) + shape)
        faces[0] = b
        faces[2] = B0
        c = fd.make_config(shape)
        _, rb = fd.compute_rhs(q, faces, c, 1.0, 1.0)
        err = rb[0] + V * A * k * np.cos(k * z)
        lead = -V * h * h * (-A * k ** 3 * np.cos(k * z)) / 24.0
        ratios.append(float((err * lead).sum() / (lead * lead).sum()))
        # nothing but B_x changes: the y and z faces stay
        assert np.abs(rb[1]).max() < 1e-12 and np.abs(rb[2]).max() < 1e-12
    return ratios


def test_ct_point_to_average_weight_is_one_24th(fd):
    """The CT edge EMFs are converted from point values to edge averages by
    q + (delta^2_A q + delta^2_B q) / 24 (_interpolate.py:38-62,
    _constrained_transport.py:140-142) before the 6th-order difference
    (_differencing.py:8-25). The finite-difference WENO face fluxes are
    primitive-function values, f - h^2 f_xx / 24 + O(h^4) for a smooth flux
    f (their differences give f_x to 5th order), so the edge EMFs carry an
    O(h^2) term that the conversion must cancel: only with the weight 1/24
    is the face update of an advected B_x = b(z) (v = (0, 0, V), B_z = B0)
    4th-order accurate; any other weight leaves an O(h^2) error whose size
    is proportional to (24 w - 1). (Planted weights 1/12, 1/48 and 0 give a
    scaled error of +36, -18 and -36 in the units below, constant in h:
    second order.)"""
    ratios = ct_scaled_error(fd)
    # the error in units of h^2 b_zzz vanishes as h -> 0 ...
    assert abs(ratios[-1]) < 0.01, ratios
    # ... at 4th order: the h^2-scaled error falls with h^2 at the fine end
    assert abs(ratios[-1]) < abs(ratios[-2]) / 2.5, ratios
