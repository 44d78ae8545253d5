"""Pins the snapshot restatement (oracle/snapshots.py) on the CPU: closed
forms for uniform states and sinusoidal fields, loop-built radial fields for
the orientation of the reference's cell centres, the independent C++ face
divergence, and schedules derived by hand from the rules of
time_integration.py:539-610, 671-678, 736-745."""

import math

import numpy as np
import pytest

from oracle import snapshots as S

GAMMA = 5.0 / 3.0


def _uniform(dim, n, rho, v, p, mhd=False):
    nv = dim + 2 + (3 if mhd else 0)
    s = np.zeros((nv,) + (n,) * dim, np.float32)
    s[0] = rho
    for a in range(dim):
        s[1 + a] = v[a]
    s[dim + 1] = p
    return s


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_state_totals(dim):
    n, box = 8, 2.0
    dx = box / n
    rho, v, p = 1.5, (0.5, -0.25, 0.125)[:dim], 0.7
    d = S.diagnostics(_uniform(dim, n, rho, v, p), dim=dim, dx=dx, box_size=box, gamma=GAMMA)
    v2 = sum(c * c for c in v)
    vol = box ** dim
    # the reference multiplies the summed density by box_size**dim (total_quantities.py:176)
    assert d["total_mass"] == pytest.approx(rho * n ** dim * box ** dim, rel=1e-6)
    assert d["internal_energy"] == pytest.approx(p / (GAMMA - 1) * vol, rel=1e-6)
    assert d["kinetic_energy"] == pytest.approx(0.5 * rho * v2 * vol, rel=1e-6)
    assert d["total_energy"] == pytest.approx((p / (GAMMA - 1) + 0.5 * rho * v2) * vol, rel=1e-6)
    # a uniform flow has no net radial momentum about the box centre (even n: symmetric centres)
    assert abs(d["radial_momentum"]) < 1e-5 * rho * math.sqrt(v2) * vol


def test_radial_momentum_of_odd_grid_is_nan():
    # a cell centre on the box centre has r_hat = 0 / 0 (total_quantities.py:52-54)
    d = S.diagnostics(_uniform(2, 5, 1.0, (0.1, 0.2), 1.0), dim=2, dx=0.2, box_size=1.0)
    assert math.isnan(d["radial_momentum"])


def test_meshgrid_default_indexing_is_xy():
    # library fact the reference relies on: component 0 varies along axis 1
    g = S.geometric_centres(3, 4, 0.25, 1.0)
    x = S.cell_centres(4, 0.25, 1.0)
    assert np.array_equal(g[1, 2, 3], np.array([x[2], x[1], x[3]], np.float32))


@pytest.mark.parametrize("dim", [2, 3])
def test_radial_momentum_orientation(dim):
    """A velocity equal to the reference's r_hat, built by explicit loops with
    component 0 from the coordinate of axis 1, has radial momentum
    sum(rho) dx^dim; the same field with axes 0 and 1 in their natural roles
    does not."""
    n, box = 6, 1.0
    dx = box / n
    c = [(i + 0.5) * dx - box / 2 for i in range(n)]
    s = _uniform(dim, n, 2.0, (0.0,) * dim, 1.0)
    nat = s.copy()
    for idx in np.ndindex(*(n,) * dim):
        o_quirk = [c[idx[1]], c[idx[0]]] + ([c[idx[2]]] if dim == 3 else [])
        o_nat = [c[i] for i in idx]
        rq = math.sqrt(sum(x * x for x in o_quirk))
        rn = math.sqrt(sum(x * x for x in o_nat))
        for a in range(dim):
            s[(1 + a,) + idx] = o_quirk[a] / rq
            nat[(1 + a,) + idx] = o_nat[a] / rn
    want = 2.0 * n ** dim * dx ** dim
    got = S.diagnostics(s, dim=dim, dx=dx, box_size=box)["radial_momentum"]
    assert got == pytest.approx(want, rel=1e-6)
    got_nat = S.diagnostics(nat, dim=dim, dx=dx, box_size=box)["radial_momentum"]
    assert got_nat < 0.95 * want


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_central_divergence_closed_form(axis):
    n = 32
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx
    B = np.zeros((3, n, n, n), np.float32)
    shape = [1, 1, 1]
    shape[axis] = n
    B[axis] = np.sin(2 * np.pi * x).reshape(shape)
    div = S.central_divergence(B, dx)
    # (sin(k(x+dx)) - sin(k(x-dx))) / (2 dx) = cos(k x) sin(k dx) / dx
    want = (np.cos(2 * np.pi * x) * np.sin(2 * np.pi * dx) / dx).reshape(shape)
    np.testing.assert_allclose(div, np.broadcast_to(want, div.shape), atol=2e-4 * 2 * np.pi)
    s = np.zeros((8, n, n, n), np.float32)
    s[0], s[4] = 1.0, 1.0
    s[5:8] = B
    d = S.diagnostics(s, dim=3, mag_row=5, div_mode=S.DIV_CENTRAL, dx=dx, box_size=1.0)
    assert d["magnetic_divergence"] == pytest.approx(np.abs(want).max(), rel=1e-4)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_face_divergence_closed_form(axis):
    n = 16
    dx = 1.0 / n
    k = 2 * np.pi
    xc = (np.arange(n) + 0.5) * dx
    shape = [1, 1, 1]
    shape[axis] = n
    b = np.zeros((3, n, n, n), np.float32)
    b[axis] = np.sin(k * (xc + dx / 2)).reshape(shape)  # index i holds face i + 1/2
    div = S.face_divergence(b, dx)
    c1, c2, c3 = 75 / 64, -25 / 384, 3 / 640
    g = 2 * (c1 * np.sin(k * dx / 2) + c2 * np.sin(3 * k * dx / 2) + c3 * np.sin(5 * k * dx / 2)) / dx
    want = (np.cos(k * xc) * g).reshape(shape)
    np.testing.assert_allclose(div, np.broadcast_to(want, div.shape), atol=1e-4 * k)
    # sixth order: the stencil's gain is k to ~(k dx)^6
    assert g == pytest.approx(k, rel=1e-5)


@pytest.fixture(scope="module")
def fd_oracle(oracle):
    from oracle import fd as _fd
    return _fd


def test_face_divergence_matches_cpp_restatement(fd_oracle):
    rng = np.random.default_rng(3)
    b = rng.standard_normal((3, 10, 12, 8)).astype(np.float32)
    c = fd_oracle.make_config((11, 10, 12, 8))
    # the C++ function returns the undivided stencil sum: compare at dx = 1
    np.testing.assert_array_equal(S.face_divergence(b, 1.0), fd_oracle.face_divergence(b[0], b[1], b[2], c))


def _const_loop(dt, **kw):
    return S.time_integration_snapshots(np.zeros(1, np.float32), lambda s, d: s, lambda s: dt, **kw)


def test_schedule_uniform_rows_at_step_start():
    # steps start at t = 0, .25, .5, .75; thresholds k/3: rows at 0, .5, .75
    r = _const_loop(0.25, t_end=1.0, num_snapshots=3)
    assert list(r["time_points"]) == [0.0, 0.5, 0.75]
    assert (r["num_iterations"], r["current_checkpoint"]) == (4, 3)


def test_schedule_one_row_per_step_and_no_end_row():
    # dt 0.6: steps start at 0 and 0.6; threshold 0.5 is passed at 0.6 but
    # 0.75 never comes before t_end, so two of four rows stay zero
    r = _const_loop(0.6, t_end=1.0, num_snapshots=4)
    assert list(r["time_points"]) == [0.0, 0.6000000238418579, 0.0, 0.0]
    assert (r["num_iterations"], r["current_checkpoint"]) == (2, 2)


def test_schedule_timepoints_clamp_and_end_row():
    # dt 0.25 clamped to the next time point; the last row is taken at t_end
    r = _const_loop(0.25, t_end=1.0, num_snapshots=3, timepoints=[0.0, 0.375, 1.0])
    assert list(r["time_points"]) == [0.0, 0.375, 1.0]
    assert (r["num_iterations"], r["current_checkpoint"]) == (5, 3)


def test_schedule_timepoints_fixed_steps():
    # fixed steps: 4 steps of 0.25 with the clamp, the loop stops at 0.875
    r = _const_loop(None, t_end=1.0, num_snapshots=3, timepoints=[0.0, 0.375, 1.0], fixed_timestep=True,
                    num_timesteps=4)
    assert list(r["time_points"]) == [0.0, 0.375, 0.0]
    assert (r["num_iterations"], r["current_checkpoint"], r["t"]) == (4, 2, 0.875)


def test_schedule_callback_mode_uses_uniform_rule_without_clamp():
    r = _const_loop(0.25, t_end=1.0, num_snapshots=2, timepoints=[0.0, 0.3], record=False)
    # no clamp, no end clamp (time points given), thresholds 0 and 0.5
    assert r["callbacks"] == [(0.0, 0), (0.5, 1)]
    assert r["num_iterations"] == 4
