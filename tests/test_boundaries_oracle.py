"""Pins the non-periodic boundaries of the CPU restatement
(oracle/fv_oracle.cpp, _geometry/boundaries.py:74-360) on the CPU:

* reflective walls are the mirror half of a doubled periodic domain (the
  mirror image with the normal velocity negated evolves into the mirror
  image), in 2D and 3D, through whole runs;
* open boundaries keep a uniform flow uniform, and cells out of the
  boundary's reach evolve exactly as with periodic boundaries;
* the magnetic update's ghost cells against an explicitly padded restatement
  of its first iterate (numpy rolls on arrays padded as boundaries.py pads
  them, including the reflective sign flip of row `axis` of the magnetic and
  velocity arrays, _set_specific_var_along_axis)."""

import numpy as np
import pytest

PER, OPEN, REFL = 0, 1, 2


def _smooth_hydro(shape, seed):
    rng = np.random.default_rng(seed)
    dim = len(shape)
    grids = np.meshgrid(*[(np.arange(n) + 0.5) / n for n in shape], indexing="ij")
    s = np.zeros((dim + 2,) + shape)
    ph = rng.uniform(0, 2 * np.pi, 8)
    s[0] = 1.0 + 0.3 * np.sin(2 * np.pi * grids[0] + ph[0]) * np.cos(2 * np.pi * grids[-1] + ph[1])
    for a in range(dim):
        s[1 + a] = 0.4 * np.sin(2 * np.pi * (grids[a] + grids[-1 - a]) + ph[2 + a])
    s[dim + 1] = 1.0 + 0.2 * np.cos(2 * np.pi * grids[0] + ph[6])
    return s


@pytest.mark.parametrize("dim,axis", [(2, 0), (2, 1), (3, 1), (3, 2)])
def test_reflective_wall_is_the_mirror_half_of_a_periodic_domain(oracle, dim, axis):
    shape = (12, 10, 8)[:dim]
    s = _smooth_hydro(shape, seed=axis)
    # doubled domain along `axis`: the mirror image with the normal velocity negated
    mirror = np.flip(s, axis=1 + axis).copy()
    mirror[1 + axis] *= -1
    doubled = np.concatenate([s, mirror], axis=1 + axis)
    bc = [PER] * 6
    bc[2 * axis] = bc[2 * axis + 1] = REFL
    n_last = shape[-1]
    kw = dict(mhd=False, riemann=1, C_cfl=0.4, t_end=0.1)
    # the oracle's dx is box / (last axis): keep it equal in both runs
    oc_half = oracle.make_config(s.shape, dim, box_size=1.0, bc=bc, **kw)
    oc_dbl = oracle.make_config(doubled.shape, dim, box_size=doubled.shape[-1] / n_last, **kw)
    half, steps, _, _ = oracle.time_integration(s, oc_half)
    full, steps2, _, _ = oracle.time_integration(doubled, oc_dbl)
    assert steps == steps2 and steps > 3
    sl = [slice(None)] * (dim + 1)
    sl[1 + axis] = slice(0, shape[axis])
    np.testing.assert_allclose(half, full[tuple(sl)], rtol=0, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_open_boundaries_keep_uniform_flow(oracle, dim):
    shape = (10, 12, 8)[:dim]
    s = np.zeros((dim + 5,) + shape)
    s[0], s[dim + 1] = 1.3, 0.8
    s[1:dim + 1] = np.array([0.4, -0.3, 0.2][:dim]).reshape((dim,) + (1,) * dim)
    s[dim + 2:] = np.array([0.5, 0.2, -0.1]).reshape((3,) + (1,) * dim)
    oc = oracle.make_config(s.shape, dim, mhd=True, riemann=0, bc=[OPEN] * 6, t_end=0.05)
    out, steps, _, _ = oracle.time_integration(s, oc)
    assert steps > 3
    np.testing.assert_allclose(out, s, rtol=1e-13, atol=1e-13)


def test_open_boundary_does_not_reach_the_interior_early(oracle):
    shape = (24, 24)
    s = _smooth_hydro(shape, seed=3)
    n_steps = 3
    kw = dict(mhd=False, riemann=0, fixed_timestep=True, num_timesteps=n_steps, t_end=0.006)
    per, _, _, _ = oracle.time_integration(s, oracle.make_config(s.shape, 2, **kw))
    opn, _, _, _ = oracle.time_integration(s, oracle.make_config(s.shape, 2, bc=[OPEN] * 4 + [PER] * 2, **kw))
    # an RK2 step reaches 4 cells (two stages of the 2-cell stencil)
    r = 4 * n_steps
    inner = (slice(None), slice(r, -r), slice(r, -r))
    np.testing.assert_array_equal(opn[inner], per[inner])
    assert np.abs(opn - per).max() > 1e-6


def _pad(arr, dim, bc, flip_rows, ng=2):
    """Pad the spatial axes as boundaries.py does: open copies the edge,
    reflective mirrors and negates row flip_rows(axis) (if any), periodic
    wraps; axes in order x, y, z, the periodic ones last."""
    out = arr.copy()
    pads = [(0, 0)] + [(ng, ng)] * dim
    out = np.pad(out, pads, mode="edge")
    def take(a, idx, ax):
        return np.take(a, idx, axis=ax)
    for a in range(dim):
        ax = 1 + a
        n = arr.shape[ax]
        lo, hi = bc[2 * a], bc[2 * a + 1]
        for side, b in ((0, lo), (1, hi)):
            if b == PER:
                continue
            for i in range(ng):
                if side == 0:
                    set_i, get_i = (ng - i - 1, ng + i) if b == REFL else (i, ng)
                else:
                    set_i, get_i = (ng + n + i, ng + n - 1 - i) if b == REFL else (ng + n + i, ng + n - 1)
                sl_set = [slice(None)] * out.ndim
                sl_set[ax] = set_i
                out[tuple(sl_set)] = take(out, get_i, ax)
                r = flip_rows(a)
                if b == REFL and r is not None:
                    sl_r = list(sl_set)
                    sl_r[0] = r
                    out[tuple(sl_r)] *= -1
    for a in range(dim):
        ax = 1 + a
        n = arr.shape[ax]
        if bc[2 * a] == PER and bc[2 * a + 1] == PER:
            for i in range(ng):
                for set_i, get_i in ((i, i + n), (ng + n + i, ng + i)):
                    sl = [slice(None)] * out.ndim
                    sl[ax] = set_i
                    out[tuple(sl)] = take(out, get_i, ax)
    return out


def _curl(f, dx, dim):
    def d(k, ax):
        return 0.5 * (np.roll(f[k], -1, ax) - np.roll(f[k], 1, ax)) / dx
    if dim == 3:
        return np.stack([d(2, 1) - d(1, 2), d(0, 2) - d(2, 0), d(1, 0) - d(0, 1)])
    return np.stack([d(2, 1), -d(2, 0), d(1, 0) - d(0, 1)])


@pytest.mark.parametrize("dim,bc", [
    (2, [REFL, OPEN, PER, PER]), (2, [OPEN, REFL, REFL, OPEN]),
    (3, [REFL, REFL, OPEN, REFL, REFL, OPEN]), (3, [PER, PER, REFL, REFL, OPEN, OPEN]),
])
def test_magnetic_first_iterate_against_padded_arrays(oracle, dim, bc):
    rng = np.random.default_rng(5)
    shape = (10, 9, 8)[:dim]
    B = rng.normal(0, 1, (3,) + shape)
    gas = np.zeros((dim + 2,) + shape)
    gas[0] = rng.uniform(0.5, 2, shape)
    gas[1:dim + 1] = rng.normal(0, 0.5, (dim,) + shape)
    gas[dim + 1] = 1.0
    dx, dt = 1.0 / shape[-1], 1e-3
    oc = oracle.make_config((dim + 5,) + shape, dim, mhd=True, riemann=0, bc=bc, max_iter=0)
    B1, _, _ = oracle.magnetic_update(B, gas, oc, dt)
    # the reference: pad B and v (row `axis` + 1 of the 3-row arrays negated
    # at reflective walls: By/vy at x, Bz/vz at y, none at z), E = B x v,
    # B1 = B - dt curl(E) on the padded arrays
    v = np.zeros((3,) + shape)
    v[:dim] = gas[1:dim + 1]
    rows = lambda a: a + 1 if a + 1 <= 2 else None
    Bp = _pad(B, dim, bc, rows)
    vp = _pad(v, dim, bc, rows)
    E = np.cross(Bp, vp, axis=0)
    B1p = Bp - dt * _curl(E, dx, dim)
    inner = (slice(None),) + (slice(2, -2),) * dim
    np.testing.assert_allclose(B1, B1p[inner], rtol=1e-12, atol=1e-12)
