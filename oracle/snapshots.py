"""CPU restatement of the snapshot rows and the snapshot schedule of
``time_integration`` (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ and __graft_entry__.smoke as the
checker; the product path never calls it.

* ``diagnostics`` -- one row of SnapshotData's quantities for a state:
  ``calculate_total_mass`` / ``_energy`` / ``_internal_energy`` /
  ``_kinetic_energy`` / ``_radial_momentum``
  (``_fluid_equations/total_quantities.py:29-176``) and max |div B|
  (``time_stepping/time_integration.py:507-522``) with ``divergence3D``
  (``_finite_volume/_magnetic_update/_vector_maths.py:33-51``) or
  ``_interface_field_divergence``
  (``_finite_difference/_maths/_differencing.py:9-52``). Per-cell terms in
  float32 in the reference's operation order; sums in float64.
* ``geometric_centres`` -- ``get_helper_data`` Cartesian centres
  (``data_classes/simulation_helper_data.py:92-131``): ``jnp.meshgrid`` with
  its default 'xy' indexing, which puts the coordinate of array axis 1 in
  component 0.
* ``time_integration_snapshots`` -- the loop of ``_time_integration`` with
  snapshots (``time_integration.py:404-610`` rows and schedule, ``:629-678``
  dt and clamps, ``:736-745`` end-time row, ``:756-801`` loop kinds), with the
  step and the CFL step supplied by the caller (e.g. ``oracle.fv``).

Pinned by closed forms and brute force in tests/test_snapshot_oracle.py
(uniform states, sinusoidal fields with known differences, loop-built radial
fields, hand-derived schedules) and, for the face divergence, by the
independent C++ restatement ``oracle.fd.face_divergence``.
"""

from __future__ import annotations

import numpy as np

F32 = np.float32

DIV_NONE, DIV_CENTRAL, DIV_FACES = 0, 1, 2


def cell_centres(n: int, dx: float, box_size: float) -> np.ndarray:
    """jnp.linspace(dx/2, box_size + dx/2, n, endpoint=False) in float32
    (simulation_helper_data.py:94-99)."""
    return np.linspace(dx / 2, box_size + dx / 2, n, endpoint=False).astype(F32)


def geometric_centres(dim: int, n: int, dx: float, box_size: float) -> np.ndarray:
    """(n,)*dim + (dim,) centres from np.meshgrid's default 'xy' indexing,
    as jnp.meshgrid(x, y[, z]) then moveaxis(0, -1) (simulation_helper_data.py:120-127)."""
    x = cell_centres(n, dx, box_size)
    return np.moveaxis(np.array(np.meshgrid(*([x] * dim))), 0, -1)


def _roll(a, s, axis):
    return np.roll(a, s, axis=axis)


def central_divergence(B: np.ndarray, dx: float) -> np.ndarray:
    """divergence3D: sum_d (roll(B_d, -1, d) - roll(B_d, 1, d)) / (2 dx) in float32."""
    den = F32(2 * dx)
    d = [(_roll(B[a], -1, a) - _roll(B[a], 1, a)) / den for a in range(3)]
    return (d[0] + d[1]) + d[2]


def int6(f: np.ndarray, axis: int) -> np.ndarray:
    """finite_difference_int6 (_differencing.py:9-30); index i holds face i + 1/2."""
    c1, c2, c3 = F32(75.0 / 64.0), F32(-25.0 / 384.0), F32(3.0 / 640.0)
    return (c1 * (f - _roll(f, 1, axis)) + c2 * (_roll(f, -1, axis) - _roll(f, 2, axis))) \
        + c3 * (_roll(f, -2, axis) - _roll(f, 3, axis))


def face_divergence(b: np.ndarray, dx: float) -> np.ndarray:
    """_interface_field_divergence: sum_d int6(b_d, d) / dx in float32."""
    d = [int6(b[a], a) / F32(dx) for a in range(3)]
    return (d[0] + d[1]) + d[2]


def _sum64(a: np.ndarray) -> float:
    return float(np.sum(a.astype(np.float64)))


def diagnostics(state, *, dim, rho_row=0, vel_row=1, p_row=None, mag_row=None, div_mode=DIV_NONE,
                gamma=5.0 / 3.0, dx, box_size) -> dict:
    """Every quantity of one snapshot row (float values)."""
    s = np.asarray(state, dtype=F32)
    p_row = dim + 1 if p_row is None else p_row
    rho = s[rho_row]
    p = s[p_row]
    v = [s[vel_row + a] for a in range(dim)]
    vol = F32(dx ** dim)
    out = {"total_mass": float(F32(_sum64(s[0])) * F32(box_size ** dim))}
    # get_absolute_velocity (_equations.py:165-197): +1e-20 in 3D only
    usq = v[0] * v[0] + v[1] * v[1]
    if dim == 3:
        usq = (usq + v[2] * v[2]) + F32(1e-20)
    u = np.sqrt(usq)
    kin = (F32(0.5) * rho) * (u * u)
    inte = p / (F32(gamma) - F32(1.0))
    out["internal_energy"] = _sum64(inte * vol)
    out["kinetic_energy"] = _sum64(kin * vol)
    out["total_energy"] = _sum64((inte + kin) * vol)
    n = s.shape[1]
    if all(m == n for m in s.shape[1:1 + dim]):
        off = geometric_centres(dim, n, dx, box_size) - F32(box_size / 2)
        sq = off[..., 0] * off[..., 0] + off[..., 1] * off[..., 1]
        if dim == 3:
            sq = sq + off[..., 2] * off[..., 2]
        nrm = np.sqrt(sq)
        with np.errstate(invalid="ignore", divide="ignore"):  # odd grids: 0 / 0 at the centre, as the reference
            ur = v[0] * (off[..., 0] / nrm) + v[1] * (off[..., 1] / nrm)
            if dim == 3:
                ur = ur + v[2] * (off[..., 2] / nrm)
        out["radial_momentum"] = _sum64((rho * ur) * vol)
    if div_mode == DIV_CENTRAL:
        out["magnetic_divergence"] = float(np.max(np.abs(central_divergence(s[mag_row:mag_row + 3], dx))))
    elif div_mode == DIV_FACES:
        out["magnetic_divergence"] = float(np.max(np.abs(face_divergence(s[mag_row:mag_row + 3], dx))))
    return out


def time_integration_snapshots(state, step, cfl_dt, *, t_end, num_snapshots, fixed_timestep=False,
                               num_timesteps=0, exact_end_time=True, timepoints=None, record=True,
                               diag=None, max_steps=None):
    """The reference loop with snapshots on (record=True) or with a snapshot
    callback (record=False: rows are only listed). ``step(state, dt)``
    returns the next state; ``cfl_dt(state)`` the CFL step. ``diag(state)``
    (optional) returns the quantities of a row. Time, dt and the schedule
    are float32, as in the reference. Returns a dict with time_points,
    states, diagnostics, num_iterations, current_checkpoint, final_state."""
    n = int(num_snapshots)
    t_end = F32(t_end)
    use_tp = timepoints is not None and record
    tps = [F32(v) for v in timepoints] if timepoints is not None else None
    rows = {"time_points": np.zeros(n, F32), "states": [None] * n, "diagnostics": [None] * n}
    cc = 0

    def take(t, st):
        nonlocal cc
        if record and cc < n:  # out-of-bounds .at[].set is dropped
            rows["time_points"][cc] = t
            rows["states"][cc] = np.array(st, copy=True)
            if diag is not None:
                rows["diagnostics"][cc] = diag(st)
        elif not record:
            rows.setdefault("callbacks", []).append((float(t), cc))
        cc += 1

    def tp_at(k):
        return tps[min(k, len(tps) - 1)]  # a JAX gather clamps its index

    t = F32(0.0)
    steps = 0
    while True:
        if fixed_timestep:
            if steps >= num_timesteps:
                break
        elif not t < t_end:
            break
        if max_steps is not None and steps >= max_steps:
            break
        # row at the start of the step (time_integration.py:539-566, 598-606)
        if use_tp:
            if abs(F32(t - tp_at(cc))) < F32(1e-12):
                take(t, state)
        elif t >= F32(F32(F32(cc) * t_end) / F32(n)):
            take(t, state)
        dt = F32(t_end / F32(num_timesteps)) if fixed_timestep else F32(cfl_dt(state))
        if use_tp:
            dt = min(dt, F32(tp_at(cc) - t))
        if exact_end_time and timepoints is None:
            dt = min(dt, F32(t_end - t))
        state = step(state, dt)
        t = F32(t + dt)
        steps += 1
        if use_tp and abs(F32(t - t_end)) < F32(1e-12):
            take(t, state)
    rows["num_iterations"] = steps
    rows["current_checkpoint"] = cc
    rows["final_state"] = state
    rows["t"] = float(t)
    return rows
