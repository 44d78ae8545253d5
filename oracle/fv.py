"""ctypes front of the CPU restatement in ``oracle/fv_oracle.cpp``.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke``
and by the CPU-baseline leg of ``bench.py`` -- as the checker, never as the
thing measured or shipped. The product path (``astronomix_amd``) never imports
this module.

Each function mirrors one reference function (see the C++ header for the
file:line map): ``time_integration`` follows
``astronomix/time_stepping/time_integration.py:629-801`` (no snapshots),
``evolve_state_fv`` follows ``_finite_volume/_state_evolution/evolve_state.py:496-574``,
``cfl_dt`` follows ``_finite_volume/_timestep_estimation/_timestep_estimator.py:118-128``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfv_oracle.so")
_lib = None


class OracleConfig(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("mhd", ctypes.c_int),
        ("riemann", ctypes.c_int),
        ("euler", ctypes.c_int),
        ("fixed_timestep", ctypes.c_int),
        ("num_timesteps", ctypes.c_int),
        ("exact_end_time", ctypes.c_int),
        ("max_iter", ctypes.c_int),
        ("limiter", ctypes.c_int),
        ("bc", ctypes.c_int * 6),
        ("gamma", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("C_cfl", ctypes.c_double),
        ("t_end", ctypes.c_double),
        ("tol", ctypes.c_double),
    ]


def build() -> str:
    """Compile the restatement with its Makefile (g++ -fopenmp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def use_library(path=None):
    """Bind the functions below to the restatement built at `path` (None: the
    default build). The mutation tests bind copies with planted mistakes."""
    global _lib
    _lib = None if path is None else _bind(ctypes.CDLL(path))


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def _bind(lib):
    P = ctypes.c_void_p
    for sfx in ("f32", "f64"):
        getattr(lib, f"oracle_time_integration_{sfx}").argtypes = [
            ctypes.POINTER(OracleConfig), P, ctypes.c_long,
            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]
        getattr(lib, f"oracle_time_integration_{sfx}").restype = ctypes.c_long
        getattr(lib, f"oracle_evolve_state_fv_{sfx}").argtypes = [
            ctypes.POINTER(OracleConfig), P, ctypes.c_double]
        getattr(lib, f"oracle_evolve_state_fv_{sfx}").restype = ctypes.c_int
        getattr(lib, f"oracle_cfl_dt_{sfx}").argtypes = [ctypes.POINTER(OracleConfig), P]
        getattr(lib, f"oracle_cfl_dt_{sfx}").restype = ctypes.c_double
        getattr(lib, f"oracle_gas_stage_{sfx}").argtypes = [
            ctypes.POINTER(OracleConfig), P, P, ctypes.c_double]
        getattr(lib, f"oracle_gas_rk2_{sfx}").argtypes = [
            ctypes.POINTER(OracleConfig), P, ctypes.c_double]
        getattr(lib, f"oracle_magnetic_update_{sfx}").argtypes = [
            ctypes.POINTER(OracleConfig), P, P, ctypes.c_double]
        getattr(lib, f"oracle_magnetic_update_{sfx}").restype = ctypes.c_int
        getattr(lib, f"oracle_riemann_{sfx}").argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_long, P, P, P]
        getattr(lib, f"oracle_minmod_{sfx}").argtypes = [ctypes.c_long, P, P, P]
        getattr(lib, f"oracle_limit_{sfx}").argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_long, P, P, P]
        getattr(lib, f"oracle_am_switch_{sfx}").argtypes = [ctypes.POINTER(OracleConfig), P, P]
        getattr(lib, f"oracle_pp_factors_{sfx}").argtypes = [ctypes.POINTER(OracleConfig), P, P]
    return lib


def _sfx(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "f32"
    if dt == np.float64:
        return "f64"
    raise TypeError(f"oracle supports float32/float64, got {dt}")


def make_config(state_shape, dim, *, mhd=True, riemann=0, euler=0, gamma=5.0 / 3.0,
                box_size=1.0, C_cfl=0.4, t_end=0.2, tol=1e-5, fixed_timestep=False,
                num_timesteps=1000, exact_end_time=True, max_iter=1000, limiter=0,
                bc=(0, 0, 0, 0, 0, 0)) -> OracleConfig:
    """bc: (x low, x high, y low, y high, z low, z high), 0 periodic, 1 open, 2 reflective."""
    n = state_shape[-1]
    sh = list(state_shape[1:]) + [1]
    return OracleConfig(
        dim=dim, nx=sh[0], ny=sh[1], nz=sh[2] if dim == 3 else 1, mhd=int(mhd),
        riemann=int(riemann), euler=int(euler), fixed_timestep=int(fixed_timestep),
        num_timesteps=int(num_timesteps), exact_end_time=int(exact_end_time),
        max_iter=int(max_iter), limiter=int(limiter), bc=(ctypes.c_int * 6)(*[int(b) for b in bc]),
        gamma=float(gamma), dx=float(box_size) / n,
        C_cfl=float(C_cfl), t_end=float(t_end), tol=float(tol))


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def time_integration(state: np.ndarray, oc: OracleConfig, max_steps: int = -1):
    """Run the step loop to t_end; returns (final_state, steps, t, magnetic iterations)."""
    lib = _load()
    s = np.ascontiguousarray(state).copy()
    t = ctypes.c_double(0.0)
    it = ctypes.c_long(0)
    steps = getattr(lib, f"oracle_time_integration_{_sfx(s.dtype)}")(
        ctypes.byref(oc), _ptr(s), int(max_steps), ctypes.byref(t), ctypes.byref(it))
    return s, int(steps), float(t.value), int(it.value)


def evolve_state_fv(state: np.ndarray, oc: OracleConfig, dt: float):
    lib = _load()
    s = np.ascontiguousarray(state).copy()
    k = getattr(lib, f"oracle_evolve_state_fv_{_sfx(s.dtype)}")(ctypes.byref(oc), _ptr(s), float(dt))
    return s, int(k)


def cfl_dt(gas: np.ndarray, oc: OracleConfig) -> float:
    lib = _load()
    g = np.ascontiguousarray(gas)
    return float(getattr(lib, f"oracle_cfl_dt_{_sfx(g.dtype)}")(ctypes.byref(oc), _ptr(g)))


def gas_stage(gas: np.ndarray, oc: OracleConfig, dt: float) -> np.ndarray:
    lib = _load()
    g = np.ascontiguousarray(gas)
    out = np.empty_like(g)
    getattr(lib, f"oracle_gas_stage_{_sfx(g.dtype)}")(ctypes.byref(oc), _ptr(g), _ptr(out), float(dt))
    return out


def gas_rk2(gas: np.ndarray, oc: OracleConfig, dt: float) -> np.ndarray:
    lib = _load()
    g = np.ascontiguousarray(gas).copy()
    getattr(lib, f"oracle_gas_rk2_{_sfx(g.dtype)}")(ctypes.byref(oc), _ptr(g), float(dt))
    return g


def magnetic_update(B: np.ndarray, gas: np.ndarray, oc: OracleConfig, dt: float):
    lib = _load()
    b = np.ascontiguousarray(B).copy()
    g = np.ascontiguousarray(gas).copy()
    k = getattr(lib, f"oracle_magnetic_update_{_sfx(b.dtype)}")(ctypes.byref(oc), _ptr(b), _ptr(g), float(dt))
    return b, g, int(k)


def riemann(wl: np.ndarray, wr: np.ndarray, dim: int, axis: int, gamma: float, solver: int = 0,
            alpha: float = 0.0):
    """Interface fluxes for (ng, n) left/right primitive states; axis 1-based;
    alpha is the Lax-Friedrichs dissipation speed (solver 3), or, for the
    hybrid / AM HLLC solvers (4, 5), the compression switch (non-zero = set)."""
    lib = _load()
    a = np.ascontiguousarray(wl)
    b = np.ascontiguousarray(wr, dtype=a.dtype)
    f = np.empty_like(a)
    getattr(lib, f"oracle_riemann_{_sfx(a.dtype)}")(dim, solver, axis, float(gamma), float(alpha), a.shape[1],
                                                    _ptr(a), _ptr(b), _ptr(f))
    return f


def am_switch(gas: np.ndarray, oc: OracleConfig) -> np.ndarray:
    """Compression switch (0/1 per cell) of the AM / hybrid HLLC solvers."""
    lib = _load()
    g = np.ascontiguousarray(gas)
    out = np.empty(g.shape[1:], dtype=np.uint8)
    getattr(lib, f"oracle_am_switch_{_sfx(g.dtype)}")(ctypes.byref(oc), _ptr(g), _ptr(out))
    return out


def pp_factors(gas: np.ndarray, oc: OracleConfig) -> np.ndarray:
    """VAN_ALBADA_PP factors of a gas state, (2 dim + 1,) + grid: alpha per
    axis, kappa per axis, beta (reconstruction.py:168-365)."""
    lib = _load()
    g = np.ascontiguousarray(gas)
    out = np.empty((2 * oc.dim + 1,) + g.shape[1:], dtype=g.dtype)
    getattr(lib, f"oracle_pp_factors_{_sfx(g.dtype)}")(ctypes.byref(oc), _ptr(g), _ptr(out))
    return out


def limit(lim: int, a: np.ndarray, b: np.ndarray, dx: float) -> np.ndarray:
    """Limited gradient from backward / forward gradients a, b (limiter id lim)."""
    lib = _load()
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    out = np.empty_like(a)
    getattr(lib, f"oracle_limit_{_sfx(a.dtype)}")(int(lim), float(dx), a.size, _ptr(a), _ptr(b), _ptr(out))
    return out


def minmod(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    lib = _load()
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    out = np.empty_like(a)
    getattr(lib, f"oracle_minmod_{_sfx(a.dtype)}")(a.size, _ptr(a), _ptr(b), _ptr(out))
    return out
