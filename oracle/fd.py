"""ctypes front of the finite-difference restatement in ``oracle/fd_oracle.cpp``.

TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke, bench.py's CPU
baseline). ``time_integration`` follows the FD branch of
``astronomix/time_stepping/time_integration.py:654-735`` with snapshots off;
``evolve_state`` follows ``_finite_difference/_state_evolution/_evolve_state.py:24-72``.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from oracle import fv as _fv

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfd_oracle.so")
_lib = None


class FdConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("positivity", ctypes.c_int),
        ("fixed_timestep", ctypes.c_int),
        ("num_timesteps", ctypes.c_int),
        ("exact_end_time", ctypes.c_int),
        ("gamma", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("C_cfl", ctypes.c_double),
        ("t_end", ctypes.c_double),
        ("dt_max", ctypes.c_double),
        ("rhomin", ctypes.c_double),
        ("pgmin", ctypes.c_double),
    ]


def use_library(path=None):
    """Bind the functions below to the restatement built at `path` (None: the
    default build). The mutation tests bind copies with planted mistakes."""
    global _lib
    _lib = None if path is None else _bind(ctypes.CDLL(path))


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            _fv.build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def _bind(lib):
    P = ctypes.c_void_p
    C = ctypes.POINTER(FdConfig)
    for s in ("f32", "f64"):
        getattr(lib, f"fd_time_integration_{s}").argtypes = [C, P, ctypes.c_long, ctypes.POINTER(ctypes.c_double)]
        getattr(lib, f"fd_time_integration_{s}").restype = ctypes.c_long
        getattr(lib, f"fd_evolve_state_{s}").argtypes = [C, P, ctypes.c_double]
        getattr(lib, f"fd_cfl_dt_{s}").argtypes = [C, P]
        getattr(lib, f"fd_cfl_dt_{s}").restype = ctypes.c_double
        getattr(lib, f"fd_weno_flux_{s}").argtypes = [C, P, ctypes.c_int, P]
        getattr(lib, f"fd_compute_rhs_{s}").argtypes = [C, P, P, ctypes.c_double, ctypes.c_double, P, P]
        getattr(lib, f"fd_eigensystem_{s}").argtypes = [C, P, P, P, P, P]
        getattr(lib, f"fd_mhd_flux_{s}").argtypes = [C, P, P]
        getattr(lib, f"fd_face_divergence_{s}").argtypes = [C, P, P, P, P]
        getattr(lib, f"fd_interface_fields_{s}").argtypes = [C, P, P]
    return lib


def _sfx(a) -> str:
    return {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64"}[np.dtype(a.dtype)]


def _p(a):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def make_config(shape, *, gamma=5.0 / 3.0, box_size=1.0, C_cfl=0.4, t_end=0.2, dt_max=0.001, rhomin=1e-14,
                pgmin=1e-14, positivity=True, fixed_timestep=False, num_timesteps=1000, exact_end_time=True):
    nx, ny, nz = shape[-3:]
    return FdConfig(nx=nx, ny=ny, nz=nz, positivity=int(positivity), fixed_timestep=int(fixed_timestep),
                    num_timesteps=int(num_timesteps), exact_end_time=int(exact_end_time), gamma=float(gamma),
                    dx=float(box_size) / nz, C_cfl=float(C_cfl), t_end=float(t_end), dt_max=float(dt_max),
                    rhomin=float(rhomin), pgmin=float(pgmin))


def time_integration(state, c: FdConfig, max_steps: int = -1):
    lib = _load()
    s = np.ascontiguousarray(state).copy()
    t = ctypes.c_double(0)
    n = getattr(lib, f"fd_time_integration_{_sfx(s)}")(ctypes.byref(c), _p(s), int(max_steps), ctypes.byref(t))
    return s, int(n), float(t.value)


def evolve_state(state, c: FdConfig, dt: float):
    lib = _load()
    s = np.ascontiguousarray(state).copy()
    getattr(lib, f"fd_evolve_state_{_sfx(s)}")(ctypes.byref(c), _p(s), float(dt))
    return s


def cfl_dt(state, c: FdConfig) -> float:
    lib = _load()
    s = np.ascontiguousarray(state)
    return float(getattr(lib, f"fd_cfl_dt_{_sfx(s)}")(ctypes.byref(c), _p(s)))


def weno_flux(q, c: FdConfig, direction: int):
    lib = _load()
    q = np.ascontiguousarray(q)
    F = np.empty_like(q)
    getattr(lib, f"fd_weno_flux_{_sfx(q)}")(ctypes.byref(c), _p(q), int(direction), _p(F))
    return F


def compute_rhs(q, b, c: FdConfig, k2: float, dt: float):
    lib = _load()
    q = np.ascontiguousarray(q)
    b = np.ascontiguousarray(b, dtype=q.dtype)
    rq = np.empty_like(q)
    rb = np.empty_like(b)
    getattr(lib, f"fd_compute_rhs_{_sfx(q)}")(ctypes.byref(c), _p(q), _p(b), float(k2), float(dt), _p(rq), _p(rb))
    return rq, rb


def eigensystem(qa, qb, c: FdConfig):
    """(L 7x8, R 8x7, cell lambdas of qa) of the x-frame interface between qa and qb."""
    lib = _load()
    qa = np.ascontiguousarray(qa)
    qb = np.ascontiguousarray(qb, dtype=qa.dtype)
    L = np.empty((7, 8), dtype=qa.dtype)
    R = np.empty((8, 7), dtype=qa.dtype)
    lam = np.empty(7, dtype=qa.dtype)
    getattr(lib, f"fd_eigensystem_{_sfx(qa)}")(ctypes.byref(c), _p(qa), _p(qb), _p(L), _p(R), _p(lam))
    return L, R, lam


def mhd_flux(q, c: FdConfig):
    """x-direction ideal-MHD flux of conserved cells q (8,) or (8, ...)."""
    lib = _load()
    q = np.asarray(q)
    flat = np.ascontiguousarray(q.reshape(8, -1).T)
    f = np.empty_like(flat)
    fn = getattr(lib, f"fd_mhd_flux_{_sfx(flat)}")
    for i in range(flat.shape[0]):
        fn(ctypes.byref(c), _p(flat[i]), _p(f[i]))
    return np.ascontiguousarray(f.T).reshape(q.shape)


def face_divergence(bx, by, bz, c: FdConfig):
    lib = _load()
    bx, by, bz = (np.ascontiguousarray(a) for a in (bx, by, bz))
    out = np.empty_like(bx)
    getattr(lib, f"fd_face_divergence_{_sfx(bx)}")(ctypes.byref(c), _p(bx), _p(by), _p(bz), _p(out))
    return out


def interface_fields(B, c: FdConfig):
    """initialize_interface_fields (_constrained_transport.py:201-213)."""
    lib = _load()
    B = np.ascontiguousarray(B)
    b = np.empty_like(B)
    getattr(lib, f"fd_interface_fields_{_sfx(B)}")(ctypes.byref(c), _p(B), _p(b))
    return b
