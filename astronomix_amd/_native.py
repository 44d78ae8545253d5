"""ctypes binding of ``libamx.so`` (the C ABI declared in ``include/amx.h``).

The library is built in-tree by ``__graft_entry__.build()`` /
``astronomix_amd.build``. There is no CPU fallback: if the library is absent,
or no GPU is visible, every entry point raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("AMX_LIB") or os.path.join(_HERE, "libamx.so")

AMX_OK = 0
AMX_HLL = 0
AMX_HLLC = 1
AMX_HLLC_LM = 2
AMX_LAX_FRIEDRICHS = 3
AMX_MINMOD = 0
AMX_BC_PERIODIC, AMX_BC_OPEN, AMX_BC_REFLECTIVE = 0, 1, 2


class AmxFvConfig(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("mhd", ctypes.c_int),
        ("riemann", ctypes.c_int),
        ("magnetic_integrator", ctypes.c_int),
        ("max_fixed_point_iterations", ctypes.c_int),
        ("gamma", ctypes.c_float),
        ("C_cfl", ctypes.c_float),
        ("fixed_point_tol", ctypes.c_float),
        ("limiter", ctypes.c_int),
        ("grid_spacing", ctypes.c_double),
        ("boundary", ctypes.c_int * 6),
    ]


class AmxFvRun(ctypes.Structure):
    _fields_ = [
        ("t_end", ctypes.c_float),
        ("fixed_timestep", ctypes.c_int),
        ("num_timesteps", ctypes.c_int),
        ("exact_end_time", ctypes.c_int),
        ("max_steps", ctypes.c_long),
    ]


class AmxFvStats(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_long),
        ("t", ctypes.c_double),
        ("fixed_point_iterations", ctypes.c_long),
        ("max_fixed_point_iterations", ctypes.c_int),
    ]


class AmxFvTiming(ctypes.Structure):
    _fields_ = [
        ("stage_ms", ctypes.c_double),
        ("stage_launches", ctypes.c_long),
        ("mag_ms", ctypes.c_double),
        ("mag_launches", ctypes.c_long),
    ]


class AmxFvSlotStatus(ctypes.Structure):
    _fields_ = [
        ("live", ctypes.c_int),
        ("finished", ctypes.c_int),
        ("steps", ctypes.c_long),
        ("t", ctypes.c_double),
        ("t_end", ctypes.c_double),
        ("fixed_point_iterations", ctypes.c_long),
    ]


class AmxFdConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("enforce_positivity", ctypes.c_int),
        ("gamma", ctypes.c_float),
        ("C_cfl", ctypes.c_float),
        ("dt_max", ctypes.c_float),
        ("minimum_density", ctypes.c_float),
        ("minimum_pressure", ctypes.c_float),
        ("pad_", ctypes.c_float),
        ("grid_spacing", ctypes.c_double),
    ]


AMX_DIV_NONE, AMX_DIV_CENTRAL, AMX_DIV_FACES = 0, 1, 2


class AmxDiagDesc(ctypes.Structure):
    _fields_ = [
        ("num_vars", ctypes.c_int),
        ("dim", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("rho_row", ctypes.c_int),
        ("vel_row", ctypes.c_int),
        ("p_row", ctypes.c_int),
        ("mag_row", ctypes.c_int),
        ("div_mode", ctypes.c_int),
        ("gamma", ctypes.c_float),
        ("grid_spacing", ctypes.c_double),
        ("box_size", ctypes.c_double),
    ]


# SnapshotData fields recorded per snapshot, in amx_snapshot_out order
SNAPSHOT_FIELDS = ("time_points", "states", "total_mass", "total_energy", "internal_energy", "kinetic_energy",
                   "radial_momentum", "magnetic_divergence")


class AmxSnapshotOut(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in SNAPSHOT_FIELDS] + [
        ("current_checkpoint", ctypes.c_int),
        ("num_iterations", ctypes.c_long),
    ]


SNAPSHOT_CALLBACK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_float, ctypes.c_int, ctypes.c_void_p)
STEP_HOOK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                             ctypes.c_void_p)


class AmxSnapshotSpec(ctypes.Structure):
    _fields_ = [
        ("num_snapshots", ctypes.c_int),
        ("use_timepoints", ctypes.c_int),
        ("timepoints", ctypes.POINTER(ctypes.c_float)),
        ("num_timepoints", ctypes.c_int),
        ("box_size", ctypes.c_double),
        ("callback", SNAPSHOT_CALLBACK),
        ("user", ctypes.c_void_p),
    ]


# Every symbol include/amx.h declares, with its ctypes signature.
_P = ctypes.c_void_p
SIGNATURES = {
    "amx_fv_create": (ctypes.c_int, [ctypes.POINTER(AmxFvConfig), ctypes.POINTER(_P)]),
    "amx_fv_destroy": (ctypes.c_int, [_P]),
    "amx_fv_workspace_bytes": (ctypes.c_int, [ctypes.POINTER(AmxFvConfig), ctypes.POINTER(ctypes.c_size_t)]),
    "amx_fv_cfl_dt": (ctypes.c_int, [_P, _P, _P, _P]),
    "amx_fv_evolve_state": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(ctypes.c_int), _P]),
    "amx_fv_evolve_state_noalloc": (ctypes.c_int, [ctypes.POINTER(AmxFvConfig), _P, _P,
                                                   ctypes.POINTER(ctypes.c_int), _P]),
    "amx_fv_time_integration": (ctypes.c_int, [_P, _P, ctypes.POINTER(AmxFvRun), ctypes.POINTER(AmxFvStats), _P]),
    "amx_fv_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "amx_fv_get_timing": (ctypes.c_int, [_P, ctypes.POINTER(AmxFvTiming)]),
    "amx_fv_set_physics": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_float]),
    "amx_fv_adjoint_step": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "amx_fv_pool_slot_bytes": (ctypes.c_int, [ctypes.POINTER(AmxFvConfig), ctypes.POINTER(ctypes.c_size_t)]),
    "amx_fv_pool_create": (ctypes.c_int, [ctypes.POINTER(AmxFvConfig), ctypes.POINTER(AmxFvRun), ctypes.c_int,
                                          ctypes.POINTER(_P)]),
    "amx_fv_pool_destroy": (ctypes.c_int, [_P]),
    "amx_fv_pool_load": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_float, ctypes.c_float, _P]),
    "amx_fv_pool_read": (ctypes.c_int, [_P, ctypes.c_int, _P, _P]),
    "amx_fv_pool_release": (ctypes.c_int, [_P, ctypes.c_int, _P]),
    "amx_fv_pool_step_launch": (ctypes.c_int, [_P, _P]),
    "amx_fv_pool_step_finish": (ctypes.c_int, [_P, _P]),
    "amx_fv_pool_set_graphs": (ctypes.c_int, [_P, ctypes.c_int]),
    "amx_fv_pool_set_batch": (ctypes.c_int, [_P, ctypes.c_int]),
    "amx_fv_pool_status": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(AmxFvSlotStatus)]),
    "amx_fd_create": (ctypes.c_int, [ctypes.POINTER(AmxFdConfig), ctypes.POINTER(_P)]),
    "amx_fd_destroy": (ctypes.c_int, [_P]),
    "amx_fd_workspace_bytes": (ctypes.c_int, [ctypes.POINTER(AmxFdConfig), ctypes.POINTER(ctypes.c_size_t)]),
    "amx_fd_cfl_dt": (ctypes.c_int, [_P, _P, _P, _P]),
    "amx_fd_evolve_state": (ctypes.c_int, [_P, _P, _P, _P]),
    "amx_fd_time_integration": (ctypes.c_int, [_P, _P, ctypes.POINTER(AmxFvRun), ctypes.POINTER(AmxFvStats), _P]),
    "amx_fd_interface_fields": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]),
    "amx_fd_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "amx_fd_get_timing": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]),
    "amx_snapshot_workspace_bytes": (ctypes.c_size_t, []),
    "amx_snapshot_record": (ctypes.c_int, [ctypes.POINTER(AmxDiagDesc), _P, ctypes.c_float,
                                           ctypes.POINTER(AmxSnapshotOut), ctypes.c_int, _P, _P]),
    "amx_fv_time_integration_snapshots": (ctypes.c_int, [_P, _P, ctypes.POINTER(AmxFvRun),
                                                         ctypes.POINTER(AmxSnapshotSpec),
                                                         ctypes.POINTER(AmxSnapshotOut),
                                                         ctypes.POINTER(AmxFvStats), _P]),
    "amx_fd_time_integration_snapshots": (ctypes.c_int, [_P, _P, ctypes.POINTER(AmxFvRun),
                                                         ctypes.POINTER(AmxSnapshotSpec),
                                                         ctypes.POINTER(AmxSnapshotOut),
                                                         ctypes.POINTER(AmxFvStats), _P]),
    "amx_fv_set_step_hook": (ctypes.c_int, [_P, _P, _P]),
    "amx_fd_set_step_hook": (ctypes.c_int, [_P, _P, _P]),
    "amx_forcing_spectrum": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_ulonglong, ctypes.c_ulonglong, _P]),
    "amx_forcing_workspace_bytes": (ctypes.c_size_t, []),
    "amx_forcing_apply": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, _P, _P, ctypes.c_float, ctypes.c_double, _P, _P, _P]),
    "amx_copy_f32": (ctypes.c_int, [_P, _P, ctypes.c_size_t, _P]),
    "amx_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "amx_last_error": (ctypes.c_char_p, []),
}

_lib = None
_lock = threading.Lock()


class AmxError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH):
    """Load libamx.so (import torch first so both share one HIP runtime)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise AmxError(
                    f"{path} is not built: the HIP extension is required (run __graft_entry__.build())")
            import torch  # noqa: F401  -- load torch's HIP runtime first
            lib = ctypes.CDLL(path)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(status: int, what: str):
    if status != AMX_OK:
        lib = load_library()
        msg = lib.amx_last_error().decode(errors="replace")
        raise AmxError(f"{what} failed: {lib.amx_status_string(status).decode()} ({msg})")


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise AmxError("no GPU visible: the astronomix_amd step runs only on the HIP device")


class FvContext:
    """Owns one amx_fv_ctx (device workspace for one grid/configuration)."""

    def __init__(self, cfg: AmxFvConfig):
        require_gpu()
        self.lib = load_library()
        self.cfg = cfg
        h = ctypes.c_void_p()
        check(self.lib.amx_fv_create(ctypes.byref(cfg), ctypes.byref(h)), "amx_fv_create")
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self.lib.amx_fv_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _stream(stream):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def cfl_dt(self, state, out, stream=None):
        check(self.lib.amx_fv_cfl_dt(self.handle, ctypes.c_void_p(state.data_ptr()),
                                     ctypes.c_void_p(out.data_ptr()), self._stream(stream)), "amx_fv_cfl_dt")

    def evolve_state(self, state, dt, stream=None) -> int:
        k = ctypes.c_int(0)
        check(self.lib.amx_fv_evolve_state(self.handle, ctypes.c_void_p(state.data_ptr()),
                                           ctypes.c_void_p(dt.data_ptr()), ctypes.byref(k),
                                           self._stream(stream)), "amx_fv_evolve_state")
        return int(k.value)

    def set_timing(self, every: int):
        """0: off; n >= 1: HIP-event timing of the kernels of every n-th step."""
        check(self.lib.amx_fv_set_timing(self.handle, max(0, int(every))), "amx_fv_set_timing")

    def get_timing(self) -> AmxFvTiming:
        t = AmxFvTiming()
        check(self.lib.amx_fv_get_timing(self.handle, ctypes.byref(t)), "amx_fv_get_timing")
        return t

    def adjoint_step(self, state, dt, lam_out, lam_in, stream=None):
        check(self.lib.amx_fv_adjoint_step(self.handle, ctypes.c_void_p(state.data_ptr()), ctypes.c_void_p(dt.data_ptr()),
                                           ctypes.c_void_p(lam_out.data_ptr()), ctypes.c_void_p(lam_in.data_ptr()),
                                           self._stream(stream)), "amx_fv_adjoint_step")

    def set_physics(self, gamma: float, C_cfl: float):
        check(self.lib.amx_fv_set_physics(self.handle, float(gamma), float(C_cfl)), "amx_fv_set_physics")
        self.cfg.gamma = float(gamma)
        self.cfg.C_cfl = float(C_cfl)

    def time_integration(self, state, run: AmxFvRun, stream=None) -> AmxFvStats:
        st = AmxFvStats()
        check(self.lib.amx_fv_time_integration(self.handle, ctypes.c_void_p(state.data_ptr()), ctypes.byref(run),
                                               ctypes.byref(st), self._stream(stream)), "amx_fv_time_integration")
        return st

    def time_integration_snapshots(self, state, run: AmxFvRun, snapshots: "SnapshotRun", stream=None) -> AmxFvStats:
        st = AmxFvStats()
        snapshots.invoke(self.lib.amx_fv_time_integration_snapshots, self.handle, state, run, st,
                         self._stream(stream), "amx_fv_time_integration_snapshots")
        return st


def pool_slot_bytes(cfg: AmxFvConfig) -> int:
    lib = load_library()
    n = ctypes.c_size_t(0)
    check(lib.amx_fv_pool_slot_bytes(ctypes.byref(cfg), ctypes.byref(n)), "amx_fv_pool_slot_bytes")
    return int(n.value)


class FvPool:
    """Owns one amx_fv_pool: `capacity` slots of independent simulations of
    one grid and scheme, stepped together (continuous batching). All calls
    go to the pool's own HIP stream."""

    def __init__(self, cfg: AmxFvConfig, capacity: int, run: AmxFvRun = None):
        import torch
        require_gpu()
        self.lib = load_library()
        self.cfg = cfg
        self.capacity = int(capacity)
        if run is None:
            run = AmxFvRun(t_end=0.0, fixed_timestep=0, num_timesteps=0, exact_end_time=1, max_steps=-1)
        self.stream = torch.cuda.Stream()
        h = ctypes.c_void_p()
        check(self.lib.amx_fv_pool_create(ctypes.byref(cfg), ctypes.byref(run), self.capacity, ctypes.byref(h)),
              "amx_fv_pool_create")
        self.handle = h
        self._s = ctypes.c_void_p(self.stream.cuda_stream)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.amx_fv_pool_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, slot: int, state, gamma: float, t_end: float):
        """state: float32 tensor (GPU or CPU) or numpy array in the reference layout."""
        import numpy as np
        import torch
        if isinstance(state, np.ndarray):
            state = torch.from_numpy(np.ascontiguousarray(state, dtype=np.float32))
        state = state.contiguous()
        if state.dtype != torch.float32:
            raise TypeError("state must be float32")
        check(self.lib.amx_fv_pool_load(self.handle, int(slot), ctypes.c_void_p(state.data_ptr()),
                                        float(gamma), float(t_end), self._s), "amx_fv_pool_load")
        self.stream.synchronize()  # the source may be a temporary

    def read(self, slot: int, out):
        check(self.lib.amx_fv_pool_read(self.handle, int(slot), ctypes.c_void_p(out.data_ptr()), self._s),
              "amx_fv_pool_read")
        self.stream.synchronize()
        return out

    def release(self, slot: int):
        check(self.lib.amx_fv_pool_release(self.handle, int(slot), self._s), "amx_fv_pool_release")

    def set_graphs(self, enable: bool):
        check(self.lib.amx_fv_pool_set_graphs(self.handle, int(bool(enable))), "amx_fv_pool_set_graphs")

    def set_batch(self, steps: int):
        check(self.lib.amx_fv_pool_set_batch(self.handle, int(steps)), "amx_fv_pool_set_batch")

    def step_launch(self):
        check(self.lib.amx_fv_pool_step_launch(self.handle, self._s), "amx_fv_pool_step_launch")

    def step_finish(self):
        check(self.lib.amx_fv_pool_step_finish(self.handle, self._s), "amx_fv_pool_step_finish")

    def status(self, slot: int) -> AmxFvSlotStatus:
        st = AmxFvSlotStatus()
        check(self.lib.amx_fv_pool_status(self.handle, int(slot), ctypes.byref(st)), "amx_fv_pool_status")
        return st


class FdContext:
    """Owns one amx_fd_ctx (device workspace of the SSPRK(5,4) FD step)."""

    def __init__(self, cfg: AmxFdConfig):
        require_gpu()
        self.lib = load_library()
        self.cfg = cfg
        h = ctypes.c_void_p()
        check(self.lib.amx_fd_create(ctypes.byref(cfg), ctypes.byref(h)), "amx_fd_create")
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self.lib.amx_fd_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    _stream = staticmethod(FvContext._stream)

    def cfl_dt(self, state, out, stream=None):
        check(self.lib.amx_fd_cfl_dt(self.handle, ctypes.c_void_p(state.data_ptr()),
                                     ctypes.c_void_p(out.data_ptr()), self._stream(stream)), "amx_fd_cfl_dt")

    def evolve_state(self, state, dt, stream=None):
        check(self.lib.amx_fd_evolve_state(self.handle, ctypes.c_void_p(state.data_ptr()),
                                           ctypes.c_void_p(dt.data_ptr()), self._stream(stream)),
              "amx_fd_evolve_state")

    def set_timing(self, enable: bool):
        check(self.lib.amx_fd_set_timing(self.handle, int(bool(enable))), "amx_fd_set_timing")

    def get_timing(self):
        ms, n = ctypes.c_double(0), ctypes.c_long(0)
        check(self.lib.amx_fd_get_timing(self.handle, ctypes.byref(ms), ctypes.byref(n)), "amx_fd_get_timing")
        return float(ms.value), int(n.value)

    def time_integration(self, state, run: AmxFvRun, stream=None) -> AmxFvStats:
        st = AmxFvStats()
        check(self.lib.amx_fd_time_integration(self.handle, ctypes.c_void_p(state.data_ptr()), ctypes.byref(run),
                                               ctypes.byref(st), self._stream(stream)), "amx_fd_time_integration")
        return st

    def time_integration_snapshots(self, state, run: AmxFvRun, snapshots: "SnapshotRun", stream=None) -> AmxFvStats:
        st = AmxFvStats()
        snapshots.invoke(self.lib.amx_fd_time_integration_snapshots, self.handle, state, run, st,
                         self._stream(stream), "amx_fd_time_integration_snapshots")
        return st


class SnapshotRun:
    """Host side of one snapshot run: the amx_snapshot_spec, the device
    arrays of amx_snapshot_out (torch tensors, kept alive here) and, in
    callback mode, the ctypes trampoline. An exception raised by the Python
    callback stops the run and is re-raised after the C call returns."""

    def __init__(self, num_snapshots: int, box_size: float, arrays: dict = None, timepoints=None,
                 callback=None):
        self.arrays = dict(arrays or {})
        self.out = AmxSnapshotOut()
        for name in SNAPSHOT_FIELDS:
            t = self.arrays.get(name)
            setattr(self.out, name, ctypes.c_void_p(t.data_ptr()) if t is not None else None)
        self._tp = None
        n_tp = 0
        if timepoints is not None:
            vals = [float(v) for v in timepoints]
            n_tp = len(vals)
            self._tp = (ctypes.c_float * max(n_tp, 1))(*vals)
        self._error = None
        self._cb = None
        if callback is not None:
            def trampoline(user, t, checkpoint, state_ptr):
                try:
                    callback(float(t), int(checkpoint), int(state_ptr or 0))
                    return 0
                except BaseException as e:  # noqa: BLE001 -- re-raised after the C call
                    self._error = e
                    return 1
            self._cb = SNAPSHOT_CALLBACK(trampoline)
        self.spec = AmxSnapshotSpec(
            num_snapshots=int(num_snapshots), use_timepoints=int(timepoints is not None),
            timepoints=ctypes.cast(self._tp, ctypes.POINTER(ctypes.c_float)) if self._tp is not None else None,
            num_timepoints=n_tp, box_size=float(box_size),
            callback=self._cb if self._cb is not None else SNAPSHOT_CALLBACK(), user=None)

    def invoke(self, fn, handle, state, run, stats, stream, what):
        status = fn(handle, ctypes.c_void_p(state.data_ptr()), ctypes.byref(run), ctypes.byref(self.spec),
                    ctypes.byref(self.out), ctypes.byref(stats), stream)
        if self._error is not None:
            err, self._error = self._error, None
            raise err
        check(status, what)

    @property
    def current_checkpoint(self) -> int:
        return int(self.out.current_checkpoint)

    @property
    def num_iterations(self) -> int:
        return int(self.out.num_iterations)


def snapshot_workspace_bytes() -> int:
    return int(load_library().amx_snapshot_workspace_bytes())


def snapshot_record(desc: AmxDiagDesc, state, t: float, arrays: dict, index: int = 0, work=None, stream=None):
    """One snapshot row of `state` into the device arrays named in `arrays`
    (SNAPSHOT_FIELDS keys, float32 tensors) at `index`."""
    import torch
    lib = load_library()
    if work is None:
        work = torch.empty(snapshot_workspace_bytes(), dtype=torch.uint8, device=state.device)
    out = AmxSnapshotOut()
    for name in SNAPSHOT_FIELDS:
        a = arrays.get(name)
        setattr(out, name, ctypes.c_void_p(a.data_ptr()) if a is not None else None)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.amx_snapshot_record(ctypes.byref(desc), ctypes.c_void_p(state.data_ptr()), float(t), ctypes.byref(out),
                                  int(index), ctypes.c_void_p(work.data_ptr()), ctypes.c_void_p(s.cuda_stream)),
          "amx_snapshot_record")


def copy_f32(dst, src, stream=None):
    """dst[:] = src[:] on the device with the library's streaming copy kernel."""
    import torch
    lib = load_library()
    assert dst.numel() == src.numel() and dst.dtype == src.dtype == torch.float32
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib.amx_copy_f32(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()), dst.numel(),
                           ctypes.c_void_p(s.cuda_stream)), "amx_copy_f32")


_ctx_cache = {}


def _field_key(cfg, f):
    v = getattr(cfg, f)
    return tuple(v) if isinstance(v, ctypes.Array) else v


def _cached(kind, cls, cfg):
    import torch
    key = (kind, torch.cuda.current_device()) + tuple(_field_key(cfg, f) for f, _ in type(cfg)._fields_)
    ctx = _ctx_cache.get(key)
    if ctx is None:
        ctx = cls(cfg)
        _ctx_cache[key] = ctx
    return ctx


def fv_context(cfg: AmxFvConfig) -> FvContext:
    """Cached context per (device, grid and scheme); gamma and C_cfl are set
    on the cached context when they differ (they are run-time parameters)."""
    import torch
    key = ("fv", torch.cuda.current_device()) + tuple(_field_key(cfg, f) for f, _ in AmxFvConfig._fields_
                                                      if f not in ("gamma", "C_cfl"))
    ctx = _ctx_cache.get(key)
    if ctx is None:
        ctx = FvContext(AmxFvConfig.from_buffer_copy(cfg))
        _ctx_cache[key] = ctx
    elif ctx.cfg.gamma != cfg.gamma or ctx.cfg.C_cfl != cfg.C_cfl:
        ctx.set_physics(cfg.gamma, cfg.C_cfl)
    return ctx


def fd_context(cfg: AmxFdConfig) -> FdContext:
    """Cached FD context per (device, configuration)."""
    return _cached("fd", FdContext, cfg)


def clear_context_cache():
    """Frees the cached contexts (their device workspaces)."""
    for ctx in list(_ctx_cache.values()):
        ctx.close()
    _ctx_cache.clear()
