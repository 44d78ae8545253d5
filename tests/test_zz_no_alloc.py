"""The allocation-free step (amx_fv_evolve_state_noalloc, through the C ABI)
against the context step (amx_fv_evolve_state) on the same inputs: equal bit
for bit, for MHD and hydro, 2D and 3D, HLL / HLLC / the LF and VAN_ALBADA_PP
pre-pass kernels, a fixed point past the first launches, and two calls in a
row on one device's scratch, the second on a smaller grid; a grid past the
scratch is refused before any device work."""

import ctypes

import numpy as np
import pytest

from astronomix_amd import _native
from astronomix_amd import workloads as W
from astronomix_amd._finite_volume import native_config
from astronomix_amd.option_classes.simulation_config import HLL, HLLC, LAX_FRIEDRICHS, VAN_ALBADA_PP

pytestmark = pytest.mark.gpu


def noalloc_step(nc, state, dt):
    lib = _native.load_library()
    k = ctypes.c_int(-1)
    _native.check(lib.amx_fv_evolve_state_noalloc(ctypes.byref(nc), ctypes.c_void_p(state.data_ptr()),
                                                  ctypes.c_void_p(dt.data_ptr()), ctypes.byref(k), None),
                  "amx_fv_evolve_state_noalloc")
    return int(k.value)


def case(name):
    if name == "blast3d":
        s = W.blast_3d(32)
        cfg = W.fv_config(3, 32, mhd=True, riemann=HLL)
        return s, cfg, 5 / 3, 0.3, 1e-5
    if name == "ot2d":
        s = W.orszag_tang_2d(48)
        cfg = W.fv_config(2, 48, mhd=True, riemann=HLL)
        return s, cfg, 5 / 3, 0.4, 1e-5
    if name == "mhd3d_tight":  # tolerance 1e-6: the fixed point runs past the first launches
        s = W.random_smooth((24, 24, 24), True, seed=1, amp=1.0)
        cfg = W.fv_config(3, 24, mhd=True, riemann=HLL)
        return s, cfg, 5 / 3, 0.8, 1e-6
    if name == "kh2d_hllc":
        s = W.kelvin_helmholtz_2d(64, seed=3)
        cfg = W.fv_config(2, 64, mhd=False, riemann=HLLC)
        return s, cfg, 5 / 3, 0.4, 1e-5
    if name == "hydro3d_lf":
        s = W.random_smooth((20, 18, 24), False, seed=4)
        cfg = W.fv_config(3, 20, mhd=False, riemann=LAX_FRIEDRICHS)
        return s, cfg, 1.4, 0.4, 1e-5
    if name == "ot2d_pp":
        s = W.orszag_tang_2d(40)
        cfg = W.fv_config(2, 40, mhd=True, riemann=HLL)._replace(limiter=VAN_ALBADA_PP)
        return s, cfg, 5 / 3, 0.4, 1e-5
    raise KeyError(name)


def check_equal(name, steps=3):
    import torch
    s, cfg, gamma, C, tol = case(name)
    nc = native_config(s.shape, cfg, gamma, C)
    nc.fixed_point_tol = tol
    ctx = _native.FvContext(_native.AmxFvConfig.from_buffer_copy(nc))
    a = torch.from_numpy(np.ascontiguousarray(s)).cuda()
    b = a.clone()
    dt = torch.empty((), dtype=torch.float32, device="cuda")
    for step in range(steps):
        ctx.cfl_dt(a, dt)
        ka = ctx.evolve_state(a, dt)
        kb = noalloc_step(nc, b, dt)
        torch.cuda.synchronize()
        assert ka == kb, (name, step, ka, kb)
        assert torch.equal(a, b), (name, step, float((a - b).abs().max()))
    return ka


@pytest.mark.parametrize("name", ["blast3d", "ot2d", "kh2d_hllc", "hydro3d_lf", "ot2d_pp"])
def test_noalloc_step_equals_context_step(gpu, name):
    print(f"[gpu] noalloc vs context step: {name}", flush=True)
    check_equal(name)


def test_noalloc_fixed_point_past_first_launches(gpu):
    print("[gpu] noalloc vs context step: tight fixed point", flush=True)
    k = check_equal("mhd3d_tight", steps=2)
    assert k > 5


def test_noalloc_two_calls_second_smaller(gpu):
    """Two calls in a row on the device's scratch: a 32^3 MHD step, then a
    smaller 2D one; each equals its context step (the second call reads
    nothing the first left behind)."""
    print("[gpu] noalloc: 32^3 then 48^2", flush=True)
    check_equal("blast3d", steps=1)
    check_equal("ot2d", steps=1)


def test_noalloc_refuses_grid_past_the_scratch(gpu):
    import torch
    print("[gpu] noalloc: grid past the scratch is refused", flush=True)
    cfg = W.fv_config(3, 128, mhd=True, riemann=HLL)
    nc = native_config((8, 128, 128, 128), cfg, 5 / 3, 0.3)
    lib = _native.load_library()
    state = torch.zeros(1, device="cuda")
    dt = torch.zeros((), device="cuda")
    st = lib.amx_fv_evolve_state_noalloc(ctypes.byref(nc), ctypes.c_void_p(state.data_ptr()),
                                         ctypes.c_void_p(dt.data_ptr()), None, None)
    assert st == 1  # AMX_ERR_INVALID
    assert b"scratch" in lib.amx_last_error()
