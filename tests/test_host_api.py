"""Host-side logic on the CPU: configuration, registry layout, state
construction, the implemented-path gate, the C ABI's exported symbols and
the bench's one-GPU contract. No GPU calls."""

import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import astronomix_amd as amx
from astronomix_amd import _native
from astronomix_amd._finite_volume import boundary_codes, unsupported_reason
from astronomix_amd.option_classes import simulation_config as sc
from astronomix_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("dim,mhd,mode,layout", [
    (2, False, sc.FINITE_VOLUME, dict(num_vars=4, pressure_index=3, velocity_index=(1, 2, -1))),
    (2, True, sc.FINITE_VOLUME, dict(num_vars=7, pressure_index=3, magnetic_index=(4, 5, 6))),
    (3, False, sc.FINITE_VOLUME, dict(num_vars=5, pressure_index=4, velocity_index=(1, 2, 3))),
    (3, True, sc.FINITE_VOLUME, dict(num_vars=8, pressure_index=4, magnetic_index=(5, 6, 7))),
    (3, True, sc.FINITE_DIFFERENCE, dict(num_vars=11, interface_magnetic_field_index=(8, 9, 10))),
    (1, False, sc.FINITE_VOLUME, dict(num_vars=3, pressure_index=2, velocity_index=1)),
])
def test_registry_layouts(dim, mhd, mode, layout):
    rv = amx.get_registered_variables(amx.SimulationConfig(dimensionality=dim, mhd=mhd, solver_mode=mode))
    for k, v in layout.items():
        got = getattr(rv, k)
        assert (tuple(got) if isinstance(v, tuple) else got) == v, k
    assert rv.density_index == 0
    assert rv.momentum_index == rv.velocity_index and rv.energy_index == rv.pressure_index


def test_registry_tracers_and_axes():
    cfg = amx.SimulationConfig(dimensionality=3, mhd=True,
                               wind_config=sc.WindConfig(trace_wind_density=True),
                               cosmic_ray_config=sc.CosmicRayConfig(cosmic_rays=True))
    rv = amx.get_registered_variables(cfg)
    assert (rv.wind_density_index, rv.cosmic_ray_n_index, rv.num_vars) == (8, 9, 10)
    assert [a.axis_in_array for a in rv.axes_info] == [1, 2, 3]
    assert [a.magnetic_index for a in rv.axes_info] == [5, 6, 7]


def test_finalize_config():
    cfg = amx.SimulationConfig(dimensionality=3, mhd=True, box_size=2.0)
    out = amx.finalize_config(cfg, (8, 32, 32, 32))
    assert out.num_cells == 32 and out.grid_spacing == 2.0 / 32
    assert out.boundary_settings == sc.BoundarySettings()  # open boundaries by default
    with pytest.raises(ValueError):
        amx.finalize_config(cfg, (8, 32, 16, 32))
    fd = amx.SimulationConfig(dimensionality=3, mhd=True, solver_mode=sc.FINITE_DIFFERENCE,
                              boundary_settings=sc.ALL_PERIODIC)
    out = amx.finalize_config(fd, (11, 16, 16, 16))
    assert out.time_integrator == sc.RK4_SSP and out.boundary_handling == sc.PERIODIC_ROLL
    with pytest.raises(ValueError):
        amx.finalize_config(fd._replace(boundary_settings=None), (11, 16, 16, 16))
    with pytest.raises(ValueError):
        amx.finalize_config(fd._replace(dimensionality=2), (11, 16, 16))


def test_construct_primitive_state_layout():
    cfg = amx.SimulationConfig(dimensionality=3, mhd=True)
    rv = amx.get_registered_variables(cfg)
    f = [np.full((4, 4, 4), float(i)) for i in range(8)]
    s = amx.construct_primitive_state(cfg, rv, density=f[0], velocity_x=f[1], velocity_y=f[2], velocity_z=f[3],
                                      magnetic_field_x=f[5], magnetic_field_y=f[6], magnetic_field_z=f[7],
                                      gas_pressure=f[4])
    assert s.shape == (8, 4, 4, 4) and s.dtype == np.float32
    for i in range(8):
        assert np.all(s[i] == i)


def test_implemented_path_gate():
    ok = W.fv_config(3, 16)
    assert unsupported_reason(ok) is None
    assert unsupported_reason(W.fv_config(2, 16, mhd=False, riemann=sc.HLLC)) is None
    # open / reflective ghost boundaries (BoundarySettings() is open on every side)
    assert unsupported_reason(ok._replace(boundary_settings=sc.BoundarySettings())) is None
    refl = sc.BoundarySettings1D(sc.REFLECTIVE_BOUNDARY, sc.REFLECTIVE_BOUNDARY)
    per = sc.BoundarySettings1D(sc.PERIODIC_BOUNDARY, sc.PERIODIC_BOUNDARY)
    mixed = sc.BoundarySettings(x=refl, y=per, z=sc.BoundarySettings1D())
    assert unsupported_reason(ok._replace(boundary_settings=mixed)) is None
    assert boundary_codes(ok._replace(boundary_settings=mixed)) == (2, 2, 0, 0, 1, 1)
    assert boundary_codes(W.fv_config(2, 16)._replace(boundary_settings=mixed)) == (2, 2, 0, 0, 0, 0)
    # ghost cells only with GHOST_CELLS handling; PERIODIC_ROLL rolls every stencil
    assert boundary_codes(ok._replace(boundary_settings=mixed, boundary_handling=sc.PERIODIC_ROLL)) == (0,) * 6
    one_sided = sc.BoundarySettings(x=sc.BoundarySettings1D(sc.PERIODIC_BOUNDARY, sc.OPEN_BOUNDARY), y=per, z=per)
    assert "one side" in unsupported_reason(ok._replace(boundary_settings=one_sided))
    jet = sc.BoundarySettings(x=per, y=sc.BoundarySettings1D(sc.MHD_JET_BOUNDARY, sc.OPEN_BOUNDARY), z=per)
    assert "jet" in unsupported_reason(ok._replace(boundary_settings=jet))
    assert "periodic" in unsupported_reason(ok._replace(boundary_settings=mixed, riemann_solver=sc.AM_HLLC))
    assert "UNSPLIT" in unsupported_reason(ok._replace(split=sc.SPLIT))
    for lim in (sc.OSHER, sc.DOUBLE_MINMOD, sc.SUPERBEE, sc.VAN_ALBADA):
        assert unsupported_reason(ok._replace(limiter=lim)) is None
    for rs in (sc.HLLC_LM, sc.LAX_FRIEDRICHS, sc.HYBRID_HLLC, sc.AM_HLLC):
        assert unsupported_reason(ok._replace(riemann_solver=rs)) is None
    assert unsupported_reason(ok._replace(limiter=sc.VAN_ALBADA_PP)) is None
    assert "periodic" in unsupported_reason(ok._replace(limiter=sc.VAN_ALBADA_PP, boundary_settings=mixed))
    assert "limiter" in unsupported_reason(ok._replace(limiter=9))
    assert "Riemann" in unsupported_reason(ok._replace(riemann_solver=7))
    assert "float32" in unsupported_reason(ok._replace(numerical_precision=sc.DOUBLE_PRECISION))
    assert "module" in unsupported_reason(ok._replace(cooling_config=sc.CoolingConfig(cooling=True)))
    assert "2D and 3D" in unsupported_reason(ok._replace(dimensionality=1))


def test_time_integration_rejects_unsupported_options_before_touching_the_gpu():
    import torch
    state, cfg, params = W.workload("orszag_tang_2d", 16)
    rv = amx.get_registered_variables(cfg)
    with pytest.raises(NotImplementedError):
        amx.time_integration(state, cfg._replace(split=sc.SPLIT), params, rv)
    grad_state = torch.from_numpy(state).requires_grad_(True)
    with pytest.raises(NotImplementedError, match="snapshots"):
        amx.time_integration(grad_state, cfg._replace(return_snapshots=True), params, rv)


def header_symbols():
    text = open(os.path.join(ROOT, "include", "amx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(amx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    path = _native.LIB_PATH
    assert os.path.exists(path), "libamx.so must be built (run __graft_entry__.build())"
    lib = ctypes.CDLL(path)
    syms = header_symbols()
    assert len(syms) >= 10
    for name in syms:
        assert hasattr(lib, name), name
    # the Python binding declares a signature for every exported entry point
    assert set(syms) <= set(_native.SIGNATURES), set(syms) - set(_native.SIGNATURES)


STRUCTS = {  # C struct in include/amx.h -> ctypes mirror in _native
    "amx_fv_config": _native.AmxFvConfig, "amx_fv_run": _native.AmxFvRun, "amx_fv_stats": _native.AmxFvStats,
    "amx_fv_timing": _native.AmxFvTiming, "amx_fv_slot_status": _native.AmxFvSlotStatus,
    "amx_fd_config": _native.AmxFdConfig, "amx_diag_desc": _native.AmxDiagDesc,
    "amx_snapshot_out": _native.AmxSnapshotOut, "amx_snapshot_spec": _native.AmxSnapshotSpec,
}


def test_native_struct_layout_matches_header(tmp_path):
    """Size and every field offset of each ctypes mirror equal the C
    compiler's layout of the header's struct."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "amx.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, cls in STRUCTS.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)


def test_no_gpu_means_loud_failure(monkeypatch):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(_native.AmxError):
        _native.require_gpu()


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
