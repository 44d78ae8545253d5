"""Planted mistakes: copies of the CPU restatement (oracle/fv_oracle.cpp) with
one plausible error each are built in a temporary directory, and each must
fail at least one of the pins in tests/test_oracle_pins.py. This shows that
the pins fix the parts of the restatement the reference's golden output
cannot: the reconstruction (slope neighbours, dx/2, primitive variables),
the Heun average, the HLLC star speed and the signs of the magnetic
sub-step's Lorentz force and induction curl; and, in copies of the
finite-difference restatement (oracle/fd_oracle.cpp), the point-to-average
weight of the CT edge EMFs, which tests/test_fd_oracle.py pins by the order
of the face update. The mistakes are planted in the oracle on the CPU only;
nothing here touches the GPU."""

import os
import subprocess

import numpy as np
import pytest

import test_oracle_pins as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "fv_oracle.cpp")

CONS_LIMIT = """        wl[v] = b + d_m1;
        wr[v] = cc - d_0;
      }
      {  // planted: limit the conserved variables and convert the face states back
        T Uq[4][5], ul[5], ur[5];
        for (int q = 0; q < 4; ++q) {
          T wq[5];
          for (int v = 0; v < ng; ++v) wq[v] = w[v * nc + id[q]];
          cons_from_prim(wq, Uq[q], g.dim, g.gamma);
        }
        for (int v = 0; v < ng; ++v) {
          const T gm = limit(g.limiter, (Uq[1][v] - Uq[0][v]) / g.dx, (Uq[2][v] - Uq[1][v]) / g.dx, g.dx);
          const T g0 = limit(g.limiter, (Uq[2][v] - Uq[1][v]) / g.dx, (Uq[3][v] - Uq[2][v]) / g.dx, g.dx);
          ul[v] = Uq[1][v] + (gm * g.dx) / T(2);
          ur[v] = Uq[2][v] - (g0 * g.dx) / T(2);
        }
        prim_from_cons(ul, wl, g.dim, g.gamma);
        prim_from_cons(ur, wr, g.dim, g.gamma);
      }
"""

# name: ([(original text, planted text)], pins that must catch it)
MUTANTS = {
    "slope_of_left_neighbour": (
        [("const T g_0 = limit(g.limiter, (cc - b) / g.dx, (d - cc) / g.dx, g.dx);",
          "const T g_0 = limit(g.limiter, (b - a) / g.dx, (cc - b) / g.dx, g.dx);")],
        ["reflection"]),
    "face_extrapolated_by_dx": (
        [("T d_m1 = (g_m1 * g.dx) / T(2);", "T d_m1 = (g_m1 * g.dx);"),
         ("T d_0 = (g_0 * g.dx) / T(2);", "T d_0 = (g_0 * g.dx);")],
        ["linear"]),
    "conserved_variable_limiting": (
        [("        wl[v] = b + d_m1;\n        wr[v] = cc - d_0;\n      }\n", CONS_LIMIT)],
        ["linear"]),
    "no_heun_average": (
        [("for (int v = 0; v < ng; ++v) um[v] = T(0.5) * (ua[v] + ub[v]);",
          "for (int v = 0; v < ng; ++v) um[v] = ub[v];")],
        ["advection"]),
    "lorentz_force_sign": (
        [("const T vn = v0[k * nc + c] - dt * phi2;", "const T vn = v0[k * nc + c] + dt * phi2;")],
        ["alfven"]),
    "induction_sign": (
        [("const T bn = B0[k * nc + c] - dt * cE[k];", "const T bn = B0[k * nc + c] + dt * cE[k];")],
        ["alfven"]),
    "electric_field_v_cross_b": (
        [("cross3(b, v, e);  // electric_field = cross(magnetic_field, velocity)",
          "cross3(v, b, e);")],
        ["alfven"]),
    "hllc_star_speed_pressure_sign": (
        [("const T s_star = (((pr - pl) + (rl * ul) * (SL - ul)) - (rr * ur) * (SR - ur)) /",
          "const T s_star = (((pl - pr) + (rl * ul) * (SL - ul)) - (rr * ur) * (SR - ur)) /")],
        ["contact", "shock_tube_hllc"]),
}


def run_pin(fv, pin):
    if pin == "reflection":
        for args in [(2, 1, False), (2, 2, True), (3, 1, True), (3, 3, False)]:
            P.test_step_commutes_with_reflections(fv, *args)
    elif pin == "linear":
        for args in [(2, 1), (2, 2), (3, 1), (3, 2), (3, 3)]:
            P.test_stage_is_exact_flux_difference_on_linear_data(fv, *args)
    elif pin == "advection":
        P.test_hydro_step_is_second_order_on_smooth_advection(fv)
    elif pin == "alfven":
        P.test_mhd_step_is_second_order_on_circularly_polarised_alfven_wave(fv)
    elif pin == "contact":
        P.test_hllc_keeps_a_moving_contact_sharper_than_hll(fv)
    elif pin == "shock_tube_hllc":
        for name in ("sod", "toro3"):
            P.test_shock_tubes_converge_to_the_exact_riemann_solution(fv, name, 1)
    else:
        raise KeyError(pin)


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    """Builds every mutant in parallel (g++ with the oracle Makefile's flags)."""
    src = open(SRC).read()
    d = tmp_path_factory.mktemp("mutants")
    procs = {}
    for name, (edits, _) in MUTANTS.items():
        text = src
        for old, new in edits:
            assert text.count(old) == 1, f"{name}: the planted site is no longer in fv_oracle.cpp"
            text = text.replace(old, new)
        cpp = d / f"{name}.cpp"
        cpp.write_text(text)
        so = d / f"lib{name}.so"
        procs[name] = (so, subprocess.Popen(
            ["g++", "-O1", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c++17",
             "-shared", "-o", str(so), str(cpp)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    libs = {}
    for name, (so, p) in procs.items():
        out, _ = p.communicate(timeout=300)
        assert p.returncode == 0, out.decode()
        libs[name] = str(so)
    return libs


def test_unmodified_restatement_passes_every_pin(oracle):
    for pin in ("reflection", "linear", "advection", "alfven", "contact", "shock_tube_hllc"):
        run_pin(oracle, pin)


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_fails_a_pin(oracle, mutant_libs, name):
    _, pins = MUTANTS[name]
    oracle.use_library(mutant_libs[name])
    try:
        caught = []
        for pin in pins:
            try:
                with np.errstate(all="ignore"):
                    run_pin(oracle, pin)
            except (AssertionError, ValueError, RuntimeError, FloatingPointError) as e:
                caught.append((pin, type(e).__name__))
        assert caught, f"planted mistake {name} passed the pins {pins}"
    finally:
        oracle.use_library(None)


FD_SRC = os.path.join(ROOT, "oracle", "fd_oracle.cpp")
P2A = """  const T sa = ((q[g.nb(c, A, -1)] - T(2) * q[c]) + q[g.nb(c, A, 1)]) / T(24);
  const T sb = ((q[g.nb(c, B, -1)] - T(2) * q[c]) + q[g.nb(c, B, 1)]) / T(24);"""
FD_MUTANTS = {
    "ct_weight_one_12th": P2A.replace("/ T(24)", "/ T(12)"),
    "ct_weight_one_48th": P2A.replace("/ T(24)", "/ T(48)"),
    "ct_no_point_to_average": P2A.replace("/ T(24)", "* T(0)"),
}


@pytest.fixture(scope="module")
def fd_mutant_libs(tmp_path_factory):
    src = open(FD_SRC).read()
    assert src.count(P2A) == 1, "the planted site is no longer in fd_oracle.cpp"
    d = tmp_path_factory.mktemp("fd_mutants")
    procs = {}
    for name, planted in FD_MUTANTS.items():
        cpp = d / f"{name}.cpp"
        cpp.write_text(src.replace(P2A, planted))
        so = d / f"lib{name}.so"
        procs[name] = (so, subprocess.Popen(
            ["g++", "-O1", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c++17",
             "-shared", "-o", str(so), str(cpp)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    libs = {}
    for name, (so, p) in procs.items():
        out, _ = p.communicate(timeout=300)
        assert p.returncode == 0, out.decode()
        libs[name] = str(so)
    return libs


@pytest.mark.parametrize("name", sorted(FD_MUTANTS))
def test_planted_ct_weight_fails_the_order_pin(oracle, fd_mutant_libs, name):
    from oracle import fd
    import test_fd_oracle as T
    fd.use_library(fd_mutant_libs[name])
    try:
        with pytest.raises(AssertionError):
            T.test_ct_point_to_average_weight_is_one_24th(fd)
    finally:
        fd.use_library(None)
