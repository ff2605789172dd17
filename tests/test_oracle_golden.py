"""The oracle (CPU restatement) against vectors written by the compiled reference.

Fixtures come from tests/golden/make_golden.py, which runs the unmodified
reference classes (Actions::Adv, Runge_Kutta_4, Explicit_EC_FP, Euler_Backward).
"""
import numpy as np
import pytest

from oracle.oracle import Config, Oracle, hind


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def make(g, **over):
    c = g["cfg"]
    kw = dict(l0=int(c["l0"]), m0=int(c["m0"]), np=int(c["np"]), nx=int(c["nx"]), ny=int(c["ny"]),
              pmax=float(c["pmax"]), dx=float(g["dx"]), dy=float(g["dy"]),
              density_np=float(c["density_np"]), zeta=float(c["zeta"]),
              small_dt=float(c["small_dt"]), nb=int(c["nb"]))
    kw.update(over)
    return Oracle(Config(**kw))


@pytest.mark.parametrize("name", ["rhs_2d_m3", "rhs_quirk_l7"])
def test_rhs_matches_reference(golden, name):
    g = golden(name)
    o = make(g)
    o.set_state(g["f_in"], g["fld_in"])
    pd, _, _ = o.axes()
    assert np.array_equal(pd, g["p"])          # same momentum axis, bit for bit
    f, fld = o.rhs()
    assert rel(f, g["f_out"]) < 1e-14
    assert rel(fld, g["fld_out"]) < 1e-14


def test_quirk_is_needed(golden):
    """Without the def-VlasovMax.cpp:918 coefficient slip the slope differs."""
    g = golden("rhs_quirk_l7")
    o = make(g, ref_quirks=0)
    o.set_state(g["f_in"], g["fld_in"])
    f, _ = o.rhs()
    assert rel(f, g["f_out"]) > 1e-4


@pytest.mark.parametrize("name,tol", [("step_2d_c3", 1e-13), ("rk4_1d_nocoll", 1e-13),
                                      ("ee_only", 0.0), ("flm_only", 0.0)])
def test_steps_match_reference(golden, name, tol):
    g = golden(name)
    c = g["cfg"]
    o = make(g)
    o.set_state(g["f_in"], g["fld_in"])
    h, n = float(c["h"]), int(c["nsteps"])
    op = int(c["op"])
    if op == 1:
        o.step(h, n)
    elif op == 2:
        for _ in range(n):
            o.rk4(h)
    else:
        t = 0.0
        for _ in range(n):
            t += h
            if op == 3:
                o.collide_ee(t)
            else:
                o.collide_flm(t, 1)
                o.collide_flm(t, 2)
    f, fld = o.get_state()
    assert rel(f, g["f_out"]) <= tol
    assert rel(fld, g["fld_out"]) <= max(tol, 1e-300) or np.abs(g["fld_out"]).max() == 0
    # and the step really changed the state
    assert rel(g["f_out"], g["f_in"]) > 1e-6


def test_m0_zero_native_matches_reference_standin(golden):
    """l0=8, m0=0 run natively by the oracle vs the reference's m0=2 stand-in."""
    g = golden("step_1d_m0")
    c = g["cfg"]
    l0 = int(c["l0"])
    o = make(g, m0=0)
    nx, np_ = int(c["nx"]), int(c["np"])
    f = np.zeros((l0 + 1, 1, nx, np_), np.complex128)
    f[:] = g["f_in"]
    fld = np.zeros((6, 1, nx), np.complex128)
    fld[0] = g["fld_in"][0]
    o.set_state(f, fld)
    o.step(float(c["h"]), int(c["nsteps"]))
    fo, fldo = o.get_state()
    assert rel(fo.real, g["f_out"]) < 1e-13
    assert np.abs(fo.imag).max() == 0.0
    assert rel(fldo[0].real, g["fld_out"][0]) < 1e-12


@pytest.mark.parametrize("name", ["imp_1d_m0", "imp_1d_m0_l8"])
def test_implicit_1d_matches_reference(golden, name):
    """Implicit-E steps (Runge_Kutta_2_Imp + Implicit_E_Field, if_implicit1D), m0=0 run natively
    by the oracle vs the reference's m0=2 y-uniform stand-in."""
    g = golden(name)
    c = g["cfg"]
    l0, nx, np_ = int(c["l0"]), int(c["nx"]), int(c["np"])
    o = make(g, m0=0)
    f = np.zeros((l0 + 1, 1, nx, np_), np.complex128)
    f[:] = g["f_in"]
    fld = np.zeros((6, 1, nx), np.complex128)
    fld[0] = g["fld_in"][0]
    o.set_state(f, fld)
    o.step_implicit_1d(float(c["h"]), int(c["nsteps"]), int(c["es"]))
    fo, fldo = o.get_state()
    assert rel(fo.real, g["f_out"]) < 1e-13
    assert np.abs(fo.imag).max() == 0.0
    assert rel(fldo[0].real, g["fld_out"][0]) < 1e-13
    assert not np.any(fldo[1:3])                         # the 1D solve sets Ey = Ez = 0
    assert rel(g["fld_out"][0], g["fld_in"][0]) > 1e-3   # the field solve changed Ex
    assert rel(g["f_out"], g["f_in"]) > 1e-6


def test_implicit_1d_field_is_the_zero_current_field(golden):
    """The 1D implicit field is the root of the linear current response: with
    E = DE J0/(J0 - J_DE), the current of the probe step with E scales linearly,
    so a second solve from the solved state changes Ex only at second order."""
    g = golden("imp_1d_m0")
    c = g["cfg"]
    l0, nx, np_ = int(c["l0"]), int(c["nx"]), int(c["np"])
    o = make(g, m0=0)
    f = np.zeros((l0 + 1, 1, nx, np_), np.complex128)
    f[:] = g["f_in"]
    fld = np.zeros((6, 1, nx), np.complex128)
    fld[0] = g["fld_in"][0]
    o.set_state(f, fld)
    o.step_implicit_1d(0.5, 1, 1)
    _, fld1 = o.get_state()
    # the solved field does not depend on the incoming Ex beyond the probe size DE ~ 1e-7 |E|
    o2 = make(g, m0=0)
    fld2 = fld.copy()
    fld2[0] *= 3.0
    o2.set_state(f, fld2)
    o2.step_implicit_1d(0.5, 1, 1)
    _, fld2o = o2.get_state()
    assert rel(fld2o[0], fld1[0]) < 1e-6
