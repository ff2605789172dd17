"""The oracle's sources (def-PLSource.cpp) against vectors written by the compiled reference.

Fixtures pls_only, ib_only, efield_only (3 calls of PLaserSource::Laser_Source,
InverseBremsstrahlung::loop, ExternalEfield::Laser_Efield) and step_src (2 full
steps of main.cpp:79-107 with the laser source and inverse bremsstrahlung on,
:90-92) come from tests/golden/make_golden.py.  The cell positions are the
reference's own axes (oracle.deck_axis, def-Input.cpp:313-325).  Plus a check
the fixtures cannot make: inverse bremsstrahlung keeps the density exactly
(it renormalises, def-PLSource.cpp:836-850) and heats.
"""
import numpy as np
import pytest

from oracle.oracle import SOURCE_KEYS, Config, Oracle, deck_axis


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def src_of(c):
    return {k: float(c["src_" + k]) for k in SOURCE_KEYS}


def oracle_for(g):
    c = g["cfg"]
    o = Oracle(Config(l0=int(c["l0"]), m0=int(c["m0"]), np=int(c["np"]), nx=int(c["nx"]), ny=int(c["ny"]),
                      pmax=float(c["pmax"]), dx=float(g["dx"]), dy=float(g["dy"]),
                      density_np=float(c["density_np"]), zeta=float(c["zeta"]), small_dt=float(c["small_dt"]),
                      nb=int(c["nb"])))
    x, xg = deck_axis(int(c["nx"]), float(c["xmin"]), float(c["xmax"]))
    y, _ = deck_axis(int(c["ny"]), float(c["ymin"]), float(c["ymax"]))
    o.set_state(g["f_in"], g["fld_in"])
    o.set_sources(src_of(c), x, y, xg)
    return o


@pytest.mark.parametrize("name,tol", [("pls_only", 1e-15), ("ib_only", 1e-15), ("efield_only", 1e-15),
                                      ("step_src", 1e-13)])
def test_sources_match_reference(golden, name, tol):
    g = golden(name)
    c = g["cfg"]
    o = oracle_for(g)
    h, n, op = float(c["h"]), int(c["nsteps"]), int(c["op"])
    if op == 1:
        o.step(h, n)
    else:
        t = 0.0
        for _ in range(n):
            t += h
            if op == 6:
                assert o.laser_source(t)
            elif op == 7:
                o.inverse_bremsstrahlung(t)
            else:
                o.laser_efield(t)
    f, fld = o.get_state()
    assert rel(f, g["f_out"]) <= tol, rel(f, g["f_out"])
    assert rel(fld, g["fld_out"]) <= tol, rel(fld, g["fld_out"])


def test_inverse_bremsstrahlung_keeps_density_and_heats(golden):
    """The renormalisation keeps the p-trapezoid of p^2 f00 from p0 to pmax (numpy's trapezoid rule
    on the p axis, def-PLSource.cpp:836-850) and the absorbed light heats every cell."""
    g = golden("ib_only")
    o = oracle_for(g)
    pd, _, _ = o.axes()
    m0 = o.moments()
    f0, _ = o.get_state()
    o.inverse_bremsstrahlung(0.5)
    o.inverse_bremsstrahlung(1.0)
    f1, _ = o.get_state()
    m1 = o.moments()
    n_before = np.trapezoid(pd ** 2 * f0[0].real, pd, axis=-1)
    n_after = np.trapezoid(pd ** 2 * f1[0].real, pd, axis=-1)
    assert np.abs(n_after / n_before - 1).max() < 1e-13
    assert np.abs(f1[0] - f0[0]).max() > 1e-4 * np.abs(f0[0]).max()  # it acted
    assert np.all(m1[10] > m0[10])  # <v^2> grows: the laser heats
