"""The oracle's implicit-E step without if_implicit1D (the 3x3 field solve of
Implicit_E_Field::advance, def-ImplicitE.cpp:472-606) against fixtures written
by the compiled reference (harness op 9, tests/golden/make_golden.py).

Tolerance.  The probe field is DE = sqrt(16 eps) |E| + 16 eps, so each column of
the conductivity keeps ~7 of the current's digits, and the solve turns a
rounding difference of one ulp into ~1e-9..1e-5 in E after one or two steps
(the "floor": the largest change of the oracle's output when its input f is
perturbed by one ulp, 3 draws, measured per case here).  The oracle forms
multipliers as the reference does (running products in Exyz and the probe
couplings) except for the advection and B-rotation ones, so it is held to the
floor times FLOOR_MULT, and each floor must be small enough to mean something.
The dilute case is ill-conditioned inside its diluted block by construction,
so there only the cells outside the block are compared, plus the exact zero
field of the two cells whose determinant underflows.
"""
import numpy as np
import pytest

from oracle.oracle import SOURCE_KEYS, Config, Oracle, deck_axis

FLOOR_MULT = 5.0
CASES = ["imp_2d_m2", "imp_2d_m3_es", "imp_2d_l7_quirk", "imp_2d_dilute", "imp_2d_src"]


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def run_oracle(g, f_in, fld_in=None, es=None, jn=None, nsteps=None):
    c = g["cfg"]
    fld_in = g["fld_in"] if fld_in is None else fld_in
    es = int(c["es"]) if es is None else es
    nsteps = int(c["nsteps"]) if nsteps is None else nsteps
    o = Oracle(Config(l0=int(c["l0"]), m0=int(c["m0"]), np=int(c["np"]), nx=int(c["nx"]), ny=int(c["ny"]),
                      pmax=float(c["pmax"]), dx=float(g["dx"]), dy=float(g["dy"]),
                      density_np=float(c["density_np"]), zeta=float(c["zeta"]), small_dt=float(c["small_dt"]),
                      nb=int(c["nb"])))
    o.set_state(f_in, fld_in)
    if "src_I_0" in c:
        x, xg = deck_axis(int(c["nx"]), float(c["xmin"]), float(c["xmax"]))
        y, _ = deck_axis(int(c["ny"]), float(c["ymin"]), float(c["ymax"]))
        o.set_sources({k: float(c["src_" + k]) for k in SOURCE_KEYS}, x, y, xg)
    if es:
        # JN of the electrostatic variant: the Ex the solver was constructed with
        o.set_es_current(np.stack([fld_in[0]] * 3) if jn is None else jn)
    z = o.step_implicit_2d(float(c["h"]), nsteps, es)
    f, fld = o.get_state()
    return f, fld, z


def outside(g):
    """Mask [ny, nx] of the cells compared: all, or those outside the dilute block."""
    c = g["cfg"]
    m = np.ones((int(c["ny"]), int(c["nx"])), bool)
    if "dilute_block" in c:
        y0, y1, x0, x1 = (int(v) for v in c["dilute_block"])
        m[y0:y1, x0:x1] = False
    return m


def floor(g, mask):
    f0, e0, _ = run_oracle(g, g["f_in"])
    rng = np.random.default_rng(7)
    wf, we = 0.0, np.zeros(6)
    for _ in range(3):
        f1, e1, _ = run_oracle(g, g["f_in"] * (1.0 + 2.220446049250313e-16 * rng.choice([-1.0, 1.0], g["f_in"].shape)))
        wf = max(wf, rel(f1[:, mask], f0[:, mask]))
        we = np.maximum(we, [rel(e1[k][mask], e0[k][mask]) for k in range(6)])
    return wf, we


@pytest.mark.parametrize("name", CASES)
def test_implicit_2d_matches_reference(golden, name):
    g = golden(name)
    c = g["cfg"]
    f, fld, z = run_oracle(g, g["f_in"])
    mask = outside(g)
    wf, we = floor(g, mask)
    ef = rel(f[:, mask], g["f_out"][:, mask])
    ee = [rel(fld[k][mask], g["fld_out"][k][mask]) for k in range(6)]
    print(name, f"f {ef:.1e} (floor {wf:.1e})", "E/B", [f"{a:.1e}/{b:.1e}" for a, b in zip(ee, we)])
    assert wf < 1e-10 and np.all(we < 1e-4)
    assert ef <= max(1e-12, FLOOR_MULT * wf)
    for k in range(6):
        assert ee[k] <= max(1e-12, FLOOR_MULT * we[k]), k
    # the solve moved all three components of E
    for k in range(3):
        assert rel(g["fld_out"][k][mask], g["fld_in"][k][mask]) > 1e-3, k
    if "zero_cells" in c:
        cells = [tuple(int(v) for v in p) for p in c["zero_cells"]]
        assert z == len(cells)
        for (iy, ix) in cells:
            assert not np.any(fld[:3, iy, ix]) and not np.any(g["fld_out"][:3, iy, ix])
        zero = np.all(fld[:3] == 0, axis=0)
        assert sorted(zip(*np.nonzero(zero))) == sorted(cells)
    else:
        assert z == 0


def test_implicit_2d_field_is_the_zero_current_field(golden):
    """The solved E is the root of the linearised current response: in the
    electrostatic variant (no Faraday, so the incoming E reaches the solve only
    through the probe size DE ~ 1e-7 |E|), one step from the same f with the
    incoming E scaled by 3 gives the same E to the probe's second order, while
    the step's own change of E is O(1)."""
    g = golden("imp_2d_m2")
    jn = np.zeros((3,) + g["fld_in"].shape[1:], np.complex128)
    _, e1, _ = run_oracle(g, g["f_in"], es=1, jn=jn, nsteps=1)
    fld3 = g["fld_in"].copy()
    fld3[:3] *= 3.0
    _, e3, _ = run_oracle(g, g["f_in"], fld_in=fld3, es=1, jn=jn, nsteps=1)
    for k in range(3):
        assert rel(e3[k], e1[k]) < 1e-6, k
        assert rel(e1[k], g["fld_in"][k]) > 1e-2, k


def test_implicit_2d_rejects_m0_below_2():
    o = Oracle(Config(l0=3, m0=1, np=12, nx=4, ny=3, pmax=0.25, dx=1.0, dy=1.0))
    with pytest.raises(ValueError):
        o.step_implicit_2d(0.1, 1, 0)
