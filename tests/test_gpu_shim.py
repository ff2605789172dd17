"""The drop-in boundary on the GPU: the reference's own classes (Runge_Kutta_4,
Runge_Kutta_2_Imp, Explicit_EC_FP, Euler_Backward, Implicit_E_Field), driven
by the reference-class test driver exactly as in main_loop (main.cpp:79-107),
with the shim (oshun_amd/shim/oshun_shim.cpp) linked ahead of the reference
translation units, so that every step runs on the GPU through the C ABI
(oracle/_ref/oshun_ref_shim, built by `make -C oracle shim`).

* RK4 decks (fixtures step_2d_c3, step_1d_m0, and step_src with the laser
  source and inverse bremsstrahlung on) and the collision-only drivers
  (ee_only, flm_only): the shim's output against the compiled reference's
  fixture at <= 1e-12 normwise in f and the fields.
* mirror-boundary decks (mirror_x_2d, mirror_y_2d): the whole arrays, guards
  included, against the fixture at <= 1e-12 (the device keeps the guards and
  refreshes them in the shim's Euler_Backward::advance_lm).
* implicit-E decks, 1D variant (imp_1d_m0, imp_1d_m0_l8) and 2D variant
  (imp_2d_m2; imp_2d_m3_es, whose JN is the Ex the shim's RK2 constructor
  copied): bitwise equal to the same steps called directly through the C ABI
  (oshun_step_implicit_1d/_2d); those paths are held to the fixtures in
  test_gpu_implicit.py and test_gpu_implicit2d.py.
Fixtures of m0 = 0 physics ran the reference as an m0 = 2 y-uniform stand-in;
the shim gets the same stand-in deck and input.
"""
import os

import numpy as np
import pytest

from oracle.oracle import SHIM_BIN, hind, num_harmonics, run_ref
from test_gpu_parity import TOL, ctx_from_golden, rel

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(SHIM_BIN), reason="shim driver not built (make -C oracle shim)")]


def _standin(g):
    """input of the fixture's reference run (m0 = 2 stand-in where the physics is m0 = 0; such
    fixtures keep the m = 0 harmonics and Ex only)"""
    c = g["cfg"]
    l0, m0 = int(c["l0"]), int(c.get("standin_m0", c["m0"]))
    f, fld = g["f_in"], g["fld_in"]
    if m0 != int(c["m0"]):
        full = np.zeros((num_harmonics(l0, m0),) + f.shape[1:], np.complex128)
        for l in range(l0 + 1):
            full[hind(l, 0, m0)] = f[hind(l, 0, int(c["m0"]))]
        f = full
    if fld.shape[0] != 6:
        six = np.zeros((6,) + fld.shape[1:], np.complex128)
        six[:fld.shape[0]] = fld
        fld = six
    return l0, m0, f, fld


def _run_shim(g):
    c = g["cfg"]
    l0, m0, f, fld = _standin(g)
    out = run_ref(l0, m0, int(c["np"]), int(c["nx"]), int(c["ny"]), f, fld, int(c["op"]),
                  nsteps=int(c["nsteps"]), nb=int(c["nb"]), pmax=float(c["pmax"]), xmin=float(c["xmin"]),
                  xmax=float(c["xmax"]), ymin=float(c["ymin"]), ymax=float(c["ymax"]), h=float(c["h"]),
                  small_dt=float(c["small_dt"]), density_np=float(c["density_np"]), zeta=float(c["zeta"]),
                  es=int(c.get("es", 0)), binary=SHIM_BIN,
                  src={k[4:]: float(v) for k, v in c.items() if k.startswith("src_")} or None,
                  bnd=(int(c.get("bndX", 0)), int(c.get("bndY", 0))), nbc=int(c.get("nbc", 4)))
    if "bndX" in c:
        return out
    fo, flo = out["f"], out["fld"]
    if m0 != int(c["m0"]):
        assert np.abs(np.delete(fo, [hind(l, 0, m0) for l in range(l0 + 1)], axis=0)).max() == 0.0
        fo = fo[[hind(l, 0, m0) for l in range(l0 + 1)]]
    return fo, flo[:g["fld_out"].shape[0]]


@pytest.mark.parametrize("name", ["step_2d_c3", "step_1d_m0", "ee_only", "flm_only", "step_src"])
def test_reference_classes_through_the_shim_match_the_reference(golden, name):
    g = golden(name)
    print(f"case shim {name}", flush=True)
    f, fld = _run_shim(g)
    assert rel(f, g["f_out"]) < TOL
    if np.abs(g["fld_out"]).max() > 0:
        assert rel(fld.real, g["fld_out"].real) < TOL


@pytest.mark.parametrize("name", ["imp_1d_m0", "imp_1d_m0_l8"])
def test_implicit_e_through_the_shim_equals_the_abi(golden, name):
    g = golden(name)
    c = g["cfg"]
    print(f"case shim {name}", flush=True)
    f, fld = _run_shim(g)
    ctx = ctx_from_golden(g)
    _, _, _, fld6 = _standin(g)
    ctx.upload(g["f_in"], fld6)
    ctx.step_implicit_1d(float(c["h"]), int(c["nsteps"]), int(c["es"]))
    fa, flda = ctx.download()
    assert np.array_equal(f, fa)
    assert np.array_equal(fld, flda[:fld.shape[0]])


@pytest.mark.parametrize("name", ["imp_2d_m2", "imp_2d_m3_es"])
def test_implicit_e_2d_through_the_shim_equals_the_abi(golden, name):
    g = golden(name)
    c = g["cfg"]
    print(f"case shim {name}", flush=True)
    f, fld = _run_shim(g)
    ctx = ctx_from_golden(g)
    ctx.upload(g["f_in"], g["fld_in"])
    if int(c["es"]):
        ctx.set_implicit_es_current(np.stack([g["fld_in"][0]] * 3))
    ctx.step_implicit_2d(float(c["h"]), int(c["nsteps"]), int(c["es"]))
    fa, flda = ctx.download()
    assert np.array_equal(f, fa)
    assert np.array_equal(fld, flda)


@pytest.mark.parametrize("name", ["mirror_x_2d", "mirror_y_2d"])
def test_mirror_deck_through_the_shim(golden, name):
    g = golden(name)
    print(f"case shim {name}", flush=True)
    out = _run_shim(g)
    assert rel(out["f_ext"], g["f_ext"]) < TOL
    assert rel(out["fld_ext"].real, g["fld_ext"].real) < TOL
    assert rel(out["f"], g["f_out"]) < TOL


@pytest.mark.parametrize("name", ["mirror_imp_x", "mirror_imp_x_es"])
def test_mirror_implicit_deck_through_the_shim_equals_the_abi(golden, name):
    """the production deck's integrator and boundary (implicit E, 2D variant, x mirror,
    RKLevel-2 guards) through the reference's classes: bitwise the ABI's steps"""
    import oshun_amd
    g = golden(name)
    c = g["cfg"]
    nb = int(c["nbc"])
    print(f"case shim {name}", flush=True)
    out = _run_shim(g)
    ctx = oshun_amd.Context(l0=int(c["l0"]), m0=int(c["m0"]), nump=int(c["np"]), numx=int(c["nx"]),
                            numy=int(c["ny"]), pmax=float(c["pmax"]), dx=float(g["dx"]), dy=float(g["dy"]),
                            density_np=float(c["density_np"]), zeta=float(c["zeta"]), small_dt=float(c["small_dt"]),
                            nb_algorithms=int(c["nb"]), bndx=int(c["bndX"]), bndy=int(c["bndY"]), nbc=nb)
    ctx.upload(g["f_ext0"], g["fld_ext0"], nbc=nb)
    if int(c["es"]):
        ctx.set_implicit_es_current(np.stack([g["fld_ext0"][0]] * 3), nbc=nb)
    ctx.step_implicit_2d(float(c["h"]), int(c["nsteps"]), int(c["es"]))
    f, fld = ctx.download(nbc=nb)
    assert np.array_equal(out["f_ext"], f)
    assert np.array_equal(out["fld_ext"], fld)
