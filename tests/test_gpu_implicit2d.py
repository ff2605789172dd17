"""Implicit-E step without if_implicit1D on the GPU (oshun_step_implicit_2d:
Runge_Kutta_2_Imp + Implicit_E_Field with the 3x3 conductivity solve, e-e,
sources, implicit f_lm) through the C ABI.

* the compiled reference's fixtures imp_2d_m2 (Ampere JN, B and E nonzero),
  imp_2d_m3_es (if_implicitES, JN = the input Ex), imp_2d_l7_quirk (the
  y-advection slip in Adv1), imp_2d_dilute (two cells with a zero determinant:
  4 attempts, E = 0 there; the ill-conditioned diluted block itself is not
  compared) and imp_2d_src (laser heating and inverse bremsstrahlung);
* the oracle at larger shapes: l0 = 6, m0 = 3, 2D (2 steps), and a POP-like
  harmonic count (l0 = 16, m0 = 3) on a small grid.

Tolerance.  The probe field is DE = sqrt(16 eps) |E| + 16 eps, so the solve
turns a rounding difference of one ulp into 1e-9..1e-5 in E after a step or
two: the reference algorithm's own noise floor, measured per case on the oracle
(the largest change of its output under a one-ulp perturbation of f, 3 draws).
The GPU forms its multipliers directly from the coefficient tables (the oracle
and the reference use running products) and its collision kernels sum in
another order, so f and the fields are held to FLOOR_MULT times that floor
(B too: from the second step on, Faraday moves B with the solved E).
"""
import numpy as np
import pytest

from oracle.oracle import SOURCE_KEYS, Config, Oracle, deck_axis
from test_gpu_parity import TOL, ctx_from_golden, rel

pytestmark = pytest.mark.gpu

FLOOR_MULT = 5.0


def _mask(c):
    m = np.ones((int(c["ny"]), int(c["nx"])), bool)
    if "dilute_block" in c:
        y0, y1, x0, x1 = (int(v) for v in c["dilute_block"])
        m[y0:y1, x0:x1] = False
    return m


def _oracle_of(c, dx, dy, f, fld, es, jn, steps, src_axes=None):
    o = Oracle(Config(l0=int(c["l0"]), m0=int(c["m0"]), np=int(c["np"]), nx=int(c["nx"]), ny=int(c["ny"]),
                      pmax=float(c["pmax"]), dx=dx, dy=dy, density_np=float(c["density_np"]),
                      zeta=float(c.get("zeta", 1.0)), small_dt=float(c["small_dt"]), nb=int(c["nb"])))
    o.set_state(f, fld)
    if src_axes:
        o.set_sources(*src_axes)
    if es:
        o.set_es_current(jn)
    o.step_implicit_2d(float(c["h"]), steps, es)
    return o.get_state()


def _floor(c, dx, dy, f, fld, es, jn, steps, mask, src_axes=None):
    f0, e0 = _oracle_of(c, dx, dy, f, fld, es, jn, steps, src_axes)
    rng = np.random.default_rng(7)
    wf, we = 0.0, np.zeros(6)
    for _ in range(3):
        f1, e1 = _oracle_of(c, dx, dy, f * (1.0 + 2.220446049250313e-16 * rng.choice([-1.0, 1.0], f.shape)), fld, es,
                            jn, steps, src_axes)
        wf = max(wf, rel(f1[:, mask], f0[:, mask]))
        we = np.maximum(we, [rel(e1[k][mask], e0[k][mask]) for k in range(6)])
    return (f0, e0), wf, we


def _check(fg, fldg, fref, fldref, wf, we, mask):
    ef = rel(fg[:, mask], fref[:, mask])
    ee = [rel(fldg[k].real[mask], fldref[k].real[mask]) for k in range(6)]
    print(f"  f {ef:.2e} (floor {wf:.1e}), E,B {['%.1e' % v for v in ee]} (floors {['%.1e' % v for v in we]})",
          flush=True)
    assert wf < 1e-10 and np.all(we < 1e-4)
    assert ef <= max(TOL, FLOOR_MULT * wf)
    for k in range(6):
        assert ee[k] <= max(TOL, FLOOR_MULT * we[k]), k


@pytest.mark.parametrize("name", ["imp_2d_m2", "imp_2d_m3_es", "imp_2d_l7_quirk", "imp_2d_dilute", "imp_2d_src"])
def test_implicit_2d_vs_reference(golden, name):
    g = golden(name)
    c = g["cfg"]
    es, steps = int(c["es"]), int(c["nsteps"])
    print(f"case implicit 2D {name}", flush=True)
    ctx = ctx_from_golden(g)
    ctx.upload(g["f_in"], g["fld_in"])
    src_axes = None
    if "src_I_0" in c:
        x, xg = deck_axis(int(c["nx"]), float(c["xmin"]), float(c["xmax"]))
        y, _ = deck_axis(int(c["ny"]), float(c["ymin"]), float(c["ymax"]))
        src_axes = ({k: float(c["src_" + k]) for k in SOURCE_KEYS}, x, y, xg)
        ctx.set_sources(*src_axes)
    jn = np.stack([g["fld_in"][0]] * 3)  # the Ex the reference's solver was built with
    if es:
        ctx.set_implicit_es_current(jn)
    z = ctx.step_implicit_2d(float(c["h"]), steps, es)
    fg, fldg = ctx.download()
    mask = _mask(c)
    _, wf, we = _floor(c, float(g["dx"]), float(g["dy"]), g["f_in"], g["fld_in"], es, jn, steps, mask, src_axes)
    _check(fg, fldg, g["f_out"], g["fld_out"], wf, we, mask)
    if "zero_cells" in c:
        cells = sorted(tuple(int(v) for v in p) for p in c["zero_cells"])
        assert z == len(cells)
        zero = np.all(fldg[:3] == 0, axis=0)
        assert sorted(zip(*np.nonzero(zero))) == cells
    else:
        assert z == 0
    assert ctx.time() == float(c["h"]) * steps


@pytest.mark.parametrize("l0,m0,np_,nx,ny,steps,es", [(6, 3, 48, 16, 12, 2, 0), (16, 3, 64, 8, 6, 1, 0),
                                                      (5, 2, 40, 12, 10, 2, 1)],
                         ids=["l6m3", "l16m3", "l5m2_es"])
def test_implicit_2d_vs_oracle(l0, m0, np_, nx, ny, steps, es):
    import oshun_amd
    from test_gpu_parity import _maxwellian_state
    print(f"case implicit 2D vs oracle l0={l0} m0={m0} np={np_} {nx}x{ny} steps={steps} es={es}", flush=True)
    c = dict(l0=l0, m0=m0, np=np_, nx=nx, ny=ny, pmax=0.25, density_np=1e21, small_dt=0.2, nb=4, h=0.5)
    dx, dy = 3.0, 2.5
    f, fld = _maxwellian_state(l0, m0, np_, nx, ny, c["pmax"], seed=70 + l0)
    x = np.arange(nx) / nx
    y = np.arange(ny) / ny
    fld[0] = 1e-4 * np.sin(2 * np.pi * x)[None, :] * (1.0 + 0.3 * np.cos(2 * np.pi * y))[:, None]
    fld[1] = 2e-5 * np.cos(2 * np.pi * x)[None, :] * np.ones((ny, 1))
    fld[2] = -1e-5 * np.sin(2 * np.pi * y)[:, None] * np.ones((1, nx))
    fld[3] = 1e-3 * np.cos(2 * np.pi * y)[:, None] * np.ones((1, nx))
    fld[4] = -1.5e-3 * np.sin(2 * np.pi * x)[None, :] * np.ones((ny, 1))
    fld[5] = 2e-3
    jn = np.zeros((3, ny, nx), np.complex128)
    jn[0] = 0.3 * fld[0]
    ctx = oshun_amd.Context(l0=l0, m0=m0, nump=np_, numx=nx, numy=ny, pmax=c["pmax"], dx=dx, dy=dy,
                            density_np=c["density_np"], small_dt=c["small_dt"], nb_algorithms=c["nb"])
    ctx.upload(f, fld)
    if es:
        ctx.set_implicit_es_current(jn)
    assert ctx.step_implicit_2d(c["h"], steps, es) == 0
    fg, fldg = ctx.download()
    mask = np.ones((ny, nx), bool)
    (fo, fldo), wf, we = _floor(c, dx, dy, f, fld, es, jn, steps, mask)
    _check(fg, fldg, fo, fldo, wf, we, mask)


def test_implicit_2d_rejects_m0_below_2():
    import oshun_amd
    ctx = oshun_amd.Context(l0=3, m0=1, nump=16, numx=4, numy=3, pmax=0.25)
    with pytest.raises(oshun_amd.OshunError):
        ctx.step_implicit_2d(0.1, 1, 0)
