"""The sources (def-PLSource.cpp) on the GPU against the compiled reference's fixtures and the oracle.

* pls_only, ib_only, efield_only (3 calls each) and step_src (2 full steps with the
  laser source and inverse bremsstrahlung inside, main.cpp:90-92): the GPU through
  the C ABI against the reference's outputs, <= 1e-12 normwise in f and fields.
* a larger 2D case (Np = 128, 32 x 8 cells, 4 steps of the full step with both
  sources) against the oracle at <= 1e-12, in FP64, and in the FP32 split against
  the oracle run on the GPU's own rounded input within the split's bound.
"""
import numpy as np
import pytest

from oracle.oracle import SOURCE_KEYS, Config, Oracle, deck_axis
from test_gpu_parity import TOL, ctx_from_golden, rel

pytestmark = pytest.mark.gpu


def _src(c):
    return {k: float(c["src_" + k]) for k in SOURCE_KEYS}


def _axes(c):
    x, xg = deck_axis(int(c["nx"]), float(c["xmin"]), float(c["xmax"]))
    y, _ = deck_axis(int(c["ny"]), float(c["ymin"]), float(c["ymax"]))
    return x, y, xg


@pytest.mark.parametrize("name", ["pls_only", "ib_only", "efield_only", "step_src"])
def test_sources_vs_reference(golden, name):
    g = golden(name)
    c = g["cfg"]
    print(f"case sources {name}", flush=True)
    ctx = ctx_from_golden(g)
    ctx.upload(g["f_in"], g["fld_in"])
    ctx.set_sources(_src(c), *_axes(c))
    h, n, op = float(c["h"]), int(c["nsteps"]), int(c["op"])
    if op == 1:
        ctx.step(h, n)
    else:
        t = 0.0
        for _ in range(n):
            t += h
            {6: ctx.laser_source, 7: ctx.inverse_bremsstrahlung, 8: ctx.laser_efield}[op](t)
    f, fld = ctx.download()
    assert rel(f, g["f_out"]) < TOL, rel(f, g["f_out"])
    assert rel(fld, g["fld_out"]) < TOL, rel(fld, g["fld_out"])


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32split"])
def test_step_with_sources_vs_oracle(precision):
    import oshun_amd
    from test_gpu_parity import _maxwellian_state
    l0, m0, np_, nx, ny = 3, 1, 128, 32, 8
    print(f"case step with sources l0={l0} m0={m0} np={np_} {nx}x{ny} precision {precision}", flush=True)
    src = dict(polynomial_Ivst=1, inverse_bremsstrahlung=1, I_0=3e14, lambda_0=0.351, ab_eta=0.5, ab_depth=30.0,
               cntr_x0=40.0, cntr_r0=15.0, spot_w0=20.0, rise_time=0.8, flat_time=3.0, fall_time=1.0,
               smaller_dt=0.1)
    x, xg = deck_axis(nx, 0.0, 100.0)
    y, _ = deck_axis(ny, 0.0, 30.0)
    dx, dy = xg[1] - xg[0], y[1] - y[0]
    f, fld = _maxwellian_state(l0, m0, np_, nx, ny, 0.3, seed=61)
    kw = dict(pmax=0.3, dx=dx, dy=dy, density_np=1e21, small_dt=0.25)
    ctx = oshun_amd.Context(l0, m0, np_, nx, ny, nb_algorithms=4, precision=precision, **kw)
    ctx.upload(f, fld)
    f_in, fld_in = ctx.download()  # the state the GPU holds (float-rounded in the split)
    ctx.set_sources(src, x, y, xg)
    ctx.step(0.5, 4)
    fg, fldg = ctx.download()
    o = Oracle(Config(l0=l0, m0=m0, np=np_, nx=nx, ny=ny, nb=4, **kw))
    o.set_state(f_in, fld_in)
    o.set_sources(src, x, y, xg)
    o.step(0.5, 4)
    fo, fldo = o.get_state()
    assert np.abs(fo[0] - f_in[0]).max() > 1e-3 * np.abs(f_in[0]).max()  # the sources acted
    if precision == 0:
        assert rel(fg, fo) < TOL, rel(fg, fo)
        assert rel(fldg.real, fldo.real) < TOL, rel(fldg.real, fldo.real)
    else:
        e00 = float(np.linalg.norm(fg[0] - fo[0]) / np.linalg.norm(fo[0]))
        print(f"FP32 split vs oracle after 4 steps: f00 {e00:.1e}", flush=True)
        assert e00 < 1e-5
