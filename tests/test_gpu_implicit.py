"""Implicit-E step of the 1D variant on the GPU (oshun_step_implicit_1d:
Runge_Kutta_2_Imp + Implicit_E_Field with if_implicit1D, e-e, implicit f_lm)
against the compiled reference's fixtures (m0 = 2 y-uniform stand-in) and
against the oracle at larger shapes, with and without the electrostatic
switch (Faraday in Adv1 moves B when Ey, Ez are nonzero).

Tolerance.  The 1D field solve is Ex = DE J0 / (J0 - J_DE) with a probe field
DE = sqrt(16 eps) |E| + 16 eps, so J0 - J_DE keeps only ~7 of J0's digits: the
reference algorithm itself turns a rounding difference of one ulp in its input
into ~1e-12 in f and ~4e-8 in Ex after three steps (measured here on the oracle,
which matches the reference bit for bit in f00, f10 and Ex).  The collision
kernels sum in a different order than the reference (scans, contracted
multiply-adds: a few ulp) and the coefficient tables of the 1D kernels are
formed on the host (an ulp from the oracle's complex ones in places), so f and
Ex are held to a multiple of that measured noise floor; the magnetic fields,
which do not pass through the solve, to the usual 1e-12."""
import numpy as np
import pytest

from test_gpu_parity import TOL, _maxwellian_state, ctx_from_golden, rel

pytestmark = pytest.mark.gpu

KW = dict(pmax=0.3, dx=0.4, density_np=1e22, small_dt=0.2, nb_algorithms=6)
FLOOR_MULT = 5.0


def _oracle(l0, np_, nx, pmax, dx, density_np, small_dt, nb, coll=1):
    from oracle.oracle import Config, Oracle
    return Oracle(Config(l0=l0, m0=0, np=np_, nx=nx, pmax=pmax, dx=dx, density_np=density_np, small_dt=small_dt,
                         nb=nb, coll_ee=coll, coll_flm=coll))


def _noise_floor(make, f, fld, h, steps, es):
    """Largest change of the oracle's f and Ex when its input is perturbed by one ulp (3 draws)."""
    o = make()
    o.set_state(f, fld)
    o.step_implicit_1d(h, steps, es)
    f0, fld0 = o.get_state()
    rng = np.random.default_rng(7)
    wf = we = 0.0
    for _ in range(3):
        o = make()
        o.set_state(f * (1.0 + 2.220446049250313e-16 * rng.choice([-1.0, 1.0], f.shape)), fld)
        o.step_implicit_1d(h, steps, es)
        f1, fld1 = o.get_state()
        wf, we = max(wf, rel(f1, f0)), max(we, rel(fld1[0], fld0[0]))
    return wf, we


@pytest.mark.parametrize("name", ["imp_1d_m0", "imp_1d_m0_l8"])
def test_implicit_1d_vs_reference(golden, name):
    g = golden(name)
    c = g["cfg"]
    l0, nx, np_ = int(c["l0"]), int(c["nx"]), int(c["np"])
    h, steps, es = float(c["h"]), int(c["nsteps"]), int(c["es"])
    print(f"case implicit {name}", flush=True)
    f = np.zeros((l0 + 1, 1, nx, np_), np.complex128)
    f[:] = g["f_in"]
    fld = np.zeros((6, 1, nx), np.complex128)
    fld[0] = g["fld_in"][0]
    ctx = ctx_from_golden(g, m0=0)
    ctx.upload(f, fld)
    ctx.step_implicit_1d(h, steps, es)
    fo, fldo = ctx.download()
    wf, we = _noise_floor(lambda: _oracle(l0, np_, nx, float(c["pmax"]), float(g["dx"]), float(c["density_np"]),
                                          float(c["small_dt"]), int(c["nb"])), f, fld, h, steps, es)
    ef, ee = rel(fo.real, g["f_out"]), rel(fldo[0].real, g["fld_out"][0])
    print(f"  f {ef:.2e} (1-ulp floor {wf:.2e}), Ex {ee:.2e} (floor {we:.2e})", flush=True)
    assert ef <= max(TOL, FLOOR_MULT * wf) and wf < 1e-10
    assert ee <= max(TOL, FLOOR_MULT * we) and we < 1e-6
    assert not np.any(fldo[1:3])
    assert ctx.time() == h * steps


def _state(l0, np_, nx):
    f, fld = _maxwellian_state(l0, 0, np_, nx, 1, KW["pmax"], seed=40 + l0)
    x = np.arange(nx) / nx
    fld[1, 0] = 3e-5 * np.cos(2 * np.pi * x)       # Ey, Ez: Faraday moves By, Bz in Adv1 (es = 0)
    fld[2, 0] = -2e-5 * np.sin(4 * np.pi * x)
    fld[4, 0] = 1e-5
    return f, fld


@pytest.mark.parametrize("es", [1, 0])
@pytest.mark.parametrize("l0,np_,nx,steps", [(8, 256, 64, 4), (3, 100, 37, 3), (12, 48, 20, 2)])
def test_implicit_1d_no_collisions_vs_oracle(l0, np_, nx, steps, es):
    """Collisions off: RK2 of Adv1/Implicit_Ex1D/Adv2, Faraday, the current and the field solve."""
    import oshun_amd
    print(f"case implicit no-collisions l0={l0} np={np_} nx={nx} es={es}", flush=True)
    f, fld = _state(l0, np_, nx)
    ctx = oshun_amd.Context(l0, 0, np_, nx, 1, collisions=0, **KW)
    ctx.upload(f, fld)
    ctx.step_implicit_1d(0.5, steps, es)
    fg, fldg = ctx.download()
    make = lambda: _oracle(l0, np_, nx, KW["pmax"], KW["dx"], KW["density_np"], KW["small_dt"],  # noqa: E731
                           KW["nb_algorithms"], coll=0)
    o = make()
    o.set_state(f, fld)
    o.step_implicit_1d(0.5, steps, es)
    fo, fldo = o.get_state()
    wf, we = _noise_floor(make, f, fld, 0.5, steps, es)
    ef, ee = rel(fg, fo), rel(fldg[0], fldo[0])
    print(f"  f {ef:.2e} (1-ulp floor {wf:.2e}), Ex {ee:.2e} (floor {we:.2e})", flush=True)
    assert np.all(np.isfinite(fg))
    assert ef <= max(TOL, FLOOR_MULT * wf) and wf < 1e-10
    assert ee <= max(TOL, FLOOR_MULT * we) and we < 1e-6
    assert not np.any(fldg[1:3])
    for c in range(3, 6):
        if np.abs(fldo[c]).max() > 0:
            assert rel(fldg[c], fldo[c]) < TOL, c
        else:
            assert not np.any(fldg[c]), c
    if es == 0:  # Faraday moved B
        assert np.abs(fldg[4] - fld[4]).max() > 0
    else:
        assert np.array_equal(fldg[3:], fld[3:])
    assert rel(fldg[0], fld[0]) > 1e-3  # the field solve replaced Ex


@pytest.mark.parametrize("l0,np_,nx,steps", [(8, 256, 64, 4), (12, 48, 20, 2)])
def test_implicit_1d_vs_oracle(l0, np_, nx, steps):
    """Collisions on (e-e, l = 1 probes, f_lm): within a multiple of the measured 1-ulp floor."""
    import oshun_amd
    print(f"case implicit vs oracle l0={l0} np={np_} nx={nx}", flush=True)
    f, fld = _state(l0, np_, nx)
    ctx = oshun_amd.Context(l0, 0, np_, nx, 1, **KW)
    ctx.upload(f, fld)
    ctx.step_implicit_1d(0.5, steps, 0)
    fg, fldg = ctx.download()
    make = lambda: _oracle(l0, np_, nx, KW["pmax"], KW["dx"], KW["density_np"], KW["small_dt"],  # noqa: E731
                           KW["nb_algorithms"])
    o = make()
    o.set_state(f, fld)
    o.step_implicit_1d(0.5, steps, 0)
    fo, fldo = o.get_state()
    wf, we = _noise_floor(make, f, fld, 0.5, steps, 0)
    ef, ee = rel(fg, fo), rel(fldg[0], fldo[0])
    print(f"  f {ef:.2e} (1-ulp floor {wf:.2e}), Ex {ee:.2e} (floor {we:.2e})", flush=True)
    assert ef <= max(TOL, FLOOR_MULT * wf) and wf < 1e-10
    assert ee <= max(TOL, FLOOR_MULT * we) and we < 1e-6
    for c in range(3, 6):
        assert rel(fldg[c], fldo[c]) < TOL if np.abs(fldo[c]).max() > 0 else not np.any(fldg[c])


def test_implicit_1d_rejects_other_runs():
    import oshun_amd
    print("case implicit rejects", flush=True)
    for l0, m0, prec in ((2, 0, 0), (4, 0, 1), (4, 1, 0)):
        ctx = oshun_amd.Context(l0, m0, 32, 8, 1, precision=prec)
        with pytest.raises(oshun_amd.OshunError):
            ctx.step_implicit_1d(0.5, 1)
