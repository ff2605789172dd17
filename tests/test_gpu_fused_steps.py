"""Fused field stages of the 1D step: with both collision operators on, the
collision launch also does the step's last RK4 field stage and, inside one
oshun_step call, the next step's first field stage (from the collided f10).
That is a scheduling change only: n steps in one call must be bitwise equal to
n calls of one step (where every first field stage is its own launch), with
all six fields nonzero so that the vacuum-field halo recomputation is checked
too; and the result must still follow the oracle.
"""
import numpy as np
import pytest

from test_gpu_parity import _maxwellian_state

pytestmark = pytest.mark.gpu

KW = dict(pmax=0.3, dx=0.4, density_np=1e22, small_dt=0.1, nb_algorithms=6)
H = 0.5


def _state(l0, np_, nx, seed):
    f, fld = _maxwellian_state(l0, 0, np_, nx, 1, KW["pmax"], seed=seed)
    x = np.arange(nx) / nx
    for c in range(1, 6):  # vacuum fields: the x-neighbour terms of Ampere/Faraday
        fld[c, 0] = 1e-4 * (c + 1) * np.cos(2 * np.pi * (c * x + 0.1 * c))
    return f, fld


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32split"])
@pytest.mark.parametrize("l0,np_,nx", [(3, 192, 130), (8, 256, 64), (2, 100, 9)])
def test_steps_in_one_call_equal_single_steps(precision, l0, np_, nx):
    import oshun_amd
    print(f"case fused steps l0={l0} np={np_} nx={nx} precision={precision}", flush=True)
    f, fld = _state(l0, np_, nx, seed=77 + l0)
    a = oshun_amd.Context(l0, 0, np_, nx, 1, precision=precision, **KW)
    b = oshun_amd.Context(l0, 0, np_, nx, 1, precision=precision, **KW)
    a.upload(f, fld)
    b.upload(f, fld)
    a.step(H, 4)
    for _ in range(4):
        b.step(H, 1)
    fa, flda = a.download()
    fb, fldb = b.download()
    assert np.all(np.isfinite(fa))
    assert np.array_equal(fa, fb)
    assert np.array_equal(flda, fldb)
    assert a.time() == b.time()
    assert np.abs(flda - fld).max() > 0 and np.abs(fa - f).max() > 0  # the state moved


def test_fused_steps_follow_the_oracle():
    import oshun_amd
    from oracle.oracle import Config, Oracle
    l0, np_, nx = 3, 192, 48
    print(f"case fused steps vs oracle l0={l0} np={np_} nx={nx}", flush=True)
    f, fld = _state(l0, np_, nx, seed=5)
    ctx = oshun_amd.Context(l0, 0, np_, nx, 1, **KW)
    ctx.upload(f, fld)
    ctx.step(H, 3)
    fg, fldg = ctx.download()
    o = Oracle(Config(l0=l0, m0=0, np=np_, nx=nx, pmax=KW["pmax"], dx=KW["dx"], density_np=KW["density_np"],
                      small_dt=KW["small_dt"], nb=KW["nb_algorithms"]))
    o.set_state(f, fld)
    o.step(H, 3)
    fo, fldo = o.get_state()
    assert np.abs(fg - fo).max() <= 1e-12 * np.abs(fo).max()
    for c in range(6):
        assert np.abs(fldg[c] - fldo[c]).max() <= 1e-12 * np.abs(fldo[c]).max(), c


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32split"])
@pytest.mark.parametrize("l0,np_,nx", [(8, 256, 130), (3, 100, 9)])
def test_collisionless_steps_in_one_call_equal_single_steps(precision, l0, np_, nx):
    """Collisions off: one field launch does a step's last field stage and the next step's first."""
    import oshun_amd
    print(f"case fused steps, collisions off, l0={l0} np={np_} nx={nx} precision={precision}", flush=True)
    f, fld = _state(l0, np_, nx, seed=91 + l0)
    a = oshun_amd.Context(l0, 0, np_, nx, 1, precision=precision, collisions=0, **KW)
    b = oshun_amd.Context(l0, 0, np_, nx, 1, precision=precision, collisions=0, **KW)
    a.upload(f, fld)
    b.upload(f, fld)
    a.step(H, 4)
    for _ in range(4):
        b.step(H, 1)
    fa, flda = a.download()
    fb, fldb = b.download()
    assert np.all(np.isfinite(fa))
    assert np.array_equal(fa, fb)
    assert np.array_equal(flda, fldb)
    assert np.abs(flda - fld).max() > 0 and np.abs(fa - f).max() > 0


def test_collisionless_fused_steps_follow_the_oracle():
    import oshun_amd
    from oracle.oracle import Config, Oracle
    l0, np_, nx = 8, 128, 70
    print(f"case fused steps, collisions off, vs oracle l0={l0} np={np_} nx={nx}", flush=True)
    f, fld = _state(l0, np_, nx, seed=6)
    ctx = oshun_amd.Context(l0, 0, np_, nx, 1, collisions=0, **KW)
    ctx.upload(f, fld)
    ctx.step(H, 3)
    fg, fldg = ctx.download()
    o = Oracle(Config(l0=l0, m0=0, np=np_, nx=nx, pmax=KW["pmax"], dx=KW["dx"], density_np=KW["density_np"],
                      small_dt=KW["small_dt"], nb=KW["nb_algorithms"], coll_ee=0, coll_flm=0))
    o.set_state(f, fld)
    o.step(H, 3)
    fo, fldo = o.get_state()
    assert np.abs(fg - fo).max() <= 1e-12 * np.abs(fo).max()
    for c in range(6):
        assert np.abs(fldg[c] - fldo[c]).max() <= 1e-12 * np.abs(fldo[c]).max(), c
