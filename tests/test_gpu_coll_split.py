"""Implicit l >= 1 collisions at Np > 256 with the planes of each column group
split among several workgroups (launches without the e-e step: oshun_collide_flm,
and the implicit-E steps' f_lm collisions).

Every plane's solve is the same in any workgroup, so the FP64 result is held
to the oracle's Implicit_FLM_step at <= 1e-12 normwise, as the other collision
parity tests are; the shapes give one to several l per workgroup, ragged
splits (plane counts not a multiple of the split) and m0 = 2 complex planes.
"""
import numpy as np
import pytest

from oracle.oracle import Config, Oracle, hind

pytestmark = pytest.mark.gpu

TOL = 1e-12
KW = dict(pmax=0.3, dx=0.4, dy=0.5, density_np=1e22, small_dt=0.1, nb=6)


def _state(l0, m0, np_, nx, ny, seed):
    rng = np.random.default_rng(seed)
    from oracle.oracle import num_harmonics
    pmin = KW["pmax"] / (2.0 * np_)
    p = np.arange(np_) * ((KW["pmax"] - pmin) / (np_ - 1)) + pmin
    x = np.arange(nx) / nx
    T = (1.0 / 511.0) * (1.0 + 0.1 * np.cos(2 * np.pi * x))
    f = np.zeros((num_harmonics(l0, m0), ny, nx, np_), np.complex128)
    f[0] = (np.exp(-p[None, :] ** 2 / (2 * T[:, None])) / (2 * np.pi * T[:, None]) ** 1.5)[None, :, :]
    for l in range(1, l0 + 1):
        for m in range(0, min(l, m0) + 1):
            r = rng.uniform(-1, 1, (ny, nx, np_))
            if m > 0:
                r = r + 1j * rng.uniform(-1, 1, (ny, nx, np_))
            f[hind(l, m, m0)] = 1e-3 * r * f[0].real
    return f, np.zeros((6, ny, nx), np.complex128)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("l0,m0,np_,nx,ny", [(16, 2, 300, 24, 1), (9, 2, 288, 40, 2), (20, 0, 320, 64, 1)])
def test_flm_collisions_with_split_planes_follow_the_oracle(l0, m0, np_, nx, ny):
    import oshun_amd
    print(f"case split-plane flm collisions l0={l0} m0={m0} np={np_} nx={nx} ny={ny}", flush=True)
    f, fld = _state(l0, m0, np_, nx, ny, seed=7 * l0 + np_)
    o = Oracle(Config(l0=l0, m0=m0, np=np_, nx=nx, ny=ny, **KW))
    o.set_state(f, fld)
    ctx = oshun_amd.Context(l0, m0, np_, nx, ny, pmax=KW["pmax"], dx=KW["dx"], dy=KW["dy"],
                            density_np=KW["density_np"], small_dt=KW["small_dt"], nb_algorithms=KW["nb"])
    ctx.upload(f, fld)
    t = 0.0
    for _ in range(2):
        t += 0.5
        for which in (1, 2):
            o.collide_flm(t, which)
            ctx.collide_flm(t, which)
    fo, _ = o.get_state()
    fg, _ = ctx.download()
    assert np.all(np.isfinite(fg))
    assert rel(fg, fo) < TOL
    assert rel(fo, f) > 1e-9  # the collisions moved the state
    ctx.close()
