"""The marched stage-pair kernels (one wavefront per x strip and p chunk).

Every wavefront walks pw rows of 60 columns; chunks at the ends of p take the
general path (rows 0 and np-1, CLEAN_LOW_PC rows, rows outside the grid),
the others the interior path.  The cases below put CLEAN_LOW_PC rows in
several chunks (l0 = 12 with 4-row chunks), leave the last chunk and the last
strip ragged, and run grids narrower than one strip (every lane wraps).

Bars: collisional RK4 steps through oshun_step against the oracle <= 1e-12
normwise in f and in every field (FP64); against the tile kernels on the same
run <= 1e-13 (the two sum the current's p partials in different groups, so
they agree to rounding, not bitwise); the FP32 split against the FP64 marched
run within 1e-5 in f (the FP32 slope's rounding over three steps).
"""
import numpy as np
import pytest

from test_gpu_parity import _maxwellian_state

pytestmark = pytest.mark.gpu

KW = dict(pmax=0.3, dx=0.4, density_np=1e22, small_dt=0.2, nb_algorithms=6)
H = 0.5


def _state(l0, np_, nx, seed):
    f, fld = _maxwellian_state(l0, 0, np_, nx, 1, KW["pmax"], seed=seed)
    x = np.arange(nx) / nx
    for c in range(1, 6):
        fld[c, 0] = 1e-4 * (c + 1) * np.cos(2 * np.pi * (c * x + 0.1 * c))
    return f, fld


def _run(l0, np_, nx, rows, precision, f, fld, nsteps):
    import oshun_amd
    ctx = oshun_amd.Context(l0, 0, np_, nx, 1, precision=precision, **KW)
    ctx.pair_rows(rows)
    assert ctx.pair_rows() == rows
    ctx.upload(f, fld)
    ctx.step(H, nsteps)
    out = ctx.download()
    ctx.close()
    return out


CASES = [(1, 64, 130, -4), (3, 100, 9, -5), (8, 256, 130, -9), (8, 70, 61, -4), (12, 96, 121, -4), (12, 50, 40, -9)]


@pytest.mark.parametrize("l0,np_,nx,rows", CASES)
def test_marched_steps_follow_the_oracle(l0, np_, nx, rows):
    from oracle.oracle import Config, Oracle
    print(f"case marched pairs vs oracle l0={l0} np={np_} nx={nx} rows={-rows}", flush=True)
    f, fld = _state(l0, np_, nx, seed=300 + l0)
    fg, fldg = _run(l0, np_, nx, rows, 0, f, fld, 3)
    o = Oracle(Config(l0=l0, m0=0, np=np_, nx=nx, pmax=KW["pmax"], dx=KW["dx"], density_np=KW["density_np"],
                      small_dt=KW["small_dt"], nb=KW["nb_algorithms"]))
    o.set_state(f, fld)
    o.step(H, 3)
    fo, fldo = o.get_state()
    assert np.abs(fg - fo).max() <= 1e-12 * np.abs(fo).max()
    for c in range(6):
        assert np.abs(fldg[c] - fldo[c]).max() <= 1e-12 * np.abs(fldo[c]).max(), c


@pytest.mark.parametrize("l0,np_,nx,rows", CASES)
def test_marched_equal_tile_kernels_to_rounding(l0, np_, nx, rows):
    print(f"case marched vs tile pairs l0={l0} np={np_} nx={nx} rows={-rows}", flush=True)
    f, fld = _state(l0, np_, nx, seed=400 + l0)
    fm, fldm = _run(l0, np_, nx, rows, 0, f, fld, 2)
    ft, fldt = _run(l0, np_, nx, 10, 0, f, fld, 2)
    assert np.abs(fm - ft).max() <= 1e-13 * np.abs(ft).max()
    for c in range(6):
        assert np.abs(fldm[c] - fldt[c]).max() <= 1e-13 * max(np.abs(fldt[c]).max(), 1e-300), c


@pytest.mark.parametrize("l0,np_,nx,rows", [(8, 256, 130, -6), (12, 96, 121, -4)])
def test_marched_fp32_split_tracks_fp64(l0, np_, nx, rows):
    print(f"case marched FP32 split l0={l0} np={np_} nx={nx} rows={-rows}", flush=True)
    f, fld = _state(l0, np_, nx, seed=500 + l0)
    f64, fld64 = _run(l0, np_, nx, rows, 0, f, fld, 3)
    f32, fld32 = _run(l0, np_, nx, rows, 1, f, fld, 3)
    assert np.all(np.isfinite(f32))
    assert np.abs(f32 - f64).max() <= 1e-5 * np.abs(f64).max()
    assert np.abs(fld32[0] - fld64[0]).max() <= 1e-5 * np.abs(fld64[0]).max()


def test_marched_rows_knob_bounds():
    import oshun_amd
    print("case marched rows knob", flush=True)
    ctx = oshun_amd.Context(3, 0, 64, 64, 1, **KW)
    for bad in (-3, -33, 7, 8):
        with pytest.raises(oshun_amd.OshunError):
            ctx.pair_rows(bad)
    assert ctx.pair_rows(-32) == -32
    ctx.close()
