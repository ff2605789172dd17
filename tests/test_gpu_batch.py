"""Batched runs (oshun_batch_*): a heterogeneous batch advanced by the captured
per-step graph must give, run by run, exactly what the same run gives when it
is advanced alone through oshun_step (same kernels, same tiles: bitwise), and
within 1e-12 after the autotuner has changed the stage-pair tile height (the
current's partial sums then group p differently).
"""
import numpy as np
import pytest

from test_gpu_parity import _maxwellian_state

pytestmark = pytest.mark.gpu

# l0, Np, Nx, collisions (3: e-e + implicit, 1: e-e only, 0: none)
RUNS = [(1, 64, 32, 3), (2, 100, 77, 3), (3, 192, 130, 3), (6, 256, 64, 3), (4, 160, 50, 1), (2, 128, 40, 0),
        (8, 96, 200, 3)]
KW = dict(pmax=0.3, dx=0.4, density_np=1e22, small_dt=0.1, nb_algorithms=6)
H = 0.5


def _setup(precision):
    import oshun_amd
    runs, states = [], []
    for i, (l0, np_, nx, coll) in enumerate(RUNS):
        runs.append(dict(l0=l0, nump=np_, numx=nx, precision=precision, collisions=coll, **KW))
        states.append(_maxwellian_state(l0, 0, np_, nx, 1, KW["pmax"], seed=1000 + i))
    b = oshun_amd.Batch(runs, [H] * len(runs))
    for i, (f, fld) in enumerate(states):
        b.run(i).upload(f, fld)
    return b, runs, states


def _alone(r, f, fld, steps, rows=10):
    import oshun_amd
    kw = {k: v for k, v in r.items() if k not in ("l0", "nump", "numx")}
    ctx = oshun_amd.Context(r["l0"], 0, r["nump"], r["numx"], 1, **kw)
    ctx.pair_rows(rows)  # the batch's default tile height is 10 (a small run alone would take 6)
    ctx.upload(f, fld)
    ctx.step(H, steps)
    fo, fldo = ctx.download()
    return fo, fldo, ctx.time()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32split"])
def test_batch_equals_runs_alone(precision):
    print(f"case batch of {len(RUNS)} runs, precision {precision}", flush=True)
    b, runs, states = _setup(precision)
    b.step(3)
    for i, r in enumerate(runs):
        f, fld = b.run(i).download()
        fs, flds, t = _alone(r, *states[i], 3)
        assert np.all(np.isfinite(f))
        assert np.array_equal(f, fs), i
        assert np.array_equal(fld, flds), i
        assert b.run(i).time() == t
        assert np.abs(f - states[i][0]).max() > 0  # the run moved


def test_batch_autotune_restores_state_and_agrees():
    print("case batch autotune", flush=True)
    b, runs, states = _setup(0)
    before = [b.run(i).download() for i in range(len(runs))]
    rows, ms = b.autotune(2)
    print(f"autotune: {rows} p rows per tile, {ms:.3f} ms per batch step", flush=True)
    assert rows in (6, 10, -6, -8, -12) and ms > 0
    assert b.pair_rows() == rows
    for i in range(len(runs)):
        f, fld = b.run(i).download()
        assert np.array_equal(f, before[i][0]) and np.array_equal(fld, before[i][1])
        assert b.run(i).time() == 0.0
    b.step(3)
    for i, r in enumerate(runs):
        f, fld = b.run(i).download()
        fs, flds, _ = _alone(r, *states[i], 3)
        assert np.abs(f - fs).max() <= 1e-12 * np.abs(fs).max(), i
        assert np.abs(fld - flds).max() <= 1e-12 * max(np.abs(flds).max(), 1e-300), i


@pytest.mark.parametrize("precision,rows", [(0, -8), (1, -6), (0, 6)], ids=["fp64-march8", "fp32-march6", "fp64-tile6"])
def test_batch_equals_runs_alone_at_set_rows(precision, rows):
    print(f"case batch of {len(RUNS)} runs, precision {precision}, pair rows {rows}", flush=True)
    b, runs, states = _setup(precision)
    assert b.pair_rows(rows) == rows
    b.step(3)
    for i, r in enumerate(runs):
        f, fld = b.run(i).download()
        fs, flds, t = _alone(r, *states[i], 3, rows=rows)
        assert np.array_equal(f, fs), i
        assert np.array_equal(fld, flds), i
        assert b.run(i).time() == t


def test_batch_pair_rows_bounds():
    import oshun_amd
    print("case batch pair rows knob", flush=True)
    b, _, _ = _setup(0)
    for bad in (14, 7, -3, -33):
        with pytest.raises(oshun_amd.OshunError):
            b.pair_rows(bad)
    assert b.pair_rows() == 10
