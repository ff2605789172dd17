"""Parity at the configurations' full sizes, against the oracle.

C4  the 512-run ensemble of bench.ensemble_specs (seed 1234: Nx U{32..1024},
    Np U{64..256}, l0 U{1..6}, T0 log-U[0.1, 5] keV, cosine amplitude
    U[0.01, 0.2]), autotuned, stepped 3 times by the captured per-step graph:
    sampled runs (smallest, median and largest Nx*Np, the first run of every l0,
    the last run) against
    Oracle.step run by run -- FP64 at <= 1e-12 normwise, and the configuration's
    own FP32-split form within the FP32 bound (relative L2 <= 1e-5 in f00 and
    f10, <= 1e-4 of max|Ex| in Ex).
T   the target shape (l0 = 8, Np = 256, Nx = 4096, collisions on): 2 FP64
    steps against the oracle at <= 1e-12.
"""
import os
import sys

import numpy as np
import pytest

from oracle.oracle import Config, Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def l2rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _sample(specs):
    size = [nx * np_ for nx, np_, _, _, _ in specs]
    order = np.argsort(size)
    pick = {int(order[0]), int(order[-1]), int(order[len(order) // 2]), len(specs) - 1}
    for l0 in range(1, 7):
        pick.add(next(i for i, s in enumerate(specs) if s[2] == l0))
    return sorted(pick)


def _oracle_run(run, state, h, steps):
    o = Oracle(Config(l0=run["l0"], m0=0, np=run["nump"], nx=run["numx"], pmax=run["pmax"], dx=run["dx"],
                      density_np=run["density_np"], small_dt=run["small_dt"], nb=run["nb_algorithms"]))
    o.set_state(*state)
    o.step(h, steps)
    return o.get_state()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32split"])
def test_c4_ensemble_sampled_runs_vs_oracle(precision):
    import bench
    import oshun_amd
    specs = bench.ensemble_specs(512)
    print(f"case C4 ensemble: 512 runs, precision {precision}, autotune + 3 steps", flush=True)
    runs = []
    for nx, np_, l0, t0, amp in specs:
        pmax = max(0.05, 6.0 * np.sqrt(t0 / 511.0))
        runs.append(dict(l0=l0, nump=np_, numx=nx, pmax=pmax, dx=0.5, precision=precision,
                         density_np=bench.DENSITY, small_dt=bench.H, nb_algorithms=bench.NB))
    b = oshun_amd.Batch(runs, [bench.H] * len(runs))
    states = {}
    sample = _sample(specs)
    for i, (nx, np_, l0, t0, amp) in enumerate(specs):
        st = bench.ensemble_state(nx, np_, l0, t0, amp, runs[i]["pmax"], seed=i)
        b.run(i).upload(*st)
        if i in sample:
            states[i] = st
    rows, ms = b.autotune(2)
    print(f"autotuned {rows} p rows per tile, {ms:.3f} ms per batch step", flush=True)
    b.step(3)
    b.synchronize()
    for i in sample:
        f, fld = b.run(i).download()
        fo, fldo = _oracle_run(runs[i], states[i], bench.H, 3)
        nx, np_, l0 = specs[i][:3]
        if precision == 0:
            e_f, e_x = rel(f, fo), rel(fld[0].real, fldo[0].real)
            print(f"run {i} (Nx={nx} Np={np_} l0={l0}): f {e_f:.1e} Ex {e_x:.1e}", flush=True)
            assert e_f <= 1e-12 and e_x <= 1e-12, i
        else:
            e = {"f00": l2rel(f[0], fo[0]), "f10": l2rel(f[1], fo[1]),
                 "Ex": float(np.abs(fld[0] - fldo[0]).max() / max(np.abs(fldo[0]).max(), 1e-300))}
            print(f"run {i} (Nx={nx} Np={np_} l0={l0}) FP32 split:", {k: f"{v:.1e}" for k, v in e.items()},
                  flush=True)
            assert e["f00"] <= 1e-5 and e["f10"] <= 1e-5 and e["Ex"] <= 1e-4, i


def test_target_shape_vs_oracle():
    import bench
    import oshun_amd
    print("case target shape l0=8 Np=256 Nx=4096, 2 FP64 steps", flush=True)
    f, fld, dx = bench.target_state()
    ctx = bench.make_ctx(oshun_amd.FP64, dx)
    ctx.upload(f, fld)
    ctx.step(bench.H, 2)
    fg, fldg = ctx.download()
    o = Oracle(Config(l0=bench.L0, m0=0, np=bench.NP, nx=bench.NX, pmax=bench.PMAX, dx=dx,
                      density_np=bench.DENSITY, small_dt=bench.H, nb=bench.NB))
    o.set_state(f, fld)
    o.step(bench.H, 2)
    fo, fldo = o.get_state()
    print(f"target: f {rel(fg, fo):.1e} Ex {rel(fldg[0].real, fldo[0].real):.1e}", flush=True)
    assert rel(fg, fo) <= 1e-12
    assert rel(fldg[0].real, fldo[0].real) <= 1e-12
