"""Benchmark: phase-space cell-updates/s per RK4 + Fokker-Planck step on one MI355X.

Workload (the target shape of the distribution-function update): 1D periodic,
l0 = 8, m0 = 0 (9 real harmonics), Np = 256 to pmax = 0.3 m_e c, Nx = 4096,
kλ_D = 0.3 Langmuir perturbation over 16 wavelengths, Maxwellian T = 1 keV,
collisions on (e-e f00 + implicit l >= 1, n = 1e21 cm^-3, Z = 1), h = 0.01/ω_p,
e-e sub-step = h.  A step = RK4 (4 Vlasov-Maxwell stages) + e-e + f1 + f_lm,
i.e. main.cpp:79-107 of the reference.  Synthetic initial data, resident in HBM
before timing.

value      = N_h*Np*Nx / t_step for the plain FP64 path (dtype f64)
fp32_split = the same for FP32 f storage + FP32 Vlasov RHS with FP64 fields and
             collisions
roofline   = dominant kernel (vlasov_pair_1d, two RK4 stages per launch):
             algorithmic bytes per launch (16 S_f per step shared by the
             step's Vlasov launches, S_f = stored f bytes, i.e. 8 S_f per
             pair launch) / mean launch duration from HIP events on the
             library's stream; traffic = mean HBM bytes per pair launch
             measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
cpu_baseline = the reference (oracle/_ref, built from the reference sources)
             on a bounded sample of the same workload, 1 core, or the oracle's
             CPU restatement when that driver is absent
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E

L0, M0, NP, NX = 8, 0, 256, 4096
PMAX = 0.3
H = 0.01
DENSITY = 1e21
NB = 6


def target_state(nx=NX, l0=L0, np_=NP, seed=1234):
    pt = np.sqrt(1.0 / 511.0)                 # T = 1 keV
    lam_d = pt                                # c/ω_p units
    k = 0.3 / lam_d
    L = 16 * 2 * np.pi / k
    dx = L / nx
    pmin = PMAX / (2.0 * np_)
    p = np.arange(np_) * ((PMAX - pmin) / (np_ - 1)) + pmin
    x = np.arange(nx) * dx
    n = 1.0 + 1e-3 * np.sin(k * x)
    f = np.zeros((l0 + 1, 1, nx, np_), np.complex128)
    f00 = n[:, None] * np.exp(-p[None, :] ** 2 / (2 * pt ** 2)) / (np.sqrt(2 * np.pi) * 2 * np.pi * pt ** 3)
    f[0, 0] = f00
    rng = np.random.default_rng(seed)
    for l in range(1, l0 + 1):
        f[l, 0] = 1e-6 * rng.uniform(-1, 1, (nx, np_)) * f00
    fld = np.zeros((6, 1, nx), np.complex128)
    return f, fld, dx


def make_ctx(precision, dx):
    import oshun_amd
    return oshun_amd.Context(L0, M0, NP, NX, 1, pmax=PMAX, dx=dx, precision=precision, density_np=DENSITY,
                             zeta=1.0, small_dt=H, nb_algorithms=NB)


def conservation(m0, m1):
    """relative change of the total electron number and kinetic energy between two moment sets"""
    n0, n1 = m0[0].sum(), m1[0].sum()
    e0 = (m0[0] * (m0[3] + m0[4] + m0[5])).sum()
    e1 = (m1[0] * (m1[3] + m1[4] + m1[5])).sum()
    return {"dN_over_N": float((n1 - n0) / n0), "dE_over_E": float((e1 - e0) / e0)}


def run_path(precision, steps, warmup, f, fld, dx):
    ctx = make_ctx(precision, dx)
    ctx.upload(f, fld)
    ctx.step(H, warmup)
    ctx.synchronize()
    m_start = ctx.moments()
    t0 = time.perf_counter()
    ctx.step(H, steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    ctx.cons = conservation(m_start, ctx.moments())
    ctx.final = ctx.download()  # after warmup + steps, for the FP32-split error against FP64
    # per-kernel durations: HIP events on the library's stream over `steps` more steps
    ctx.profile(True)
    ctx.step(H, steps)
    prof = ctx.profile_read()
    ctx.profile(False)
    fo, _ = ctx.download()
    if not np.all(np.isfinite(fo.real)):
        raise RuntimeError("non-finite state after the timed steps")
    return ctx, ms, prof


def host_cores():
    """CPU threads this process may use (the GPU host's share: at most 16)"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _ref_wall(procs, nsteps, inp, td):
    """wall time of `procs` concurrent reference drivers running `nsteps` steps of the same slab"""
    import subprocess
    from oracle import oracle as orc
    with open(inp, "rb") as fh:
        raw = bytearray(fh.read())
    path = os.path.join(td, f"in_{nsteps}.bin")
    hdr = np.frombuffer(raw[:52], np.int32).copy()
    hdr[6] = nsteps
    raw[:52] = hdr.tobytes()
    with open(path, "wb") as fh:
        fh.write(raw)
    t0 = time.perf_counter()
    ps = [subprocess.Popen([orc.REF_BIN, path, os.path.join(td, f"out_{nsteps}_{k}.bin")],
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL) for k in range(procs)]
    rcs = [p.wait() for p in ps]
    if any(rcs):
        raise RuntimeError(f"reference driver failed: {rcs}")
    return time.perf_counter() - t0


def cpu_baseline(budget_s=20.0):
    """The reference on a bounded sample of the workload (same l0, Np, h, collisions) on the host's cores:
    P concurrent reference drivers (oracle/_ref, the unmodified reference classes), one periodic x slab of
    Nx = 64 each, as the reference's MPI ranks would hold slabs (its only parallelism); the halo exchange
    between ranks is not included."""
    import tempfile
    from oracle import oracle as orc
    nx = 64
    f, fld, dx = target_state(nx=nx)
    xmin, xmax = 0.0, dx * (nx - 1)
    if orc.have_ref():
        # the reference needs m0 >= 2: run the m0 = 0 physics as a y-uniform m0 = 2 stand-in
        nh2 = orc.num_harmonics(L0, 2)
        f2 = np.zeros((nh2, 1, nx, NP), np.complex128)
        for l in range(L0 + 1):
            f2[orc.hind(l, 0, 2)] = f[l]
        P = host_cores()
        with tempfile.TemporaryDirectory() as td:
            inp = os.path.join(td, "in.bin")
            orc.write_ref_input(inp, L0, 2, NP, nx, 1, f2, fld, op=1, nsteps=1, nb=NB, pmax=PMAX, xmin=xmin,
                                xmax=xmax, h=H, small_dt=H, density_np=DENSITY)
            setup = _ref_wall(P, 1, inp, td)
            nsteps = max(1, min(6, int(budget_s / max(setup, 1e-3))))
            wall = _ref_wall(P, nsteps + 1, inp, td)
        per_step = max(wall - setup, 1e-9) / nsteps  # wall time of one step of all P slabs together
        cells = (L0 + 1) * NP * nx * P
        return {"value": cells / per_step, "unit": "cell-updates/s", "cores": P, "kind": "reference",
                "sample": f"OSHUN-1 classes (oracle/_ref) as {P} concurrent processes, one per core, each a "
                          f"periodic slab of Nx=64 (+4 guards each side): l0=8 as m0=2 y-uniform stand-in (24 "
                          f"harmonics stored, 9 useful), Np=256, {nsteps} RK4+FP steps minus a 1-step run; value "
                          f"counts the 9 useful harmonics of all {P} slabs",
                "s_per_step": per_step, "slab_cells": (L0 + 1) * NP * nx}
    # fall back to the oracle's restatement
    o = orc.Oracle(orc.Config(l0=L0, m0=M0, np=NP, nx=nx, pmax=PMAX, dx=dx, density_np=DENSITY, small_dt=H,
                              nb=NB))
    o.set_state(f, fld)
    t0 = time.perf_counter()
    o.step(H, 1)
    dt1 = time.perf_counter() - t0
    n = max(1, min(50, int(budget_s / max(dt1, 1e-3))))
    t0 = time.perf_counter()
    o.step(H, n)
    per_step = (time.perf_counter() - t0) / n
    cells = (L0 + 1) * NP * nx
    return {"value": cells / per_step, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle CPU restatement, l0=8 m0=0 Np=256 Nx=64, {n} steps", "s_per_step": per_step}


def ensemble_specs(nruns=512, seed=1234):
    """The batched-ensemble workload: per run, drawn in this order from
    numpy.random.default_rng(seed): Nx ~ U{32..1024}, Np ~ U{64..256},
    l0 ~ U{1..6}, T0 ~ log-uniform [0.1, 5] keV, perturbation amplitude
    ~ U[0.01, 0.2].  1D periodic, m0 = 0, FP32 Vlasov RHS with FP64 fields and
    collisions (e-e + implicit l >= 1 at n = 1e21 cm^-3), h = e-e sub-step =
    0.01/omega_p, pmax = 6 p_t (at least 0.05), dx = 0.5 c/omega_p, initial
    state a Maxwellian with T(x) = T0 (1 + a cos(2 pi x / L))."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(nruns):
        nx = int(rng.integers(32, 1025))
        np_ = int(rng.integers(64, 257))
        l0 = int(rng.integers(1, 7))
        t0 = float(np.exp(rng.uniform(np.log(0.1), np.log(5.0))))
        amp = float(rng.uniform(0.01, 0.2))
        specs.append((nx, np_, l0, t0, amp))
    return specs


def ensemble_state(nx, np_, l0, t0, amp, pmax, seed):
    pmin = pmax / (2.0 * np_)
    p = np.arange(np_) * ((pmax - pmin) / (np_ - 1)) + pmin
    T = (t0 / 511.0) * (1.0 + amp * np.cos(2 * np.pi * np.arange(nx) / nx))
    f = np.zeros((l0 + 1, 1, nx, np_), np.complex128)
    f[0, 0] = np.exp(-p[None, :] ** 2 / (2 * T[:, None])) / (np.sqrt(2 * np.pi) * 2 * np.pi * T[:, None] ** 1.5)
    rng = np.random.default_rng(seed)
    for l in range(1, l0 + 1):
        f[l, 0] = 1e-6 * rng.uniform(-1, 1, (nx, np_)) * f[0, 0].real
    return f, np.zeros((6, 1, nx), np.complex128)


def run_ensemble(steps, warmup, nruns=512):
    import oshun_amd
    specs = ensemble_specs(nruns)
    runs, hs = [], []
    for nx, np_, l0, t0, amp in specs:
        pmax = max(0.05, 6.0 * np.sqrt(t0 / 511.0))
        runs.append(dict(l0=l0, nump=np_, numx=nx, pmax=pmax, dx=0.5, precision=oshun_amd.FP32_SPLIT,
                         density_np=DENSITY, small_dt=H, nb_algorithms=NB))
        hs.append(H)
    b = oshun_amd.Batch(runs, hs)
    for i, (nx, np_, l0, t0, amp) in enumerate(specs):
        b.run(i).upload(*ensemble_state(nx, np_, l0, t0, amp, runs[i]["pmax"], seed=i))
    rows, tuned_ms = b.autotune(3)
    b.step(warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    dt = time.perf_counter() - t0
    for i in (0, len(specs) // 2, len(specs) - 1):
        f, _ = b.run(i).download()
        if not np.all(np.isfinite(f.real)):
            raise RuntimeError(f"non-finite state in run {i}")
    cells = sum((l0 + 1) * np_ * nx for nx, np_, l0, _, _ in specs)
    return {"metric": "runs*steps/s (batched ensemble, one hipGraph launch per step)",
            "value": nruns * steps / dt, "unit": "run-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": dt * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 f / f64 fields+collisions", "data": "synthetic (seeded draws, see config)",
            "config": {"workload": f"{nruns} independent 1D runs, Nx U[32,1024], Np U[64,256], l0 U{{1..6}}, "
                                   "T0 log-U[0.1,5] keV, cosine T perturbation U[0.01,0.2], seed 1234, "
                                   "collisions on, h=0.01",
                       "cells": cells, "autotuned_p_rows_per_tile": rows, "autotune_ms_per_step": tuned_ms},
            "cell_updates_per_s": cells * steps / dt}


def _config_cases():
    """The reference's synthetic configurations as throughput cases (shapes, collisions and step
    sizes of each; states are their initial conditions): name, (l0, m0, Np, Nx, Ny), Context
    keywords, h, f, fld."""
    from oshun_amd import hind, num_harmonics

    def nu_ei_over_wp(t_kev, n=1e21, z=1.0, lnl=10.0):  # NRL formulary
        return 2.91e-6 * z * n * lnl * (t_kev * 1e3) ** -1.5 / (5.64e4 * np.sqrt(n))

    def lam_ei(t_kev):
        return np.sqrt(t_kev / 511.0) / nu_ei_over_wp(t_kev)

    def f00(np_, pmax, t_kev, n=1.0):
        pmin = pmax / (2.0 * np_)
        p = np.arange(np_) * ((pmax - pmin) / (np_ - 1)) + pmin
        pt2 = np.asarray(t_kev)[..., None] / 511.0
        return np.asarray(n)[..., None] * np.exp(-p ** 2 / (2 * pt2)) / (np.sqrt(2 * np.pi) * 2 * np.pi * pt2 ** 1.5)

    cases = []
    # C1: heat-flow relaxation, Nx = 64 on L = 100 lambda_ei, Np = 64, l0 = 2, collisions (Z = 1), h = 0.1
    nx, np_, l0 = 64, 64, 2
    x = np.arange(nx) / nx
    f = np.zeros((l0 + 1, 1, nx, np_), np.complex128)
    f[0, 0] = f00(np_, 0.25, 1.0 + 0.1 * np.cos(2 * np.pi * x))
    cases.append(("C1 heat flow, l0=2 Np=64 Nx=64, collisions", (l0, 0, np_, nx, 1),
                  dict(pmax=0.25, dx=100 * lam_ei(1.0) / nx, density_np=1e21, small_dt=0.1, nb_algorithms=4),
                  0.1, f, np.zeros((6, 1, nx), np.complex128)))
    # C2: collisionless Langmuir wave, Nx = 256, Np = 256, l0 = 8, k lambda_D = 0.3, h = 0.01
    nx, np_, l0 = 256, 256, 8
    k = 0.3 / np.sqrt(1.0 / 511.0)
    xs = np.arange(nx) * (2 * np.pi / k / nx)
    f = np.zeros((l0 + 1, 1, nx, np_), np.complex128)
    f[0, 0] = f00(np_, 0.3, np.ones(nx), 1.0 + 1e-3 * np.sin(k * xs))
    cases.append(("C2 Langmuir, l0=8 Np=256 Nx=256, collisions off", (l0, 0, np_, nx, 1),
                  dict(pmax=0.3, dx=2 * np.pi / k / nx, density_np=1e21, small_dt=0.01, nb_algorithms=6,
                       collisions=0), 0.01, f, np.zeros((6, 1, nx), np.complex128)))
    # C3: 2D magnetised hot spot, 256 x 256, Np = 128, l0 = m0 = 4, e-e + e-i, uniform Bz
    nx = ny = 256
    np_, l0, m0 = 128, 4, 4
    lam = lam_ei(0.5)
    d = 20.0 * lam / 16.0
    xx = (np.arange(nx) - nx / 2) * d
    r2 = xx[None, :] ** 2 + xx[:, None] ** 2
    f = np.zeros((num_harmonics(l0, m0), ny, nx, np_), np.complex128)
    f[hind(0, 0, m0)] = f00(np_, 0.25, 0.5 + 1.5 * np.exp(-r2 / (2 * (20.0 * lam) ** 2)))
    fld = np.zeros((6, ny, nx), np.complex128)
    fld[5] = 0.5 * nu_ei_over_wp(0.5)
    cases.append(("C3 2D hot spot, l0=m0=4 Np=128 256x256, collisions, Bz", (l0, m0, np_, nx, ny),
                  dict(pmax=0.25, dx=d, dy=d, density_np=1e21, small_dt=0.5, nb_algorithms=4), 0.5, f, fld))
    # C5: adjoint configuration's forward step, Nx = 512, Np = 192, l0 = 3, h = 0.1
    nx, np_, l0 = 512, 192, 3
    x = np.arange(nx) / nx
    rng = np.random.default_rng(1234)
    T = 1.0 + sum(0.05 / 8 * rng.uniform(-1, 1) * np.cos(2 * np.pi * (m + 1) * x + rng.uniform(0, 2 * np.pi))
                  for m in range(8))
    f = np.zeros((l0 + 1, 1, nx, np_), np.complex128)
    f[0, 0] = f00(np_, 0.25, T)
    cases.append(("C5 heat flow (adjoint config forward), l0=3 Np=192 Nx=512, collisions", (l0, 0, np_, nx, 1),
                  dict(pmax=0.25, dx=100 * lam_ei(1.0) / nx, density_np=1e21, small_dt=0.1, nb_algorithms=4),
                  0.1, f, np.zeros((6, 1, nx), np.complex128)))
    return cases


def _config_cpu(name, shape, kw, h, f, fld, budget_s=6.0):
    """The reference classes (oracle/_ref, one core) on a bounded sample of one configuration: the same
    l0, m0, Np, h and collisions on fewer x (and y) cells; m0 < 2 runs as a y-uniform m0 = 2 stand-in
    (the reference needs l0 >= 3, m0 >= 2), so C1 (l0 = 2) is timed on the oracle's restatement instead.
    Returns the rate in useful cell-updates/s (harmonics of the configuration, not of the stand-in)."""
    from oracle import oracle as orc
    l0, m0, np_, nx, ny = shape
    sx, sy = min(nx, 16 if ny == 1 else 8), min(ny, 8)
    fs = np.ascontiguousarray(f[:, :sy, :sx, :])
    flds = np.ascontiguousarray(fld[:, :sy, :sx])
    coll = kw.get("collisions", 3)
    cells = orc.num_harmonics(l0, m0) * np_ * sx * sy
    if l0 >= 3 and orc.have_ref():
        m2 = max(m0, 2)
        f2 = np.zeros((orc.num_harmonics(l0, m2), sy, sx, np_), np.complex128)
        for l in range(l0 + 1):
            for m in range(min(l, m0) + 1):
                f2[orc.hind(l, m, m2)] = fs[orc.hind(l, m, m0)]
        dx, dy = kw["dx"], kw.get("dy", 1.0)
        rk = dict(nb=kw["nb_algorithms"], pmax=kw["pmax"], xmin=0.0, xmax=dx * (sx - 1), ymin=0.0,
                  ymax=dy * max(sy - 1, 1), h=h, small_dt=kw["small_dt"], density_np=kw["density_np"])
        op = 1 if coll else 2
        t0 = time.perf_counter()
        orc.run_ref(l0, m2, np_, sx, sy, f2, flds, op=op, nsteps=1, **rk)
        setup = time.perf_counter() - t0
        n = max(1, min(8, int(budget_s / max(setup, 1e-3))))
        t0 = time.perf_counter()
        orc.run_ref(l0, m2, np_, sx, sy, f2, flds, op=op, nsteps=n + 1, **rk)
        per_step = max(time.perf_counter() - t0 - setup, 1e-9) / n
        stand = " as an m0 = 2 stand-in" if m0 < 2 else ""
        return {"value": cells / per_step, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                "sample": f"OSHUN-1 classes (oracle/_ref){stand}, Nx={sx} Ny={sy}, {n} steps minus a 1-step run"}
    # the restatement is fast enough for the whole (small) configuration
    cells = orc.num_harmonics(l0, m0) * np_ * nx * ny
    o = orc.Oracle(orc.Config(l0=l0, m0=m0, np=np_, nx=nx, ny=ny, pmax=kw["pmax"], dx=kw["dx"],
                              density_np=kw["density_np"], small_dt=kw["small_dt"], nb=kw["nb_algorithms"],
                              coll_ee=1 if coll else 0, coll_flm=1 if coll else 0))
    o.set_state(f, fld)
    t0 = time.perf_counter()
    o.step(h, 1)
    d1 = time.perf_counter() - t0
    n = max(1, min(2000, int(budget_s / max(d1, 1e-4))))
    t0 = time.perf_counter()
    o.step(h, n)
    per_step = (time.perf_counter() - t0) / n
    return {"value": cells / per_step, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle CPU restatement (the reference needs l0 >= 3), Nx={nx} Ny={ny}, {n} steps"}


def run_configs(steps, warmup, only=None, prec_only=None, no_cpu=False):
    """cell-updates/s of the RK4 + FP step on each synthetic configuration, FP64 and FP32 split, with the
    step's algorithmic HBM bytes (16 S_f) over its time as the roofline fraction.  only: a configuration
    name prefix ("C3"); prec_only: "fp64" or "fp32"."""
    import oshun_amd
    out = []
    for name, (l0, m0, np_, nx, ny), kw, h, f, fld in _config_cases():
        if only and not name.startswith(only):
            continue
        cells = oshun_amd.num_harmonics(l0, m0) * np_ * nx * ny
        row = {"config": name, "cells": cells}
        for prec, tag in ((oshun_amd.FP64, "fp64"), (oshun_amd.FP32_SPLIT, "fp32_split")):
            if prec_only and not tag.startswith(prec_only):
                continue
            ctx = oshun_amd.Context(l0, m0, np_, nx, ny, precision=prec, **kw)
            ctx.upload(f, fld)
            ctx.step(h, warmup)
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.step(h, steps)
            ctx.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / steps
            fo, _ = ctx.download()
            if not np.all(np.isfinite(fo.real)):
                raise RuntimeError(f"non-finite state in {name} ({tag})")
            S = ctx.state_bytes()
            row[tag] = {"ms_per_step": ms, "cell_updates_per_s": cells / (ms * 1e-3),
                        "step_roofline_frac": 16.0 * S / (ms * 1e-3) / HBM_PEAK,
                        "stage_kernel": ctx.stage_kernel()}
            ctx.close()
        if not no_cpu:
            row["cpu_baseline"] = _config_cpu(name, (l0, m0, np_, nx, ny), kw, h, f, fld)
        out.append(row)
    return {"metric": "phase-space cell-updates/s per RK4+FP step on the synthetic configurations",
            "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "data": "synthetic",
            "configs": out}


def run_implicit(steps, warmup, no_cpu=False):
    """The implicit-E step of the 1D variant (oshun_step_implicit_1d: Runge_Kutta_2_Imp +
    Implicit_E_Field with if_implicit1D, e-e, implicit f_lm) on the single-run workload, FP64,
    electrostatic (if_implicitES), with the reference's op-5 driver timed on a sample."""
    import oshun_amd
    f, fld, dx = target_state()
    ctx = make_ctx(oshun_amd.FP64, dx)
    ctx.upload(f, fld)
    ctx.step_implicit_1d(H, warmup, 1)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step_implicit_1d(H, steps, 1)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ctx.profile(True)
    ctx.step_implicit_1d(H, steps, 1)
    prof = ctx.profile_read()
    ctx.profile(False)
    fo, fldo = ctx.download()
    if not (np.all(np.isfinite(fo.real)) and np.all(np.isfinite(fldo.real))):
        raise RuntimeError("non-finite state after the timed implicit steps")
    cells = (L0 + 1) * NP * NX
    cpu = None
    if not no_cpu:
        from oracle import oracle as orc
        if orc.have_ref():
            nx = 64
            fs, flds, dxs = target_state(nx=nx)
            nh2 = orc.num_harmonics(L0, 2)
            f2 = np.zeros((nh2, 1, nx, NP), np.complex128)
            for l in range(L0 + 1):
                f2[orc.hind(l, 0, 2)] = fs[l]
            kw = dict(nb=NB, pmax=PMAX, xmin=0.0, xmax=dxs * (nx - 1), h=H, small_dt=H, density_np=DENSITY, es=1)
            t0 = time.perf_counter()
            orc.run_ref(L0, 2, NP, nx, 1, f2, flds, op=5, nsteps=1, **kw)
            setup = time.perf_counter() - t0
            n = 4
            t0 = time.perf_counter()
            orc.run_ref(L0, 2, NP, nx, 1, f2, flds, op=5, nsteps=n + 1, **kw)
            per_step = max(time.perf_counter() - t0 - setup, 1e-9) / n
            cpu = {"value": (L0 + 1) * NP * nx / per_step, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                   "sample": f"OSHUN-1 classes (oracle/_ref, op 5), l0=8 as m0=2 y-uniform stand-in, Np=256, Nx=64, "
                             f"{n} implicit-E steps minus a 1-step run; value counts the 9 useful harmonics",
                   "s_per_step": per_step}
    return {"metric": "phase-space cell-updates/s per implicit-E RK2+FP step (1D variant, FP64)",
            "value": cells / (ms * 1e-3), "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (the single-run workload's initial state)",
            "config": {"workload": "1D periodic, l0=8 m0=0, Np=256, Nx=4096, h=0.01, implicit_E with if_implicit1D "
                                   "and if_implicitES, e-e+e-i collisions on", "cells": cells},
            "kernels": {k: {"ms_per_step": v[0] / steps, "launches_per_step": v[1] / steps}
                        for k, v in prof.items() if v[1]},
            "cpu_baseline": cpu}


# the three production decks (def-Input.cpp:60-200 of POP-marocchino-secIII.B, POP-tzoufras-secIII.A,
# POP-tzoufras-secIV.B): implicit_E without if_implicit1D, RKLevel 2, their own grids, boundaries,
# collision and laser parameters and time step (h = dt_out / num_h with CLF = clf_dp * dp, def-RK.cpp:
# 407-413); the decks spread these grids over 32, 64 and 128 MPI ranks, here they run on one GPU
_LASER = dict(lambda_0=0.351, I_0=1.0e16, ab_eta=1.0, cntr_r0=0.0, smaller_dt=0.005)
POP_DECKS = (
    ("marocchino-secIII.B", dict(l0=48, m0=2, np=288, nx=640, ny=2, pmax=0.4, xmin=-8574.383315, xmax=8574.383315,
                                 ymin=-5000.0, ymax=5000.0, es=1, bndx=0, density_np=1.5e21, zeta=1.0,
                                 small_dt=0.16, nb=4, t_stop=72801.67589, n_out=300, src=None)),
    ("tzoufras-secIII.A", dict(l0=31, m0=2, np=432, nx=192, ny=2, pmax=0.6, xmin=-39880.82246, xmax=39880.82246,
                               ymin=-1000.0, ymax=1000.0, es=1, bndx=0, density_np=4.5e20, zeta=2.0, small_dt=0.15,
                               nb=6, t_stop=358927.4021, n_out=300,
                               src=dict(_LASER, inverse_bremsstrahlung=1, spot_w0=99999725.41, cntr_x0=0.0,
                                        ab_depth=5640.0, rise_time=1e-9, flat_time=358927.4021,
                                        fall_time=178352.46))),
    ("tzoufras-secIV.B", dict(l0=16, m0=3, np=300, nx=96, ny=48, pmax=0.6, xmin=-23780.328, xmax=17835.246,
                              ymin=-44588.115, ymax=44588.115, es=0, bndx=1, density_np=1.0e23, zeta=4.0,
                              small_dt=0.09, nb=6, t_stop=2140299.52, n_out=960,
                              src=dict(_LASER, linear_Ivst=1, spot_w0=14862.705, cntr_x0=-10701.0, ab_depth=7134.0,
                                       rise_time=178352.46, flat_time=1783524.6, fall_time=178352.46))),
)
POP_NBC = 2  # RKLevel


def pop_h(d):
    """the deck's step: num_h = int(dt_out / CLF) + 1, h = dt_out / num_h, CLF = clf_dp * dp (clf_dp = 30)"""
    pmin = d["pmax"] / (2.0 * d["np"])
    clf = 30.0 * (d["pmax"] - pmin) / (d["np"] - 1)
    dt_out = d["t_stop"] / d["n_out"]
    return dt_out / (int(dt_out / clf) + 1)


def pop_state(d, nx=None, seed=7):
    """Maxwellian f00 (T = 1 keV, 5% density and 10% temperature ripples in x), 1e-3 random f_lm, small E and B"""
    from oracle.oracle import hind, num_harmonics
    l0, m0, np_, ny = d["l0"], d["m0"], d["np"], d["ny"]
    nx = nx or d["nx"]
    rng = np.random.default_rng(seed)
    pmin = d["pmax"] / (2.0 * np_)
    p = np.arange(np_) * ((d["pmax"] - pmin) / (np_ - 1)) + pmin
    x = np.arange(nx) / nx
    y = np.arange(ny) / ny
    T = (1.0 + 0.1 * np.cos(2 * np.pi * x))[None, :] * np.ones((ny, 1))
    pt = np.sqrt(T / 511.0)
    n = (1.0 + 0.05 * np.sin(2 * np.pi * x))[None, :] * np.ones((ny, 1))
    f = np.zeros((num_harmonics(l0, m0), ny, nx, np_), np.complex128)
    f[0] = (n / (np.sqrt(2 * np.pi) * 2 * np.pi * pt ** 3))[..., None] * np.exp(-p ** 2 / (2 * pt[..., None] ** 2))
    for l in range(1, l0 + 1):
        for m in range(min(l, m0) + 1):
            r = rng.uniform(-1, 1, (ny, nx, np_)) + (1j * rng.uniform(-1, 1, (ny, nx, np_)) if m else 0.0)
            f[hind(l, m, m0)] = 1e-3 * r * f[0].real
    fld = np.zeros((6, ny, nx), np.complex128)
    fld[0] = 1e-4 * np.sin(2 * np.pi * x)[None, :]
    fld[1] = 2e-5 * np.cos(2 * np.pi * x)[None, :]
    fld[2] = -1e-5 * np.sin(2 * np.pi * y)[:, None]
    fld[5] = 2e-3
    return f, fld


def run_pop(steps, warmup, no_cpu=False, deck=None):
    """The implicit-E step without if_implicit1D (oshun_step_implicit_2d) on the three production decks'
    grids, boundaries (tzoufras-secIV.B: x mirror with RKLevel-2 guards), collision and laser settings and
    time step, FP64; the reference's op-9 driver (oracle/_ref) timed on a 4-cell x slab of each."""
    import oshun_amd
    from oracle import oracle as orc
    out = []
    for name, d in POP_DECKS:
        if deck and name != deck:
            continue
        h = pop_h(d)
        dx = (d["xmax"] - d["xmin"]) / (d["nx"] - 1)
        dy = (d["ymax"] - d["ymin"]) / (d["ny"] - 1)
        f, fld = pop_state(d)
        nbc = POP_NBC if d["bndx"] else 0
        ctx = oshun_amd.Context(l0=d["l0"], m0=d["m0"], nump=d["np"], numx=d["nx"], numy=d["ny"], pmax=d["pmax"],
                                dx=dx, dy=dy, density_np=d["density_np"], zeta=d["zeta"], small_dt=d["small_dt"],
                                nb_algorithms=d["nb"], bndx=d["bndx"], nbc=nbc)
        ctx.upload(f, fld)
        if d["src"]:
            x, xg = orc.deck_axis(d["nx"], d["xmin"], d["xmax"], guards=POP_NBC, whole=bool(d["bndx"]))
            y, _ = orc.deck_axis(d["ny"], d["ymin"], d["ymax"], guards=POP_NBC, whole=bool(d["bndx"]))
            ctx.set_sources(d["src"], x, y, xg)
        ctx.step_implicit_2d(h, warmup, d["es"])
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.step_implicit_2d(h, steps, d["es"])
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        ctx.profile(True)
        ctx.step_implicit_2d(h, steps, d["es"])
        prof = ctx.profile_read()
        ctx.profile(False)
        fo, fldo = ctx.download()
        if not (np.all(np.isfinite(fo)) and np.all(np.isfinite(fldo))):
            raise RuntimeError(f"{name}: non-finite state after the timed steps")
        nh = orc.num_harmonics(d["l0"], d["m0"])
        cells = nh * d["np"] * d["nx"] * d["ny"]
        rec = {"deck": name, "l0": d["l0"], "m0": d["m0"], "np": d["np"], "nx": d["nx"], "ny": d["ny"],
               "es": d["es"], "bndx": d["bndx"], "sources": bool(d["src"]), "h": h, "ms_per_step": ms,
               "cell_updates_per_s": cells / (ms * 1e-3), "cells": cells, "state_bytes": ctx.state_bytes(),
               "kernels": {k: {"ms_per_step": v[0] / steps, "launches_per_step": v[1] / steps}
                           for k, v in prof.items() if v[1]}}
        del ctx
        if not no_cpu and orc.have_ref():
            nxs = 4
            fs, flds = pop_state(d, nx=nxs)
            kw = dict(nb=d["nb"], pmax=d["pmax"], xmin=0.0, xmax=dx * (nxs - 1), ymin=d["ymin"], ymax=d["ymax"],
                      h=h, small_dt=d["small_dt"], density_np=d["density_np"], zeta=d["zeta"], es=d["es"],
                      src=d["src"], bnd=(d["bndx"], 0), nbc=POP_NBC)
            t0 = time.perf_counter()
            orc.run_ref(d["l0"], d["m0"], d["np"], nxs, d["ny"], fs, flds, op=9, nsteps=1, **kw)
            t1 = time.perf_counter()
            orc.run_ref(d["l0"], d["m0"], d["np"], nxs, d["ny"], fs, flds, op=9, nsteps=3, **kw)
            t2 = time.perf_counter()
            per_step = max((t2 - t1) - (t1 - t0), 1e-9) / 2
            sc = nh * d["np"] * nxs * d["ny"]
            rec["cpu_baseline"] = {"value": sc / per_step, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                                   "sample": f"OSHUN-1 classes (oracle/_ref, op 9), the deck's l0, m0, Np, Ny, "
                                             f"boundary and sources on Nx = {nxs} (+ guards), a 3-step minus a "
                                             f"1-step run",
                                   "s_per_step": per_step}
            rec["gpu_over_cpu_core"] = rec["cell_updates_per_s"] / rec["cpu_baseline"]["value"]
        out.append(rec)
        print(json.dumps(rec), file=sys.stderr, flush=True)
    return {"metric": "phase-space cell-updates/s per implicit-E RK2+FP step (2D variant, FP64), production decks",
            "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "dtype": "f64",
            "data": "synthetic (Maxwellian with ripples and 1e-3 random harmonics on each deck's grid)",
            "decks": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32"], default=None,
                    help="profiling aid: run one path only and print its timings (no JSON contract line)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="profiling aid: also launch one plain copy of S_f bytes (counter calibration)")
    ap.add_argument("--batch", action="store_true",
                    help="the batched ensemble (512 independent 1D runs) instead of the single-run workload")
    ap.add_argument("--configs", action="store_true",
                    help="throughput of the RK4 + FP step on the synthetic configurations C1, C2, C3, C5")
    ap.add_argument("--config", default=None, help="with --configs: only the configuration of this name (C3)")
    ap.add_argument("--implicit", action="store_true",
                    help="the implicit-E step of the 1D variant instead of the RK4 step (FP64)")
    ap.add_argument("--pop", action="store_true",
                    help="the implicit-E step without if_implicit1D on the three production decks' grids (FP64)")
    ap.add_argument("--deck", default=None, choices=[n for n, _ in POP_DECKS],
                    help="with --pop: only the deck of this name")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for {args.gpus}); "
              "a single run shards along x with one halo exchange per step", file=sys.stderr)
        sys.exit(2)
    import oshun_amd

    if args.configs:
        print(json.dumps(run_configs(args.steps, args.warmup, args.config, args.only, args.no_cpu)))
        return
    if args.implicit:
        print(json.dumps(run_implicit(args.steps, args.warmup, args.no_cpu)))
        return
    if args.pop:
        print(json.dumps(run_pop(args.steps, args.warmup, args.no_cpu, args.deck)))
        return
    if args.batch:
        print(json.dumps(run_ensemble(args.steps, args.warmup)))
        return
    f, fld, dx = target_state()
    cells = (L0 + 1) * NP * NX
    if args.only:
        prec = oshun_amd.FP64 if args.only == "fp64" else oshun_amd.FP32_SPLIT
        ctx, ms, prof = run_path(prec, args.steps, args.warmup, f, fld, dx)
        if args.copy_calib:
            ctx.diag_copy(ctx.state_bytes())
            ctx.synchronize()
        print(json.dumps({"path": args.only, "ms_per_step": ms, "state_bytes": ctx.state_bytes(),
                          "kernels": {k: v for k, v in prof.items()}}))
        return
    ctx64, ms64, prof64 = run_path(oshun_amd.FP64, args.steps, args.warmup, f, fld, dx)
    S64 = ctx64.state_bytes()
    ctx32, ms32, prof32 = run_path(oshun_amd.FP32_SPLIT, args.steps, args.warmup, f, fld, dx)
    S32 = ctx32.state_bytes()

    def pair_kernel(ctx):
        # the stage-pair family the context runs (marched kernels: negative rows)
        return "vlasov_march_1d" if ctx.pair_rows() < 0 else "vlasov_pair_1d"

    def roof(prof, S, steps, ctx, prec_key):
        # the Vlasov launches of one RK4 step (two fused stage pairs on the 1D
        # path) share the 16 S_f algorithmic bytes of the step
        ms, n = prof["vlasov_stage"]
        avg_s = ms / max(n, 1) * 1e-3
        per_step = n / steps
        bytes_per_launch = 16.0 * S / per_step
        ach = bytes_per_launch / avg_s
        r = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
             "frac": ach / HBM_PEAK, "traffic": None, "kernel": pair_kernel(ctx),
             "kernel_avg_us": avg_s * 1e6, "bytes_per_launch": bytes_per_launch,
             "launches_per_step": per_step}
        # HBM bytes per launch measured by rocprofv3 PMC (FETCH_SIZE, WRITE_SIZE)
        # on this workload, calibrated on a plain copy (profiles/traffic_*.json):
        # the mean over the two pair launches of the family that ran
        tpath = os.path.join(ROOT, "profiles", "traffic_l8_np256_nx4096.json")
        if os.path.exists(tpath):
            with open(tpath) as fh:
                tr = json.load(fh).get(prec_key, {})
            pairs = [v for name, v in tr.get("kernels", {}).items() if name.startswith("oshun::" + r["kernel"])]
            if pairs:
                r["traffic"] = sum(v["read_bytes"] + v["write_bytes"] for v in pairs) / len(pairs)
                r["traffic_source"] = os.path.relpath(tpath, ROOT)
                # the bytes the launches really move, at the same mean duration
                r["traffic_gbs"] = r["traffic"] / avg_s / 1e9
                r["traffic_frac"] = r["traffic"] / avg_s / HBM_PEAK
        return r

    def breakdown(prof, steps):
        return {k: {"ms_per_step": v[0] / steps, "launches_per_step": v[1] / steps}
                for k, v in prof.items() if v[1]}

    r64 = roof(prof64, S64, args.steps, ctx64, "fp64")
    r32 = roof(prof32, S32, args.steps, ctx32, "fp32")
    (f64, fl64), (f32, fl32) = ctx64.final, ctx32.final

    def l2(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    split_err = {"err_f00": l2(f32[0], f64[0]), "err_f10": l2(f32[1], f64[1]), "err_Ex": l2(fl32[0], fl64[0]),
                 "after_steps": args.warmup + args.steps, "norm": "relative L2 against the FP64 path"}
    step_frac64 = 16.0 * S64 / (ms64 * 1e-3) / HBM_PEAK
    step_frac32 = 16.0 * S32 / (ms32 * 1e-3) / HBM_PEAK
    cpu = None if args.no_cpu else cpu_baseline()
    value = cells / (ms64 * 1e-3)
    line = {
        "metric": "phase-space cell-updates/s per RK4+FP step (FP64 vs FP32 split), % HBM roofline",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Maxwellian T=1 keV, n=1+1e-3 sin kx, f_l>=1 seeded 1e-6 noise)",
        "config": {"workload": "1D periodic Langmuir/collisional, l0=8 m0=0 (9 harmonics), Np=256 pmax=0.3, "
                               "Nx=4096, k lambda_D=0.3 x16 wavelengths, h=0.01, e-e+e-i collisions on "
                               "(n=1e21 cm^-3, Z=1, small_dt=h, NB=6)",
                   "cells": cells, "state_bytes_fp64": S64, "state_bytes_fp32": S32},
        "roofline": r64,
        "step_roofline_frac": step_frac64,
        "conservation": ctx64.cons,
        "fp32_split": {"value": cells / (ms32 * 1e-3), "ms_per_step": ms32, "dtype": "f32 f / f64 fields+collisions",
                       "speedup_vs_fp64": ms64 / ms32, "roofline": r32, "step_roofline_frac": step_frac32,
                       "conservation": ctx32.cons, **split_err},
        "kernels_fp64": breakdown(prof64, args.steps),
        "kernels_fp32_split": breakdown(prof32, args.steps),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
