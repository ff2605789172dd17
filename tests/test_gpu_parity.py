"""Parity of the HIP path (through the C ABI) with the reference and the oracle.

Tolerances (the FP64 path is held to 1e-12 relative in f_lm, E and B):
  * against fixtures written by the compiled reference: <= 1e-12 normwise
    (max|gpu - ref| / max|ref| over the whole array);
  * against the oracle at larger sizes: <= 1e-12 normwise.
"""
import numpy as np
import pytest

from oracle.oracle import Config, Oracle, hind, num_harmonics

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def ctx_from_golden(g, **over):
    import oshun_amd
    c = g["cfg"]
    kw = dict(l0=int(c["l0"]), m0=int(c["m0"]), nump=int(c["np"]), numx=int(c["nx"]), numy=int(c["ny"]),
              pmax=float(c["pmax"]), dx=float(g["dx"]), dy=float(g["dy"]), density_np=float(c["density_np"]),
              zeta=float(c["zeta"]), small_dt=float(c["small_dt"]), nb_algorithms=int(c["nb"]))
    kw.update(over)
    return oshun_amd.Context(**kw)


@pytest.mark.parametrize("name", ["rhs_2d_m3", "rhs_quirk_l7"])
def test_rhs_vs_reference(golden, name):
    g = golden(name)
    print(f"case {name}", flush=True)
    ctx = ctx_from_golden(g)
    ctx.upload(g["f_in"], g["fld_in"])
    f, fld = ctx.rhs()
    assert rel(f, g["f_out"]) < TOL
    assert rel(fld.real, g["fld_out"].real) < TOL


@pytest.mark.parametrize("name", ["step_2d_c3", "rk4_1d_nocoll", "ee_only", "flm_only"])
def test_steps_vs_reference(golden, name):
    g = golden(name)
    c = g["cfg"]
    print(f"case {name}", flush=True)
    ctx = ctx_from_golden(g)
    ctx.upload(g["f_in"], g["fld_in"])
    h, n, op = float(c["h"]), int(c["nsteps"]), int(c["op"])
    if op == 1:
        ctx.step(h, n)
    elif op == 2:
        for _ in range(n):
            ctx.rk4_step(h)
    else:
        t = 0.0
        for _ in range(n):
            t += h
            if op == 3:
                ctx.collide_ee(t)
            else:
                ctx.collide_flm(t, 1)
                ctx.collide_flm(t, 2)
    f, fld = ctx.download()
    assert rel(f, g["f_out"]) < TOL
    if np.abs(g["fld_out"]).max() > 0:
        assert rel(fld.real, g["fld_out"].real) < TOL


def test_m0_zero_vs_reference_standin(golden):
    """l0 = 8, m0 = 0 on the GPU vs the reference run as an m0 = 2 y-uniform stand-in."""
    g = golden("step_1d_m0")
    c = g["cfg"]
    print("case step_1d_m0", flush=True)
    ctx = ctx_from_golden(g, m0=0)
    l0, nx, np_ = int(c["l0"]), int(c["nx"]), int(c["np"])
    f = np.zeros((l0 + 1, 1, nx, np_), np.complex128)
    f[:] = g["f_in"]
    fld = np.zeros((6, 1, nx), np.complex128)
    fld[0] = g["fld_in"][0]
    ctx.upload(f, fld)
    ctx.step(float(c["h"]), int(c["nsteps"]))
    fo, fldo = ctx.download()
    assert rel(fo.real, g["f_out"]) < TOL
    assert rel(fldo[0].real, g["fld_out"][0]) < TOL


def _maxwellian_state(l0, m0, np_, nx, ny, pmax, seed, amp=1e-3):
    rng = np.random.default_rng(seed)
    nh = num_harmonics(l0, m0)
    pmin = pmax / (2.0 * np_)
    p = np.arange(np_) * ((pmax - pmin) / (np_ - 1)) + pmin
    x = np.arange(nx) / nx
    T = (1.0 + 0.1 * np.cos(2 * np.pi * x))[None, :] * np.ones((ny, 1))
    pt = np.sqrt(T / 511.0)
    n = 1.0 + 1e-3 * np.sin(2 * np.pi * 4 * x)[None, :]
    f = np.zeros((nh, ny, nx, np_), np.complex128)
    f[0] = (n / (np.sqrt(2 * np.pi) * 2 * np.pi * pt ** 3))[..., None] * \
        np.exp(-p[None, None, :] ** 2 / (2 * pt[..., None] ** 2))
    for l in range(1, l0 + 1):
        for m in range(min(l, m0) + 1):
            r = rng.uniform(-1, 1, (ny, nx, np_))
            if m:
                r = r + 1j * rng.uniform(-1, 1, (ny, nx, np_))
            f[hind(l, m, m0)] = amp * r * f[0].real
    fld = np.zeros((6, ny, nx), np.complex128)
    fld[0] = 1e-5 * np.sin(2 * np.pi * x)[None, :]
    if m0 >= 1:
        fld[5] = 1e-3
    return f, fld


# (l0, m0, Np, Nx, Ny, steps, h, e-e sub-step); h and the e-e sub-step keep the
# explicit e-e operator inside its stability region (dt ~ dv^2 / D)
STEP_CASES = [
    (8, 0, 256, 64, 1, 4, 0.5, 0.1),    # target harmonics and momentum grid, short x
    (2, 0, 64, 64, 1, 5, 0.5, 0.1),     # the l0 = 2 heat-flow shape (reference cannot run l0 < 3)
    (1, 0, 100, 37, 1, 3, 0.5, 0.1),    # ragged: Np not a multiple of 64, Nx not a multiple of 64
    (3, 0, 192, 48, 1, 3, 0.5, 0.1),    # adjoint-config harmonics, Np = 192
    (4, 4, 128, 12, 10, 2, 0.5, 0.1),   # 2D magnetised, complex harmonics (hot-spot shape, small)
    (6, 3, 40, 9, 6, 2, 0.5, 0.1),      # exercises the y-advection coefficient slip
    (5, 1, 512, 8, 1, 2, 0.05, 0.01),   # Np = 512 (8 rows per lane)
]


@pytest.mark.parametrize("l0,m0,np_,nx,ny,steps,h,sdt", STEP_CASES)
def test_steps_vs_oracle(l0, m0, np_, nx, ny, steps, h, sdt):
    print(f"case l0={l0} m0={m0} np={np_} nx={nx} ny={ny}", flush=True)
    import oshun_amd
    pmax = 0.3
    f, fld = _maxwellian_state(l0, m0, np_, nx, ny, pmax, seed=l0 * 100 + np_)
    kw = dict(l0=l0, m0=m0, pmax=pmax, dx=0.4, dy=0.5, density_np=1e22, small_dt=sdt, nb=6)
    o = Oracle(Config(np=np_, nx=nx, ny=ny, **kw))
    o.set_state(f, fld)
    o.step(h, steps)
    fo, fldo = o.get_state()
    assert np.all(np.isfinite(fo))
    ctx = oshun_amd.Context(l0, m0, np_, nx, ny, pmax=pmax, dx=0.4, dy=0.5, density_np=1e22, small_dt=sdt,
                            nb_algorithms=6)
    ctx.upload(f, fld)
    ctx.step(h, steps)
    fg, fldg = ctx.download()
    assert rel(fg, fo) < TOL
    for k in range(6):
        if np.abs(fldo[k]).max() > 0:
            assert rel(fldg[k].real, fldo[k].real) < TOL, k
    assert rel(fo, f) > 1e-7  # the state moved


def test_fp32_split_rhs_close_to_oracle():
    """FP32 Vlasov slope (FP32 storage and accumulation) vs the FP64 oracle: float rounding only."""
    import oshun_amd
    l0, m0, np_, nx = 8, 0, 256, 64
    f, fld = _maxwellian_state(l0, m0, np_, nx, 1, 0.3, seed=4)
    fld[0] = 1e-3 * np.cos(2 * np.pi * np.arange(nx) / nx)[None, :]
    o = Oracle(Config(l0=l0, m0=m0, np=np_, nx=nx, pmax=0.3, dx=0.5))
    o.set_state(f, fld)
    so, sflo = o.rhs()
    ctx = oshun_amd.Context(l0, m0, np_, nx, 1, pmax=0.3, dx=0.5, precision=oshun_amd.FP32_SPLIT)
    ctx.upload(f, fld)
    sg, sflg = ctx.rhs()
    assert np.all(np.isfinite(sg.real))
    for l in range(l0 + 1):
        assert rel(sg[l].real, so[l].real) < 2e-5, l
    assert rel(sflg[0].real, sflo[0].real) < 1e-5


def test_fp32_split_close_to_fp64():
    """FP32 storage + FP32 Vlasov RHS, FP64 fields and collisions: error vs FP64 stays small."""
    import oshun_amd
    l0, m0, np_, nx = 8, 0, 256, 64
    f, fld = _maxwellian_state(l0, m0, np_, nx, 1, 0.3, seed=3)
    out = {}
    for prec in (oshun_amd.FP64, oshun_amd.FP32_SPLIT):
        # dx = 0.5 keeps v h / dx = 0.29 inside the RK4 stability region
        ctx = oshun_amd.Context(l0, m0, np_, nx, 1, pmax=0.3, dx=0.5, density_np=1e22, small_dt=0.1,
                                precision=prec)
        ctx.upload(f, fld)
        ctx.step(0.5, 10)
        out[prec] = ctx.download()
    f64, fl64 = out[oshun_amd.FP64]
    f32, fl32 = out[oshun_amd.FP32_SPLIT]
    assert np.all(np.isfinite(f64.real)) and np.all(np.isfinite(f32.real))
    e_f00 = np.linalg.norm(f32[0] - f64[0]) / np.linalg.norm(f64[0])
    e_ex = np.linalg.norm(fl32[0] - fl64[0]) / np.linalg.norm(fl64[0])
    print("fp32 split rel L2: f00", e_f00, "Ex", e_ex)
    assert e_f00 < 1e-5
    assert e_ex < 1e-4


def _drift_state(l0, m0, np_, nx, ny, pmax, precision_bg=False):
    """Drifting Maxwellians with T(x, y) and drift u(x, y) in all three directions, projected onto the
    harmonics: nonzero f11, f21, f22, so that every moment term (VxVy Vy in Qx among them) is exercised."""
    from tests.sph import maxwellian3, project
    nh = num_harmonics(l0, m0)
    pmin = pmax / (2.0 * np_)
    p = np.arange(np_) * ((pmax - pmin) / (np_ - 1)) + pmin
    f = np.zeros((nh, ny, nx, np_), np.complex128)
    for iy in range(ny):
        for ix in range(nx):
            ph = 2 * np.pi * (ix / nx + 0.13)
            tt = 1.0 + 0.2 * np.sin(ph) + 0.1 * np.cos(2 * np.pi * (iy + 0.3) / ny)
            pt = np.sqrt(tt / 511.0)
            u = (0.4 * pt * np.cos(ph), 0.5 * pt * np.sin(ph + iy), -0.3 * pt)
            f[:, iy, ix, :] = project(maxwellian3(1.0 + 0.1 * np.cos(ph), pt, u), p, l0, m0, nmu=24, nphi=24)
    return f, np.zeros((6, ny, nx), np.complex128)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_moments_vs_oracle(prec):
    """All 18 moments (def-Parallel.cpp:1677-2253) of the GPU against the oracle at <= 1e-12 normwise
    (FP32 split: against the oracle run on the GPU's own float-rounded state)."""
    import oshun_amd
    l0, m0, np_, nx, ny = 3, 2, 96, 12, 5
    f, fld = _drift_state(l0, m0, np_, nx, ny, 0.3)
    precision = oshun_amd.FP64 if prec == "fp64" else oshun_amd.FP32_SPLIT
    ctx = oshun_amd.Context(l0, m0, np_, nx, ny, pmax=0.3, dx=0.7, dy=1.1, density_np=3e21, zeta=2.0,
                            precision=precision)
    ctx.upload(f, fld)
    mg = ctx.moments()
    fs, _ = ctx.download()  # the state the GPU holds (float-rounded in the split)
    o = Oracle(Config(l0=l0, m0=m0, np=np_, nx=nx, ny=ny, pmax=0.3, dx=0.7, dy=1.1, density_np=3e21, zeta=2.0))
    o.set_state(fs, fld)
    mo = o.moments()
    assert np.abs(mo[7]).max() > 0 and np.abs(mo[9]).max() > 0  # VxVy, VyVz seen
    for k in range(len(oshun_amd.MOMENT_NAMES)):
        assert rel(mg[k], mo[k]) < TOL, (k, oshun_amd.MOMENT_NAMES[k], rel(mg[k], mo[k]))


def test_conservation_at_target_size():
    """Full target shape (l0=8, Np=256, Nx=4096, collisions on), 20 steps:
    total number and energy are conserved (size-independent properties)."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    import oshun_amd
    f, fld, dx = bench.target_state()
    ctx = bench.make_ctx(oshun_amd.FP64, dx)
    ctx.upload(f, fld)
    m0 = ctx.moments()
    ctx.step(bench.H, 20)
    m1 = ctx.moments()
    N0, N1 = m0[0].sum(), m1[0].sum()
    E0 = (m0[0] * (m0[3] + m0[4] + m0[5])).sum()
    E1 = (m1[0] * (m1[3] + m1[4] + m1[5])).sum()
    print("dN/N", (N1 - N0) / N0, "dE/E", (E1 - E0) / E0)
    assert abs(N1 - N0) < 1e-9 * N0
    assert abs(E1 - E0) < 1e-6 * E0
    g, gf = ctx.download()
    assert np.all(np.isfinite(g.real))
