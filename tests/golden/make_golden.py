"""Write the golden fixtures of the reference's hot path.

Runs oracle/_ref/oshun_ref -- a driver linked with the unmodified reference
translation units (see oracle/Makefile) -- on small seeded states and stores
inputs and outputs as .npz files next to this script.  Run from the repo root:

    make -C oracle all ref && python tests/golden/make_golden.py

Each fixture holds: the integer/float configuration, the interior input state
(f: complex128[nh,ny,nx,np], fld: complex128[6,ny,nx]), the axis and cell sizes
the reference actually used, and the interior output state.

Cases (the reference needs l0 >= 3 and m0 >= 2, def-Input.cpp:73):
  rhs_2d_m3      Actions::Adv slope, l0=4 m0=3, 2D, random complex state
  rhs_quirk_l7   Actions::Adv slope, l0=7 m0=3 (exercises def-VlasovMax.cpp:918)
  step_1d_m0     3 full steps (RK4 + e-e + f1 + flm), l0=8 m0=0 physics run
                 through the reference as an m0=2, y-uniform stand-in with all
                 m>=1 harmonics and Ey, Ez, B zero; only m=0 is stored
  step_2d_c3     2 full steps, l0=4 m0=4, 2D, uniform Bz (a small C3 shape)
  rk4_1d_nocoll  4 RK4 steps, collisions off (density 1e-30), l0=3 m0=2
  ee_only        3 e-e calls (Explicit_EC_FP::loop)
  flm_only       3 implicit f_lm calls (Euler_Backward::advance_1/advance_lm)
  imp_1d_m0      3 implicit-E steps (Runge_Kutta_2_Imp, Implicit_E_Field with
                 if_implicit1D, e-e, f1, flm), l0=5 m0=0 physics through the
                 m0=2 y-uniform stand-in, if_implicitES on
  imp_1d_m0_l8   the same at l0=8, Np=64, if_implicitES off (Faraday in
                 Adv1; the transverse fields are zero)
  pls_only       3 PLaserSource::Laser_Source calls (polynomial rise), 2D
  ib_only        3 InverseBremsstrahlung::loop calls, 2D
  efield_only    3 ExternalEfield::Laser_Efield calls (Ex polarisation), 2D
  step_src       2 full steps with the laser source and inverse
                 bremsstrahlung on (main.cpp:79-107 with :90-92), 2D
  imp_2d_m2      2 implicit-E steps without if_implicit1D (the 3x3 field
                 solve, Ampere JN), l0=3 m0=2, 2D, E and B nonzero
  imp_2d_m3_es   2 such steps, l0=4 m0=3, if_implicitES (JN = the input Ex)
  imp_2d_l7_quirk 1 step, l0=7 m0=3 (the y-advection slip inside Adv1)
  imp_2d_dilute  1 step with two cells diluted by 1e-150 (zero determinant:
                 4 attempts, E = 0 there)
  imp_2d_src     2 steps with the laser source and inverse bremsstrahlung
  mirror_x_2d    3 full steps with an x mirror boundary (whole arrays with
                 guards kept: f_ext0, f_ext), l0=4 m0=2
  mirror_y_2d    3 full steps with a y mirror boundary, l0=3 m0=3
  mirror_xy_rk4  4 RK4 steps with both mirrors, l0=3 m0=2
  mirror_imp_x   2 implicit-E steps (2D variant, RKLevel-2 guards), x mirror
  mirror_imp_x_es 2 such steps, l0=3 m0=3, if_implicitES (JN = the initial
                 Ex, guards included); both with the reference's own one-ulp
                 noise floor of the field solve (floor_f, floor_fld).  (A y
                 mirror leaves f00's guards stale, def-Parallel.cpp:1258, and
                 the field solve next to it is then ill-conditioned: ~1e-3
                 one-ulp floor; not a useful fixture.)
  mirror_imp_x_src 2 such steps, x mirror, with linear laser heating and
                 inverse bremsstrahlung (the production deck's whole step)

`python tests/golden/make_golden.py name ...` writes only the named fixtures.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.oracle import hind, num_harmonics, run_ref  # noqa: E402

# the source decks: 351 nm light, polynomial ramp over 1/omega_p, flat top to 5
SRC = dict(linear_Ivst=0, polynomial_Ivst=1, inverse_bremsstrahlung=0, polarization_direction=0, I_0=3.0e14,
           lambda_0=0.351, ab_eta=0.5, ab_depth=8.0, cntr_x0=20.0, cntr_r0=10.0, spot_w0=15.0, rise_time=1.0,
           flat_time=4.0, fall_time=1.0, smaller_dt=0.2)


def p_axis(np_, pmax):
    pmin = pmax / (2.0 * np_)
    d = (pmax - pmin) / float(np_ - 1)
    return np.arange(np_) * d + pmin


def maxwellian_state(l0, m0, np_, nx, ny, pmax, seed, amp=1e-3, bz=0.0, eyz=True):
    """Perturbed Maxwellian f00 (def-Setup.cpp:498-520 formula) + small f_lm."""
    rng = np.random.default_rng(seed)
    nh = num_harmonics(l0, m0)
    p = p_axis(np_, pmax)
    x = np.arange(nx) / nx
    y = np.arange(ny) / max(ny, 1)
    T = (1.0 + 0.1 * np.cos(2 * np.pi * x))[None, :] * (1.0 + 0.05 * np.sin(2 * np.pi * y))[:, None]
    pt = np.sqrt(T / 511.0)
    n = (1.0 + 0.05 * np.sin(2 * np.pi * x))[None, :] * np.ones((ny, 1))
    f = np.zeros((nh, ny, nx, np_), np.complex128)
    f[0] = (n / (np.sqrt(2 * np.pi) * 2 * np.pi * pt ** 3))[..., None] * \
        np.exp(-p[None, None, :] ** 2 / (2 * pt[..., None] ** 2))
    for l in range(1, l0 + 1):
        for m in range(min(l, m0) + 1):
            r = rng.uniform(-1, 1, (ny, nx, np_))
            if m > 0:
                r = r + 1j * rng.uniform(-1, 1, (ny, nx, np_))
            f[hind(l, m, m0)] = amp * r * f[0].real
    fld = np.zeros((6, ny, nx), np.complex128)
    fld[0] = 1e-4 * np.sin(2 * np.pi * x)[None, :] * np.ones((ny, 1))
    if eyz:
        fld[1] = 2e-5 * np.cos(2 * np.pi * x)[None, :]
        fld[2] = -1e-5 * np.sin(2 * np.pi * y)[:, None]
    fld[5] = bz
    return f, fld


def random_state(l0, m0, np_, nx, ny, seed):
    rng = np.random.default_rng(seed)
    nh = num_harmonics(l0, m0)
    f = rng.uniform(-1, 1, (nh, ny, nx, np_)) + 1j * rng.uniform(-1, 1, (nh, ny, nx, np_))
    for l in range(l0 + 1):
        f[hind(l, 0, m0)] = f[hind(l, 0, m0)].real
    fld = rng.uniform(-1, 1, (6, ny, nx)) + 0j
    return f, fld


def save(name, cfg, fin, fldin, res, keep_m0_only=False, m0_phys=None):
    l0, m0 = cfg["l0"], cfg["m0"]
    fout = res["f"]
    if keep_m0_only:
        idx = [hind(l, 0, m0) for l in range(l0 + 1)]
        fin = fin[idx].real.copy()
        fout = fout[idx].real.copy()
        fldin = fldin[0:1].real.copy()
        fldout = res["fld"][0:1].real.copy()
        cfg = dict(cfg, m0=m0_phys, standin_m0=m0)
    else:
        fldout = res["fld"]
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, f_in=fin, fld_in=fldin, f_out=fout, fld_out=fldout, p=res["p"],
                        dx=res["dx"], dy=res["dy"], h=res["h"], t=res["t"],
                        **{"cfg_" + k: v for k, v in cfg.items()})
    print(f"{name}: {os.path.getsize(path)} bytes")


def source_cases(only):
    base = dict(l0=3, m0=2, np=48, nx=6, ny=3, pmax=0.3, xmin=0.0, xmax=50.0, ymin=0.0, ymax=30.0,
                nsteps=3, nb=4, h=0.5, small_dt=0.2, density_np=1e21, zeta=1.0)
    cases = (("pls_only", 6, dict(SRC)), ("ib_only", 7, dict(SRC, polynomial_Ivst=0, inverse_bremsstrahlung=1)),
             ("efield_only", 8, dict(SRC, polynomial_Ivst=0, polarization_direction=1)),
             ("step_src", 1, dict(SRC, inverse_bremsstrahlung=1)))
    for name, op, src in cases:
        if only and name not in only:
            continue
        c = dict(base, op=op, nsteps=2 if op == 1 else 3)
        f, fld = maxwellian_state(c["l0"], c["m0"], c["np"], c["nx"], c["ny"], c["pmax"], 40 + op)
        r = run_ref(c["l0"], c["m0"], c["np"], c["nx"], c["ny"], f, fld, op=op, nsteps=c["nsteps"],
                    pmax=c["pmax"], xmin=c["xmin"], xmax=c["xmax"], ymin=c["ymin"], ymax=c["ymax"], h=c["h"],
                    small_dt=c["small_dt"], density_np=c["density_np"], nb=c["nb"], src=src)
        save(name, dict(c, **{"src_" + k: v for k, v in src.items()}), f, fld, r)


def implicit_2d_state(c, seed):
    """Perturbed Maxwellian with all three E components and B nonzero."""
    f, fld = maxwellian_state(c["l0"], c["m0"], c["np"], c["nx"], c["ny"], c["pmax"], seed, bz=2e-3)
    x = np.arange(c["nx"]) / c["nx"]
    y = np.arange(c["ny"]) / c["ny"]
    fld[3] = 1e-3 * np.cos(2 * np.pi * y)[:, None] * np.ones((1, c["nx"]))
    fld[4] = -1.5e-3 * np.sin(2 * np.pi * x)[None, :] * np.ones((c["ny"], 1))
    return f, fld


def implicit_2d_cases(only):
    """Implicit-E steps without if_implicit1D (harness op 9): the 3x3 field solve."""
    base = dict(pmax=0.25, xmin=0.0, xmax=60.0, ymin=0.0, ymax=40.0, op=9, nb=4, h=0.5, small_dt=0.2,
                density_np=1e21, zeta=1.0)
    cases = (
        ("imp_2d_m2", dict(l0=3, m0=2, np=24, nx=6, ny=4, nsteps=2, es=0), 51, None, None),
        # electrostatic: JN is the Ex the solver was built with (the input's)
        ("imp_2d_m3_es", dict(l0=4, m0=3, np=32, nx=5, ny=4, nsteps=2, es=1), 52, None, None),
        # l0 = 7, m0 = 3: the y-advection coefficient slip inside Adv1
        ("imp_2d_l7_quirk", dict(l0=7, m0=3, np=20, nx=4, ny=5, nsteps=1, es=0), 53, None, None),
        # a 5 x 6 block of cells diluted by 1e-150: the two cells two cells
        # inside it stay dilute through Adv1's two advection stages, their
        # determinant underflows to zero, the probes are retried 4 times and
        # those cells get E = 0.  They lie 4 or more cells from the edges, so
        # that no guard cell holds a copy of them (the reference counts zero
        # determinants on guard cells too and stops the run above 8)
        ("imp_2d_dilute", dict(l0=3, m0=2, np=16, nx=11, ny=9, nsteps=1, es=0), 54,
         ((slice(2, 7), slice(3, 9)), ((4, 5), (4, 6))), None),
        # sources after the e-e collisions (polynomial laser heating, inverse bremsstrahlung)
        ("imp_2d_src", dict(l0=3, m0=2, np=48, nx=6, ny=3, nsteps=2, es=0, pmax=0.3, xmax=50.0, ymax=30.0), 55,
         None, dict(SRC, inverse_bremsstrahlung=1)),
    )
    for name, cc, seed, dilute, src in cases:
        if only and name not in only:
            continue
        c = dict(base, **cc)
        f, fld = implicit_2d_state(c, seed)
        if dilute:
            f[:, dilute[0][0], dilute[0][1], :] *= 1e-150
        r = run_ref(c["l0"], c["m0"], c["np"], c["nx"], c["ny"], f, fld, op=9, nsteps=c["nsteps"],
                    pmax=c["pmax"], xmin=c["xmin"], xmax=c["xmax"], ymin=c["ymin"], ymax=c["ymax"], h=c["h"],
                    small_dt=c["small_dt"], density_np=c["density_np"], nb=c["nb"], es=c["es"], src=src)
        assert np.all(np.isfinite(r["f"])) and np.all(np.isfinite(r["fld"])), name
        extra = {"src_" + k: v for k, v in src.items()} if src else {}
        if dilute:
            ys, xs = dilute[0]
            extra["dilute_block"] = np.array([ys.start, ys.stop, xs.start, xs.stop])
            extra["zero_cells"] = np.array(dilute[1])
            zero = np.all(r["fld"][:3] == 0, axis=0)
            assert sorted(zip(*np.nonzero(zero))) == sorted(dilute[1]), (name, np.nonzero(zero))
        save(name, dict(c, **extra), f, fld, r)


def mirror_cases(only):
    """Full steps with mirror boundaries (the harness's restatement of sameNode_mirror_X/Y), whole
    arrays with their 4 guards stored before the first step and after the last."""
    base = dict(pmax=0.25, xmin=0.0, xmax=60.0, ymin=0.0, ymax=40.0, nb=4, h=0.4, small_dt=0.15,
                density_np=1e22, zeta=1.0, nbc=4)
    cases = (
        # x mirror: the l = l0 harmonics are never refreshed (def-Parallel.cpp:1168)
        ("mirror_x_2d", dict(l0=4, m0=2, np=16, nx=6, ny=5, op=1, nsteps=3, bndX=1, bndY=0), 61),
        # y mirror: l = 0 and l = l0 are never refreshed (:1258); conjugate images
        ("mirror_y_2d", dict(l0=3, m0=3, np=12, nx=5, ny=6, op=1, nsteps=3, bndX=0, bndY=1), 62),
        # both, RK4 only
        ("mirror_xy_rk4", dict(l0=3, m0=2, np=12, nx=5, ny=5, op=2, nsteps=4, bndX=1, bndY=1), 63),
        # implicit-E steps (op 9) with RKLevel = 2 guards, as the production deck with bndX = 1
        ("mirror_imp_x", dict(l0=4, m0=2, np=16, nx=6, ny=5, op=9, nsteps=2, es=0, bndX=1, bndY=0, nbc=2), 64),
        ("mirror_imp_x_es", dict(l0=3, m0=3, np=12, nx=5, ny=6, op=9, nsteps=2, es=1, bndX=1, bndY=0, nbc=2), 65),
        # the production deck's whole step: implicit E, x mirror, RKLevel 2, linear laser heating
        # (and inverse bremsstrahlung)
        ("mirror_imp_x_src", dict(l0=3, m0=2, np=48, nx=6, ny=4, op=9, nsteps=2, es=0, bndX=1, bndY=0, nbc=2,
                                  pmax=0.3, xmax=50.0, ymax=30.0,
                                  src=dict(SRC, linear_Ivst=1, polynomial_Ivst=0, inverse_bremsstrahlung=1)), 66),
    )

    def rel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))

    for name, cc, seed in cases:
        if only and name not in only:
            continue
        c = dict(base, **cc)
        src = c.pop("src", None)
        f, fld = implicit_2d_state(c, seed)

        def ref(fin):
            return run_ref(c["l0"], c["m0"], c["np"], c["nx"], c["ny"], fin, fld, op=c["op"], nsteps=c["nsteps"],
                           pmax=c["pmax"], xmin=c["xmin"], xmax=c["xmax"], ymin=c["ymin"], ymax=c["ymax"],
                           h=c["h"], small_dt=c["small_dt"], density_np=c["density_np"], nb=c["nb"],
                           zeta=c["zeta"], es=c.get("es", 0), bnd=(c["bndX"], c["bndY"]), nbc=c["nbc"], src=src)
        r = ref(f)
        for k in ("f", "fld", "f_ext0", "fld_ext0", "f_ext", "fld_ext"):
            assert np.all(np.isfinite(r[k])), (name, k)
        extra = {}
        if c["op"] == 9:
            # the implicit field solve's noise floor, measured on the reference itself: the largest
            # change of its whole arrays when the input f is perturbed by one ulp (3 draws)
            rng = np.random.default_rng(7)
            wf, we = 0.0, np.zeros(6)
            for _ in range(3):
                r1 = ref(f * (1.0 + 2.220446049250313e-16 * rng.choice([-1.0, 1.0], f.shape)))
                wf = max(wf, rel(r1["f_ext"], r["f_ext"]))
                we = np.maximum(we, [rel(r1["fld_ext"][k], r["fld_ext"][k]) if np.any(r["fld_ext"][k]) else 0.0
                                     for k in range(6)])
            extra = dict(floor_f=wf, floor_fld=we)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, f_in=f, fld_in=fld, f_out=r["f"], fld_out=r["fld"], p=r["p"], dx=r["dx"],
                            dy=r["dy"], h=r["h"], t=r["t"], f_ext0=r["f_ext0"], fld_ext0=r["fld_ext0"],
                            f_ext=r["f_ext"], fld_ext=r["fld_ext"], **extra, **{"cfg_" + k: v for k, v in c.items()},
                            **{"cfg_src_" + k: v for k, v in (src or {}).items()})
        print(f"{name}: {os.path.getsize(path)} bytes", extra)


def main(only=()):
    if only:
        source_cases(only)
        implicit_2d_cases(only)
        mirror_cases(only)
        return
    # 1) RHS, 2D complex
    c = dict(l0=4, m0=3, np=16, nx=6, ny=4, pmax=0.3, xmin=-2.0, xmax=3.0, ymin=-1.0, ymax=2.0,
             op=0, nsteps=0, nb=4, h=0.01, small_dt=1.0, density_np=1e21, zeta=1.0)
    f, fld = random_state(c["l0"], c["m0"], c["np"], c["nx"], c["ny"], 1234)
    r = run_ref(c["l0"], c["m0"], c["np"], c["nx"], c["ny"], f, fld, op=0, pmax=c["pmax"],
                xmin=c["xmin"], xmax=c["xmax"], ymin=c["ymin"], ymax=c["ymax"])
    save("rhs_2d_m3", c, f, fld, r)

    # 2) RHS with the y-advection coefficient slip
    c = dict(c, l0=7, m0=3, np=10, nx=5, ny=4)
    f, fld = random_state(c["l0"], c["m0"], c["np"], c["nx"], c["ny"], 99)
    r = run_ref(c["l0"], c["m0"], c["np"], c["nx"], c["ny"], f, fld, op=0, pmax=c["pmax"],
                xmin=c["xmin"], xmax=c["xmax"], ymin=c["ymin"], ymax=c["ymax"])
    save("rhs_quirk_l7", c, f, fld, r)

    # 3) full steps, 1D m0=0 physics through the m0=2 stand-in
    c = dict(l0=8, m0=2, np=32, nx=12, ny=1, pmax=0.25, xmin=0.0, xmax=100.0, ymin=0.0, ymax=1.0,
             op=1, nsteps=3, nb=4, h=0.5, small_dt=0.2, density_np=1e21, zeta=1.0)
    f, fld = maxwellian_state(c["l0"], c["m0"], c["np"], c["nx"], 1, c["pmax"], 7, eyz=False)
    for l in range(c["l0"] + 1):
        for m in range(1, min(l, c["m0"]) + 1):
            f[hind(l, m, c["m0"])] = 0.0
    fld[1:] = 0.0
    r = run_ref(c["l0"], c["m0"], c["np"], c["nx"], 1, f, fld, op=1, nsteps=c["nsteps"],
                pmax=c["pmax"], xmin=c["xmin"], xmax=c["xmax"], ymin=c["ymin"], ymax=c["ymax"],
                h=c["h"], small_dt=c["small_dt"], density_np=c["density_np"], nb=c["nb"])
    # the stand-in's m>=1 harmonics and Ey, Ez, B must stay exactly zero
    for l in range(c["l0"] + 1):
        for m in range(1, min(l, c["m0"]) + 1):
            assert not np.any(r["f"][hind(l, m, c["m0"])]), (l, m)
    assert not np.any(r["fld"][1:])
    save("step_1d_m0", c, f, fld, r, keep_m0_only=True, m0_phys=0)

    # 4) full steps, 2D, l0=m0=4, uniform Bz, e-e and e-i on
    c = dict(l0=4, m0=4, np=12, nx=6, ny=6, pmax=0.25, xmin=0.0, xmax=60.0, ymin=0.0, ymax=60.0,
             op=1, nsteps=2, nb=4, h=0.4, small_dt=0.15, density_np=1e22, zeta=2.0)
    f, fld = maxwellian_state(c["l0"], c["m0"], c["np"], c["nx"], c["ny"], c["pmax"], 11, bz=2e-3)
    r = run_ref(c["l0"], c["m0"], c["np"], c["nx"], c["ny"], f, fld, op=1, nsteps=c["nsteps"],
                pmax=c["pmax"], xmin=c["xmin"], xmax=c["xmax"], ymin=c["ymin"], ymax=c["ymax"],
                h=c["h"], small_dt=c["small_dt"], density_np=c["density_np"], nb=c["nb"],
                zeta=c["zeta"])
    save("step_2d_c3", c, f, fld, r)

    # 5) RK4 only, collisions off
    c = dict(l0=3, m0=2, np=24, nx=10, ny=1, pmax=0.3, xmin=-50.0, xmax=50.0, ymin=0.0, ymax=1.0,
             op=2, nsteps=4, nb=4, h=0.3, small_dt=1.0, density_np=1e-30, zeta=1.0)
    f, fld = maxwellian_state(c["l0"], c["m0"], c["np"], c["nx"], 1, c["pmax"], 5)
    r = run_ref(c["l0"], c["m0"], c["np"], c["nx"], 1, f, fld, op=2, nsteps=c["nsteps"],
                pmax=c["pmax"], xmin=c["xmin"], xmax=c["xmax"], ymin=c["ymin"], ymax=c["ymax"],
                h=c["h"], density_np=c["density_np"])
    save("rk4_1d_nocoll", c, f, fld, r)

    # 6) e-e only and 7) f_lm only, Np=64
    for name, op in (("ee_only", 3), ("flm_only", 4)):
        c = dict(l0=3, m0=2, np=64, nx=4, ny=1, pmax=0.3, xmin=0.0, xmax=10.0, ymin=0.0, ymax=1.0,
                 op=op, nsteps=3, nb=6, h=0.7, small_dt=0.25, density_np=1e23, zeta=1.0)
        f, fld = maxwellian_state(c["l0"], c["m0"], c["np"], c["nx"], 1, c["pmax"], 3 + op)
        r = run_ref(c["l0"], c["m0"], c["np"], c["nx"], 1, f, fld, op=op, nsteps=c["nsteps"],
                    pmax=c["pmax"], xmin=c["xmin"], xmax=c["xmax"], ymin=c["ymin"],
                    ymax=c["ymax"], h=c["h"], small_dt=c["small_dt"],
                    density_np=c["density_np"], nb=c["nb"])
        save(name, c, f, fld, r)

    # 8), 9) implicit-E steps, 1D variant, m0=0 physics through the m0=2 stand-in
    for name, l0, np_, nx, es, seed in (("imp_1d_m0", 5, 32, 12, 1, 21), ("imp_1d_m0_l8", 8, 64, 10, 0, 22)):
        c = dict(l0=l0, m0=2, np=np_, nx=nx, ny=1, pmax=0.25, xmin=0.0, xmax=100.0, ymin=0.0, ymax=1.0,
                 op=5, nsteps=3, nb=4, h=0.5, small_dt=0.2, density_np=1e21, zeta=1.0, es=es)
        f, fld = maxwellian_state(c["l0"], c["m0"], c["np"], c["nx"], 1, c["pmax"], seed, eyz=False)
        for l in range(c["l0"] + 1):
            for m in range(1, min(l, c["m0"]) + 1):
                f[hind(l, m, c["m0"])] = 0.0
        fld[1:] = 0.0
        r = run_ref(c["l0"], c["m0"], c["np"], c["nx"], 1, f, fld, op=5, nsteps=c["nsteps"],
                    pmax=c["pmax"], xmin=c["xmin"], xmax=c["xmax"], ymin=c["ymin"], ymax=c["ymax"],
                    h=c["h"], small_dt=c["small_dt"], density_np=c["density_np"], nb=c["nb"], es=es)
        for l in range(c["l0"] + 1):
            for m in range(1, min(l, c["m0"]) + 1):
                assert not np.any(r["f"][hind(l, m, c["m0"])]), (l, m)
        assert not np.any(r["fld"][1:])
        save(name, c, f, fld, r, keep_m0_only=True, m0_phys=0)

    source_cases(())
    implicit_2d_cases(())
    mirror_cases(())


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
