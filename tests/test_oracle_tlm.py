"""The oracle's tangent-linear model (Step1D<R> in oracle/oshun_oracle.cpp),
the FP64 reference the GPU adjoint is checked against.  Pinned here on the CPU:
its double instantiation reproduces the oracle's own step (itself pinned to the
compiled reference), and its dual-number tangent matches central differences
of the oracle's step.
"""
import numpy as np

from oracle.oracle import Config, Oracle


def _state(l0, np_, nx, pmax, seed):
    rng = np.random.default_rng(seed)
    pmin = pmax / (2.0 * np_)
    p = np.arange(np_) * ((pmax - pmin) / (np_ - 1)) + pmin
    x = np.arange(nx) / nx
    pt2 = (1.0 + 0.1 * np.cos(2 * np.pi * x)) / 511.0
    f = np.zeros((l0 + 1, nx, np_))
    f[0] = np.exp(-p[None, :] ** 2 / (2 * pt2[:, None])) / (np.sqrt(2 * np.pi) * 2 * np.pi * pt2[:, None] ** 1.5)
    f[1:] = 1e-2 * rng.uniform(-1, 1, (l0, nx, np_)) * f[0][None]
    fld = np.zeros((6, nx))
    fld[0] = 2e-3 * np.sin(2 * np.pi * x)
    fld[1] = 1e-3 * np.cos(2 * np.pi * x)
    fld[5] = 5e-4 * np.sin(4 * np.pi * x)
    return f, fld


CFG = dict(l0=3, m0=0, np=32, nx=16, pmax=0.3, dx=0.4, density_np=1e22, small_dt=0.1, nb=4)
H, STEPS = 0.2, 3


def test_tlm_forward_matches_the_oracle_step():
    cfg = Config(**CFG)
    o = Oracle(cfg)
    f, fld = _state(cfg.l0, cfg.np, cfg.nx, cfg.pmax, 3)
    o.set_state(f[:, None].astype(np.complex128), fld[:, None].astype(np.complex128))
    o.step(H, STEPS)
    fo, go = o.get_state()
    ft, gt = o.tlm_1d(f, fld, h=H, nsteps=STEPS)
    for a, b in ((ft, fo[:, 0].real), (gt, go[:, 0].real)):
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def test_tlm_tangent_matches_central_differences():
    cfg = Config(**CFG)
    o = Oracle(cfg)
    f, fld = _state(cfg.l0, cfg.np, cfg.nx, cfg.pmax, 4)
    rng = np.random.default_rng(5)
    df = rng.standard_normal(f.shape) * np.abs(f[0])[None] * np.array([1.0, 1e-2, 1e-2, 1e-2])[:, None, None]
    dg = rng.standard_normal(fld.shape) * 1e-3
    _, _, dfo, dgo = o.tlm_1d(f, fld, df, dg, h=H, nsteps=STEPS)
    eps = 1e-6
    fp, gp = o.tlm_1d(f + eps * df, fld + eps * dg, h=H, nsteps=STEPS)
    fm, gm = o.tlm_1d(f - eps * df, fld - eps * dg, h=H, nsteps=STEPS)
    for t, fd in ((dfo, (fp - fm) / (2 * eps)), (dgo, (gp - gm) / (2 * eps))):
        scale = np.abs(fd).max()
        assert np.abs(t - fd).max() <= 1e-6 * scale
