"""Checks of the oracle that do not need the reference: closed forms and invariants.

Each is chosen so that a plausible slip (wrong sign, dropped term, wrong index,
transposed operand) fails it.
"""
import numpy as np

from oracle.oracle import Config, Oracle, hind, tridiag


def test_free_streaming_closed_form():
    """f00 = a(p) sin(kx), everything else 0: the only slope is
    d f10/dt = v(p) (f00(x-dx) - f00(x+dx)) / (2 dx)   (continuum: -v df00/dx).
    Pins the x-advection sign, the (l-m+1)/(2l+1) factor at l=0 and v = p/gamma."""
    nx, np_, dx = 16, 12, 0.7
    o = Oracle(Config(l0=3, m0=0, np=np_, nx=nx, dx=dx, pmax=0.4))
    pd, pc, _ = o.axes()
    k = 2 * np.pi / (nx * dx)
    x = np.arange(nx) * dx
    a = 1.0 + pd ** 2
    f = np.zeros((4, 1, nx, np_), np.complex128)
    f[0, 0] = np.sin(k * x)[:, None] * a[None, :]
    o.set_state(f, np.zeros((6, 1, nx)))
    s, fld = o.rhs()
    v = pc / np.sqrt(1 + pc ** 2)
    expect = v[None, :] * a[None, :] * (-2.0 * np.cos(k * x) * np.sin(k * dx))[:, None] / (2 * dx)
    assert np.allclose(s[1, 0].real, expect, rtol=1e-12, atol=1e-14)
    assert np.abs(s[[0, 2, 3]]).max() == 0.0
    assert np.abs(fld).max() == 0.0


def test_electric_acceleration_closed_form():
    """Uniform Ex = E0, f00 = c0 + c2 p^2, rest 0:
    d f10/dt = E0 * (f(p+dp) - f(p-dp)) / (2 dp) = 2 c2 p E0 in the interior,
    and exactly 2 c2 p0 E0 at p0 (the G00 extrapolation is exact on quadratics)."""
    nx, np_ = 4, 20
    o = Oracle(Config(l0=2, m0=0, np=np_, nx=nx, pmax=0.3))
    pd, pc, _ = o.axes()
    E0, c0, c2 = 3e-3, 2.0, -5.0
    f = np.zeros((3, 1, nx, np_), np.complex128)
    f[0, 0] = (c0 + c2 * pc ** 2)[None, :]
    fld = np.zeros((6, 1, nx), np.complex128)
    fld[0] = E0
    o.set_state(f, fld)
    s, _ = o.rhs()
    got = s[1, 0, 0].real
    assert np.allclose(got[:-1], 2 * c2 * pc[:-1] * E0, rtol=1e-10, atol=0)
    # pmax closure: one-sided difference f(N-1)-f(N-2) over 2 dp
    dp = pc[1] - pc[0]
    assert np.isclose(got[-1], E0 * c2 * (pc[-1] ** 2 - pc[-2] ** 2) / (2 * dp), rtol=1e-10)
    # f00 feeds only f10 (through G00); f20 gets nothing from f00
    assert np.abs(s[2]).max() == 0.0


def test_thomas_vs_dense():
    rng = np.random.default_rng(0)
    n = 57
    a, c = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    b = 3.0 + rng.uniform(0, 1, n)
    d = rng.uniform(-1, 1, n)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    x = tridiag(a, b, c, d)
    assert np.allclose(A @ x, d, rtol=0, atol=1e-13)


def _maxwellian(p, T=1.0, n=1.0):
    pt = np.sqrt(T / 511.0)
    return n / (np.sqrt(2 * np.pi) * 2 * np.pi * pt ** 3) * np.exp(-p ** 2 / (2 * pt ** 2))


def test_ee_conserves_number_and_energy():
    """Bell's scheme is conservative in flux form: the slope integrates to zero
    against the cell volumes v^2 dv up to the boundary flux at p0, which is
    O(v0^3) (def-FokkerPlanck.cpp:247-260), and nearly so against v^4 dv."""
    np_ = 96
    o = Oracle(Config(l0=1, m0=0, np=np_, nx=3, pmax=0.3, density_np=1e22, nb=4))
    pd, _, v = o.axes()
    f = _maxwellian(pd, 1.0) + 0.3 * _maxwellian(pd - 0.05, 0.3)   # non-Maxwellian
    fh = o.ee_getslope(f)
    vol = np.zeros(np_)
    vol[0] = v[0] ** 2 * v[1] / 2
    vol[1:-1] = v[1:-1] ** 2 * (v[2:] - v[:-2]) / 2
    scale = np.abs(fh * vol).sum()
    assert abs((fh * vol).sum()) < 1e-3 * scale
    assert abs((fh * vol * v ** 2).sum()) < 1e-2 * np.abs(fh * vol * v ** 2).sum()
    # a Maxwellian is (nearly) stationary, the bump-on-tail is not
    fm = _maxwellian(pd, 1.0)
    fhm = o.ee_getslope(fm)
    assert np.abs(fhm[:np_ // 2]).max() < 0.05 * np.abs(fh[:np_ // 2]).max()


def test_flm_solve_equals_dense_backward_euler():
    """The implicit l>=1 step solves (I + Dt*C_l) f_new = f_old with the
    tridiagonal C_l assembled from f00; check against numpy's dense solve and
    that only the diagonal depends on l (through -l(l+1)/2 * scattering)."""
    np_ = 40
    o = Oracle(Config(l0=4, m0=0, np=np_, nx=3, pmax=0.3, density_np=1e23))
    pd, _, _ = o.axes()
    f00 = _maxwellian(pd, 0.7)
    rng = np.random.default_rng(3)
    col = rng.uniform(-1, 1, np_) + 0j
    dt = 5.0
    sub, diag, sup, sc = o.flm_coeffs(f00, dt)
    for el in (1, 2, 4):
        b = diag.copy()
        if el > 1:
            b[0] = 0.0
        b += 1.0 + 0.5 * el * (el + 1) * sc
        A = np.diag(b) + np.diag(sub[1:], -1) + np.diag(sup[:-1], 1)
        want = np.linalg.solve(A, col.real)
        got = o.flm_solve(f00, dt, el, col)
        assert np.allclose(got.real, want, rtol=1e-11, atol=1e-14)
    # pure e-i scattering damps higher harmonics more: |f_l| decreases with l
    n1 = np.abs(o.flm_solve(f00, dt, 1, col)).sum()
    n4 = np.abs(o.flm_solve(f00, dt, 4, col)).sum()
    assert n4 < n1 < np.abs(col).sum()


def test_density_conserved_by_full_step():
    """Periodic advection + e-e + implicit f_lm keep the total f00 number."""
    nx, np_ = 8, 48
    o = Oracle(Config(l0=3, m0=0, np=np_, nx=nx, pmax=0.3, dx=2.0, density_np=1e22, small_dt=0.3))
    pd, _, v = o.axes()
    x = np.arange(nx)
    f = np.zeros((4, 1, nx, np_), np.complex128)
    f[0, 0] = (1 + 0.2 * np.cos(2 * np.pi * x / nx))[:, None] * _maxwellian(pd)[None, :]
    f[1, 0] = 1e-3 * np.sin(2 * np.pi * x / nx)[:, None] * _maxwellian(pd)[None, :]
    o.set_state(f, np.zeros((6, 1, nx)))
    vol = np.zeros(np_)
    vol[0] = v[0] ** 2 * v[1] / 2
    vol[1:-1] = v[1:-1] ** 2 * (v[2:] - v[:-2]) / 2
    n0 = (f[0].real.sum(axis=(0, 1)) * vol).sum()
    o.step(0.5, 4)
    g, _ = o.get_state()
    n1 = (g[0].real.sum(axis=(0, 1)) * vol).sum()
    # advection moves f00 in x only; e-e conserves its own discrete number.
    # The p-end closures leak a little; the bound is loose but catches sign slips.
    assert abs(n1 - n0) < 1e-6 * abs(n0)


def test_moments_closed_forms():
    """FP64 moments (def-Parallel.cpp:1677-2134) on a drifting Maxwellian:
    n = 1, <v_x^2> = T/m up to the p/gamma correction, Vx = u, and Qx = 0 at rest."""
    np_, nx = 256, 4
    o = Oracle(Config(l0=2, m0=0, np=np_, nx=nx, pmax=0.3))
    pd, _, _ = o.axes()
    pt = np.sqrt(1.0 / 511.0)
    f = np.zeros((3, 1, nx, np_), np.complex128)
    f[0, 0] = _maxwellian(pd)[None, :]
    o.set_state(f, np.zeros((6, 1, nx)))
    m = o.moments()
    assert abs(m[0].mean() - 1.0) < 1e-5
    assert abs(m[3].mean() / pt ** 2 - 1.0) < 0.02          # relativistic 1/gamma^2 weight
    assert np.allclose(m[3], m[4]) and np.allclose(m[4], m[5])  # isotropic
    assert np.abs(m[1]).max() == 0.0 and np.abs(m[6]).max() == 0.0
    u = 0.01 * pt
    f[1, 0] = u * pd / pt ** 2 * f[0, 0].real                 # first harmonic of a slow drift
    o.set_state(f, np.zeros((6, 1, nx)))
    m = o.moments()
    assert abs(m[1].mean() / u - 1.0) < 0.01
