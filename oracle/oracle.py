"""ctypes front end of the oracle (test infrastructure only).

* :class:`Oracle` wraps liboshun_oracle.so, the CPU restatement of the
  reference's hot path (oshun_oracle.cpp cites each reference file:line).
* :func:`run_ref` runs oracle/_ref/oshun_ref, a driver linked with the
  unmodified reference translation units (built only where the reference
  tree exists; see oracle/Makefile).

Exchange layout for distribution functions: complex128[nh, ny, nx, np]
(p fastest, harmonic order ind(l, m) of def-State.cpp:417-427); fields:
complex128[6, ny, nx] in the order Ex, Ey, Ez, Bx, By, Bz.
"""
import ctypes
import os
import subprocess
import tempfile
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboshun_oracle.so")
REF_BIN = os.path.join(HERE, "_ref", "oshun_ref")
SHIM_BIN = os.path.join(HERE, "_ref", "oshun_ref_shim")

# planes of Oracle.moments(), in the order of the library's oshun_moments
MOMENT_NAMES = ("n", "Vx", "Vy", "VxVx", "VyVy", "VzVz", "Qx", "VxVy", "VxVz", "VyVz", "Vsq", "T_eV", "P_Mbar",
                "ND", "nu", "Qy", "Kx", "Ky")


def num_harmonics(l0, m0):
    """N_h of the (l, m) trapezoid, def-State.cpp:417."""
    return ((m0 + 1) * (2 * l0 - m0 + 2)) // 2


def hind(l, m, m0):
    """Harmonic index ind(l, m), def-State.cpp:422-426."""
    if l < m0 + 1:
        return l * (l + 1) // 2 + m
    return l * (m0 + 1) - m0 * (m0 + 1) // 2 + m


def deck_dx(xmin, xmax, n, guards=4):
    """Cell size exactly as the reference derives it from a deck.

    xglob = Axis(n, xmin, xmax); x(0) = xglob(0) - RKLevel*xglob.dx();
    x(1) = x(0) + xglob.dx(); dx = x(1) - x(0)   (def-Input.cpp:313-325,
    matrices.h:860-880).
    """
    d = (xmax - xmin) / float(n - 1)
    xg0 = 0.0 * d + xmin
    xg1 = 1.0 * d + xmin
    gdx = xg1 - xg0
    x0 = xg0 - guards * gdx
    x1 = x0 + gdx
    return x1 - x0


@dataclass
class Config:
    l0: int
    m0: int
    np: int
    nx: int
    ny: int = 1
    pmax: float = 0.3
    pmin: float = None  # defaults to pmax/(2 np) (def-Input.cpp:79)
    dx: float = 1.0
    dy: float = 1.0
    density_np: float = 1.0e21
    zeta: float = 1.0
    small_dt: float = 1.0
    nb: int = 4
    coll_ee: int = 1
    coll_flm: int = 1
    ref_quirks: int = 1

    def __post_init__(self):
        if self.pmin is None:
            self.pmin = self.pmax / (2.0 * self.np)

    @property
    def nh(self):
        return num_harmonics(self.l0, self.m0)

    def f_shape(self):
        return (self.nh, self.ny, self.nx, self.np)

    def fld_shape(self):
        return (6, self.ny, self.nx)


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        D = ctypes.c_double
        I = ctypes.c_int
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.oo_create.restype = P
        lib.oo_create.argtypes = [ip, dp]
        lib.oo_destroy.argtypes = [P]
        lib.oo_num_harmonics.argtypes = [P]
        lib.oo_num_harmonics.restype = I
        lib.oo_set_state.argtypes = [P, dp, dp]
        lib.oo_get_state.argtypes = [P, dp, dp]
        lib.oo_set_time.argtypes = [P, D]
        lib.oo_time.argtypes = [P]
        lib.oo_time.restype = D
        lib.oo_rhs.argtypes = [P, dp, dp]
        lib.oo_rk4.argtypes = [P, D]
        lib.oo_collide_ee.argtypes = [P, D]
        lib.oo_collide_flm.argtypes = [P, D, I]
        lib.oo_step.argtypes = [P, D, I]
        lib.oo_step_implicit_1d.argtypes = [P, D, I, I]
        lib.oo_step_implicit_1d.restype = I
        lib.oo_step_implicit_2d.argtypes = [P, D, I, I]
        lib.oo_step_implicit_2d.restype = I
        lib.oo_set_es_current.argtypes = [P, dp]
        lib.oo_ee_getslope.argtypes = [P, dp, dp]
        lib.oo_flm_solve.argtypes = [P, dp, D, I, dp]
        lib.oo_flm_coeffs.argtypes = [P, dp, D, dp, dp, dp, dp]
        lib.oo_tridiag.argtypes = [I, dp, dp, dp, dp, dp]
        lib.oo_axes.argtypes = [P, dp, dp, dp]
        lib.oo_moments.argtypes = [P, dp]
        lib.oo_set_sources.argtypes = [P, dp, dp, dp, dp, I]
        lib.oo_laser_source.argtypes = [P, D]
        lib.oo_laser_source.restype = I
        lib.oo_inverse_bremsstrahlung.argtypes = [P, D]
        lib.oo_laser_efield.argtypes = [P, D]
        lib.oo_tlm_1d.argtypes = [P, dp, dp, dp, dp, D, I, dp, dp, dp, dp]
        lib.oo_tlm_1d.restype = I
        _lib = lib
    return _lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """CPU restatement of one OSHUN state plus its RK4/collision drivers."""

    def __init__(self, cfg: Config):
        lib = _load()
        self.cfg = cfg
        iv = (ctypes.c_int * 9)(cfg.l0, cfg.m0, cfg.np, cfg.nx, cfg.ny, cfg.ref_quirks, cfg.nb,
                                cfg.coll_ee, cfg.coll_flm)
        dv = (ctypes.c_double * 7)(cfg.pmin, cfg.pmax, cfg.dx, cfg.dy, cfg.density_np, cfg.zeta,
                                   cfg.small_dt)
        self._h = lib.oo_create(iv, dv)
        if not self._h:
            raise ValueError(f"oracle rejected config {cfg}")
        self._lib = lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.oo_destroy(h)
            self._h = None

    def set_state(self, f, fld):
        f = np.ascontiguousarray(f, dtype=np.complex128).reshape(self.cfg.f_shape())
        fld = np.ascontiguousarray(fld, dtype=np.complex128).reshape(self.cfg.fld_shape())
        self._lib.oo_set_state(self._h, _dptr(f.view(np.float64)), _dptr(fld.view(np.float64)))

    def get_state(self):
        f = np.zeros(self.cfg.f_shape(), np.complex128)
        fld = np.zeros(self.cfg.fld_shape(), np.complex128)
        self._lib.oo_get_state(self._h, _dptr(f.view(np.float64)), _dptr(fld.view(np.float64)))
        return f, fld

    def set_time(self, t):
        self._lib.oo_set_time(self._h, float(t))

    def time(self):
        return self._lib.oo_time(self._h)

    def rhs(self):
        f = np.zeros(self.cfg.f_shape(), np.complex128)
        fld = np.zeros(self.cfg.fld_shape(), np.complex128)
        self._lib.oo_rhs(self._h, _dptr(f.view(np.float64)), _dptr(fld.view(np.float64)))
        return f, fld

    def rk4(self, h):
        self._lib.oo_rk4(self._h, float(h))

    def collide_ee(self, tnew):
        self._lib.oo_collide_ee(self._h, float(tnew))

    def collide_flm(self, tnew, which):
        self._lib.oo_collide_flm(self._h, float(tnew), int(which))

    def step(self, h, nsteps=1):
        self._lib.oo_step(self._h, float(h), int(nsteps))

    def step_implicit_1d(self, h, nsteps=1, es=0):
        """main.cpp:79-107 with implicit_E and if_implicit1D (RK2, implicit Ex); l0 >= 3."""
        if self._lib.oo_step_implicit_1d(self._h, float(h), int(nsteps), int(es)) != 0:
            raise ValueError("the 1D implicit-E step needs l0 >= 3")

    def step_implicit_2d(self, h, nsteps=1, es=0):
        """main.cpp:79-107 with implicit_E and not if_implicit1D (RK2, 3x3 implicit field
        solve); l0 >= 3, m0 >= 2.  Returns the largest number of zero-determinant cells
        left in a step; raises where the reference exits (more than 8)."""
        z = self._lib.oo_step_implicit_2d(self._h, float(h), int(nsteps), int(es))
        if z < 0:
            raise ValueError("the 2D implicit-E step needs l0 >= 3 and m0 >= 2")
        if z > 8:
            raise FloatingPointError(f"implicit field: zero determinant in {z} cells")
        return z

    def set_es_current(self, jn):
        """JN of the electrostatic 2D implicit field: complex [3][ny][nx]."""
        jn = np.ascontiguousarray(jn, np.complex128).reshape(3, self.cfg.ny, self.cfg.nx)
        self._lib.oo_set_es_current(self._h, _dptr(jn.view(np.float64)))

    def ee_getslope(self, f00):
        f00 = np.ascontiguousarray(f00, np.float64)
        out = np.zeros_like(f00)
        self._lib.oo_ee_getslope(self._h, _dptr(f00), _dptr(out))
        return out

    def flm_solve(self, f00, dt, el, col):
        f00 = np.ascontiguousarray(f00, np.float64)
        col = np.ascontiguousarray(col, np.complex128).copy()
        self._lib.oo_flm_solve(self._h, _dptr(f00), float(dt), int(el), _dptr(col.view(np.float64)))
        return col

    def flm_coeffs(self, f00, dt):
        f00 = np.ascontiguousarray(f00, np.float64)
        n = self.cfg.np
        out = [np.zeros(n) for _ in range(4)]
        self._lib.oo_flm_coeffs(self._h, _dptr(f00), float(dt), *[_dptr(o) for o in out])
        return out  # sub, diag, sup, scatt

    def set_sources(self, src, x, y, xglob):
        """Source parameters (SOURCE_KEYS) and the positions of the interior cells (x, y) and of the
        global x axis (def-PLSource.cpp constructors); the source clocks start at the current time."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        xg = np.ascontiguousarray(xglob, np.float64)
        sv = source_vector(src)
        self._lib.oo_set_sources(self._h, _dptr(sv), _dptr(x), _dptr(y), _dptr(xg), int(xg.size))

    def laser_source(self, tnew):
        """PLaserSource::Laser_Source (def-PLSource.cpp:138-252); False when a heating rate alpha
        leaves [0, 0.99] (the reference exits there)."""
        return self._lib.oo_laser_source(self._h, float(tnew)) == 0

    def inverse_bremsstrahlung(self, tnew):
        """InverseBremsstrahlung::loop (def-PLSource.cpp:780-861)."""
        self._lib.oo_inverse_bremsstrahlung(self._h, float(tnew))

    def laser_efield(self, tnew):
        """ExternalEfield::Laser_Efield (def-PLSource.cpp:368-418)."""
        self._lib.oo_laser_efield(self._h, float(tnew))

    def moments(self):
        """The 18 moments of MOMENT_NAMES per (y, x) (def-Parallel.cpp:1677-2253, FP64)."""
        out = np.zeros((len(MOMENT_NAMES), self.cfg.ny, self.cfg.nx))
        self._lib.oo_moments(self._h, _dptr(out))
        return out

    def axes(self):
        n = self.cfg.np
        pd, pc, vr = np.zeros(n), np.zeros(n), np.zeros(n)
        self._lib.oo_axes(self._h, _dptr(pd), _dptr(pc), _dptr(vr))
        return pd, pc, vr


    def tlm_1d(self, f, fld, df=None, dfld=None, h=0.1, nsteps=1):
        """Tangent-linear model of nsteps steps from t = 0 of a 1D m0 = 0 run
        (Step1D in oshun_oracle.cpp).  f, df: real [l0+1, nx, np]; fld, dfld: real
        [6, nx].  Returns (f_N, fld_N) and, when df is given, (df_N, dfld_N) too."""
        c = self.cfg
        nf, ng = (c.l0 + 1, c.nx, c.np), (6, c.nx)
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(nf)
        fld = np.ascontiguousarray(fld, dtype=np.float64).reshape(ng)
        fo, go = np.zeros(nf), np.zeros(ng)
        if df is None:
            rc = self._lib.oo_tlm_1d(self._h, _dptr(f), _dptr(fld), None, None, float(h), int(nsteps),
                                     _dptr(fo), _dptr(go), None, None)
            if rc:
                raise ValueError("tlm_1d: 1D runs with m0 = 0 only")
            return fo, go
        df = np.ascontiguousarray(df, dtype=np.float64).reshape(nf)
        dfld = np.ascontiguousarray(dfld, dtype=np.float64).reshape(ng)
        dfo, dgo = np.zeros(nf), np.zeros(ng)
        rc = self._lib.oo_tlm_1d(self._h, _dptr(f), _dptr(fld), _dptr(df), _dptr(dfld), float(h), int(nsteps),
                                 _dptr(fo), _dptr(go), _dptr(dfo), _dptr(dgo))
        if rc:
            raise ValueError("tlm_1d: 1D runs with m0 = 0 only")
        return fo, go, dfo, dgo


def tridiag(a, b, c, d):
    """Thomas algorithm of def-Nmethods.cpp:142-167 (a[0], c[-1] unused)."""
    lib = _load()
    n = len(b)
    a, b, c, d = (np.ascontiguousarray(v, np.float64) for v in (a, b, c, d))
    x = np.zeros(n)
    lib.oo_tridiag(n, _dptr(a), _dptr(b), _dptr(c), _dptr(d), _dptr(x))
    return x


def have_ref():
    return os.path.exists(REF_BIN)


# the deck's source parameters, in the order of the driver's source block and of oo_set_sources
SOURCE_KEYS = ("linear_Ivst", "polynomial_Ivst", "inverse_bremsstrahlung", "polarization_direction", "I_0",
               "lambda_0", "ab_eta", "ab_depth", "cntr_x0", "cntr_r0", "spot_w0", "rise_time", "flat_time",
               "fall_time", "smaller_dt")


def source_vector(src):
    return np.array([float(src.get(k, 0.0)) for k in SOURCE_KEYS], np.float64)


def deck_axis(n, vmin, vmax, guards=4, whole=False):
    """Interior cell positions of the reference's x (or y) axis and its global axis, computed as
    Inputdata::decl_input does (def-Input.cpp:313-325, Axis of matrices.h:860-866): the global axis
    i*d + min, then the local one from its first cell minus the guards, cumulatively by glob.dx().
    whole: the positions of all n + 2 guards cells instead of the interior's."""
    d = (vmax - vmin) / float(n - 1)
    glob = np.arange(n) * d + vmin
    gdx = glob[1] - glob[0]
    loc = np.empty(n + 2 * guards)
    loc[0] = glob[0] - guards * gdx
    for i in range(1, loc.size):
        loc[i] = loc[i - 1] + gdx
    return (loc.copy() if whole else loc[guards:guards + n].copy()), glob


def write_ref_input(path, l0, m0, np_, nx, ny, f, fld, op, nsteps=1, nb=4, pmax=0.3, xmin=-1.0, xmax=1.0,
                    ymin=-1.0, ymax=1.0, h=0.01, small_dt=1.0, density_np=1e21, zeta=1.0, es=0, src=None,
                    bnd=(0, 0), nbc=4):
    """The driver's input file (layout in ref_harness.cpp's header comment)."""
    nh = num_harmonics(l0, m0)
    f = np.ascontiguousarray(f, np.complex128).reshape(nh, ny, nx, np_)
    fld = np.ascontiguousarray(fld, np.complex128).reshape(6, ny, nx)
    with open(path, "wb") as fh:
        np.array([l0, m0, np_, nx, ny, op, nsteps, nb, es, 1 if src is not None else 0, int(bnd[0]), int(bnd[1]),
                  int(nbc)], np.int32).tofile(fh)
        np.array([pmax, xmin, xmax, ymin, ymax, h, small_dt, density_np, zeta], np.float64).tofile(fh)
        if src is not None:
            source_vector(src).tofile(fh)
        f.tofile(fh)
        fld.tofile(fh)


def run_ref(l0, m0, np_, nx, ny, f, fld, op, nsteps=1, nb=4, pmax=0.3, xmin=-1.0, xmax=1.0,
            ymin=-1.0, ymax=1.0, h=0.01, small_dt=1.0, density_np=1e21, zeta=1.0, es=0, binary=None, src=None,
            bnd=(0, 0), nbc=4):
    """Run the reference driver; returns dict(dx, dy, h, t, p, f, fld).
    op 5 and 9 (implicit-E step, 1D and 2D variant) take es = if_implicitES.  binary: another build of
    the driver (SHIM_BIN: the same driver with the GPU shim linked in).  bnd = (bndX, bndY): with a
    mirror boundary the result also holds the whole arrays with their nbc guards, of the initial
    state after its guard fill (f_ext0, fld_ext0) and of the final state (f_ext, fld_ext)."""
    binary = binary or REF_BIN
    if not os.path.exists(binary):
        raise RuntimeError(f"driver not built: {binary} (make -C oracle ref / shim)")
    nh = num_harmonics(l0, m0)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        write_ref_input(fin, l0, m0, np_, nx, ny, f, fld, op, nsteps, nb, pmax, xmin, xmax, ymin, ymax, h,
                        small_dt, density_np, zeta, es, src, bnd, nbc)
        r = subprocess.run([binary, fin, fout], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{os.path.basename(binary)} failed ({r.returncode}): {r.stdout[-2000:]}{r.stderr[-2000:]}")
        raw = np.fromfile(fout, np.float64)
    hdr = raw[:4]
    p = raw[4:4 + np_]
    rest = raw[4 + np_:].view(np.complex128)
    nf = nh * ny * nx * np_
    out = dict(dx=hdr[0], dy=hdr[1], h=hdr[2], t=hdr[3], p=p,
               f=rest[:nf].reshape(nh, ny, nx, np_).copy(),
               fld=rest[nf:nf + 6 * ny * nx].reshape(6, ny, nx).copy())
    if bnd[0] or bnd[1]:
        nxg, nyg = nx + 2 * nbc, ny + 2 * nbc
        ne, nfe = nh * nyg * nxg * np_, 6 * nyg * nxg
        ext = rest[nf + 6 * ny * nx:]
        for k, tag in enumerate(("0", "")):
            blk = ext[k * (ne + nfe):(k + 1) * (ne + nfe)]
            out["f_ext" + tag] = blk[:ne].reshape(nh, nyg, nxg, np_).copy()
            out["fld_ext" + tag] = blk[ne:].reshape(6, nyg, nxg).copy()
    return out
