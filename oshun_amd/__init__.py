"""MI355X-native distribution-function update of OSHUN (Vlasov-Maxwell RK4 + Fokker-Planck).

Python host mirror over the C ABI of ``liboshun_hip.so`` (declared in
include/oshun_hip.h).  The classes keep the reference's names and call
pattern (OSHUN-1, decl-RK.h:48-79, decl-fokkerplanck.h:140-163,289-311):

    ctx = Context(grid, phys)           # device state (the reference's Stat Y)
    rk = Runge_Kutta_4(ctx)             # rk.advance_p1()     -> t += h
    ee = Explicit_EC_FP(ctx)            # ee.loop(rk.time())
    eb = Euler_Backward(ctx)            # eb.advance_1(t); eb.advance_lm(t)

or ``ctx.step(h, nsteps)`` for the whole step body of main.cpp:79-107.

Host arrays use the reference layout: complex128[nh, ny(+2nbc), nx(+2nbc), np]
for f (p fastest, harmonic order ind(l, m)) and complex128[6, ny, nx] for
Ex, Ey, Ez, Bx, By, Bz.  The library is required: nothing here computes on the
CPU, and a missing or unloadable library raises.
"""
import ctypes
import os

import numpy as np

__all__ = ["Grid", "Phys", "Context", "Batch", "Runge_Kutta_4", "Explicit_EC_FP", "Euler_Backward", "num_harmonics",
           "hind", "deck_pmin", "FP64", "FP32_SPLIT", "COLL_EE", "COLL_FLM", "library_path", "OshunError"]

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "liboshun_hip.so"

FP64 = 0
FP32_SPLIT = 1
COLL_EE = 1
COLL_FLM = 2
# planes of Context.moments() (oshun_moments): def-Parallel.cpp:1677-2253 in FP64
MOMENT_NAMES = ("n", "Vx", "Vy", "VxVx", "VyVy", "VzVz", "Qx", "VxVy", "VxVz", "VyVz", "Vsq", "T_eV", "P_Mbar",
                "ND", "nu", "Qy", "Kx", "Ky")


class OshunError(RuntimeError):
    pass


def library_path():
    return os.path.join(HERE, _LIB_NAME)


def num_harmonics(l0, m0):
    """N_h of the (l, m) trapezoid, def-State.cpp:417."""
    return ((m0 + 1) * (2 * l0 - m0 + 2)) // 2


def hind(l, m, m0):
    """ind(l, m) of def-State.cpp:422-426."""
    if l < m0 + 1:
        return l * (l + 1) // 2 + m
    return l * (m0 + 1) - m0 * (m0 + 1) // 2 + m


def deck_pmin(pmax, nump):
    """The reference decks' pmin = pmax / (2 nump) (def-Input.cpp:79)."""
    return pmax / (2.0 * nump)


class Grid(ctypes.Structure):
    _fields_ = [("l0", ctypes.c_int), ("m0", ctypes.c_int), ("nump", ctypes.c_int), ("numx", ctypes.c_int),
                ("numy", ctypes.c_int), ("pmin", ctypes.c_double), ("pmax", ctypes.c_double),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("precision", ctypes.c_int),
                ("bndx", ctypes.c_int), ("bndy", ctypes.c_int), ("nbc", ctypes.c_int)]


class Phys(ctypes.Structure):
    _fields_ = [("density_np", ctypes.c_double), ("zeta", ctypes.c_double), ("small_dt", ctypes.c_double),
                ("nb_algorithms", ctypes.c_int), ("collisions", ctypes.c_int), ("ref_quirks", ctypes.c_int)]


class Sources(ctypes.Structure):
    """oshun_sources: the deck's laser parameters (def-PLSource.cpp)."""
    _fields_ = [("linear_Ivst", ctypes.c_int), ("polynomial_Ivst", ctypes.c_int),
                ("inverse_bremsstrahlung", ctypes.c_int), ("polarization_direction", ctypes.c_int),
                ("I_0", ctypes.c_double), ("lambda_0", ctypes.c_double), ("ab_eta", ctypes.c_double),
                ("ab_depth", ctypes.c_double), ("cntr_x0", ctypes.c_double), ("cntr_r0", ctypes.c_double),
                ("spot_w0", ctypes.c_double), ("rise_time", ctypes.c_double), ("flat_time", ctypes.c_double),
                ("fall_time", ctypes.c_double), ("smaller_dt", ctypes.c_double)]


_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise OshunError(f"{path} is missing: build it with `python -m oshun_amd._build` (hipcc, gfx950)")
    # torch ships a HIP runtime of its own; imported after this library's runtime
    # (the system one) has initialised the GPU, it finds no GPU at all.  Imported
    # first, both runtimes work in one process, in either order of first use.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    D = ctypes.c_double
    I = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    lib.oshun_create.argtypes = [ctypes.POINTER(Grid), ctypes.POINTER(Phys), ctypes.POINTER(P)]
    lib.oshun_destroy.argtypes = [P]
    lib.oshun_last_error.argtypes = [P]
    lib.oshun_last_error.restype = ctypes.c_char_p
    lib.oshun_num_harmonics.argtypes = [P]
    lib.oshun_upload_state.argtypes = [P, dp, dp, I]
    lib.oshun_download_state.argtypes = [P, dp, dp, I]
    lib.oshun_upload_harmonics.argtypes = [P, ctypes.POINTER(dp), ctypes.POINTER(dp), I]
    lib.oshun_download_harmonics.argtypes = [P, ctypes.POINTER(dp), ctypes.POINTER(dp), I]
    lib.oshun_set_time.argtypes = [P, D]
    lib.oshun_get_time.argtypes = [P, dp]
    lib.oshun_rhs.argtypes = [P, dp, dp]
    lib.oshun_rk4_step.argtypes = [P, D]
    lib.oshun_collide_ee.argtypes = [P, D]
    lib.oshun_collide_flm.argtypes = [P, D, I]
    lib.oshun_step.argtypes = [P, D, I]
    lib.oshun_step_implicit_1d.argtypes = [P, D, I, I]
    lib.oshun_step_implicit_2d.argtypes = [P, D, I, I]
    lib.oshun_set_implicit_es_current.argtypes = [P, dp, I]
    lib.oshun_implicit_zero_dets.argtypes = [P]
    lib.oshun_fill_guards.argtypes = [P]
    lib.oshun_synchronize.argtypes = [P]
    lib.oshun_get_stream.argtypes = [P, ctypes.POINTER(P)]
    lib.oshun_device_state.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(P)]
    lib.oshun_plane_info.argtypes = [P, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_longlong),
                                     ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I)]
    lib.oshun_profile_enable.argtypes = [P, I]
    lib.oshun_profile_read.argtypes = [P, dp, ctypes.POINTER(ctypes.c_longlong), I]
    lib.oshun_diag_copy.argtypes = [P, ctypes.c_longlong]
    lib.oshun_stage_kernel.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.oshun_pair_rows.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.oshun_moments.argtypes = [P, dp]
    lib.oshun_set_sources.argtypes = [P, ctypes.POINTER(Sources), dp, dp, dp, I]
    lib.oshun_laser_source.argtypes = [P, D]
    lib.oshun_inverse_bremsstrahlung.argtypes = [P, D]
    lib.oshun_laser_efield.argtypes = [P, D]
    lib.oshun_moments_noalloc.argtypes = [P, P]
    lib.oshun_batch_create.argtypes = [ctypes.POINTER(Grid), ctypes.POINTER(Phys), dp, I, ctypes.POINTER(P)]
    lib.oshun_batch_destroy.argtypes = [P]
    lib.oshun_batch_last_error.argtypes = [P]
    lib.oshun_batch_last_error.restype = ctypes.c_char_p
    lib.oshun_batch_size.argtypes = [P]
    lib.oshun_batch_run.argtypes = [P, I]
    lib.oshun_batch_run.restype = P
    lib.oshun_batch_step.argtypes = [P, I]
    lib.oshun_batch_synchronize.argtypes = [P]
    lib.oshun_batch_autotune.argtypes = [P, I, ctypes.POINTER(I), dp]
    lib.oshun_batch_pair_rows.argtypes = [P, I, ctypes.POINTER(I)]
    lib.oshun_copy_device_state.argtypes = [P, P, P, I]
    lib.oshun_adjoint_step.argtypes = [P, D, P, P]
    lib.oshun_adjoint_forward.argtypes = [P, D, I, I, ctypes.POINTER(ctypes.c_longlong)]
    lib.oshun_adjoint_backward.argtypes = [P, ctypes.c_longlong, P, P]
    lib.oshun_adjoint_release.argtypes = [P]
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    """One device-resident OSHUN state (the reference's ``Stat Y`` plus its operators)."""

    def __init__(self, l0, m0, nump, numx, numy=1, pmax=0.3, pmin=None, dx=1.0, dy=1.0, precision=FP64,
                 density_np=1e21, zeta=1.0, small_dt=1.0, nb_algorithms=4, collisions=COLL_EE | COLL_FLM,
                 ref_quirks=1, bndx=0, bndy=0, nbc=0):
        """bndx, bndy: 0 periodic, 1 mirror; with a mirror the device keeps nbc guard cells per side
        (the deck's RKLevel) and upload/download take nbc = that number (whole arrays) or 0."""
        lib = _load()
        if pmin is None:
            pmin = deck_pmin(pmax, nump)
        self.grid = Grid(l0, m0, nump, numx, numy, pmin, pmax, dx, dy, precision, bndx, bndy, nbc)
        self.phys = Phys(density_np, zeta, small_dt, nb_algorithms, collisions, ref_quirks)
        h = ctypes.c_void_p()
        rc = lib.oshun_create(ctypes.byref(self.grid), ctypes.byref(self.phys), ctypes.byref(h))
        if rc != 0:
            raise OshunError(f"oshun_create failed ({rc}) for {l0=} {m0=} {nump=} {numx=} {numy=} {bndx=} {bndy=} {nbc=}")
        self._h = h
        self._lib = lib
        self._owned = True
        self.l0, self.m0, self.np, self.nx, self.ny = l0, m0, nump, numx, numy
        self.nh = num_harmonics(l0, m0)
        self.precision = precision

    @classmethod
    def _view(cls, lib, handle, grid, phys, owner):
        """A non-owning view of a context that another object (a Batch) owns."""
        self = cls.__new__(cls)
        self._h = ctypes.c_void_p(handle)
        self._lib = lib
        self._owned = False
        self._owner = owner
        self.grid, self.phys = grid, phys
        self.l0, self.m0, self.np, self.nx, self.ny = grid.l0, grid.m0, grid.nump, grid.numx, grid.numy
        self.nh = num_harmonics(grid.l0, grid.m0)
        self.precision = grid.precision
        return self

    def _chk(self, rc):
        if rc != 0:
            msg = self._lib.oshun_last_error(self._h)
            raise OshunError(f"oshun call failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                self._lib.oshun_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ------------------------------------------------------------
    def f_shape(self, nbc=0):
        return (self.nh, self.ny + 2 * nbc, self.nx + 2 * nbc, self.np)

    def fld_shape(self, nbc=0):
        return (6, self.ny + 2 * nbc, self.nx + 2 * nbc)

    def device_shape(self):
        """Shape of the device-state planes: (ncomp, ny, np, nx), x fastest."""
        return (self.plane_info()["ncomp"], self.ny, self.np, self.nx)

    def fld_device_shape(self):
        return (6, self.ny, self.nx)

    def upload(self, f, fld, nbc=0):
        f = np.ascontiguousarray(f, np.complex128).reshape(self.f_shape(nbc))
        fld = np.ascontiguousarray(fld, np.complex128).reshape(self.fld_shape(nbc))
        self._chk(self._lib.oshun_upload_state(self._h, _dptr(f.view(np.float64)), _dptr(fld.view(np.float64)),
                                               nbc))

    def download(self, nbc=0):
        f = np.zeros(self.f_shape(nbc), np.complex128)
        fld = np.zeros(self.fld_shape(nbc), np.complex128)
        self._chk(self._lib.oshun_download_state(self._h, _dptr(f.view(np.float64)),
                                                 _dptr(fld.view(np.float64)), nbc))
        return f, fld

    def fields(self):
        """Ex, Ey, Ez, Bx, By, Bz as float64[6, ny, nx] (no distribution transfer)."""
        fld = np.zeros(self.fld_shape(), np.complex128)
        self._chk(self._lib.oshun_download_state(self._h, None, _dptr(fld.view(np.float64)), 0))
        return fld.real.copy()

    def set_time(self, t):
        self._chk(self._lib.oshun_set_time(self._h, float(t)))

    def time(self):
        t = ctypes.c_double()
        self._chk(self._lib.oshun_get_time(self._h, ctypes.byref(t)))
        return t.value

    # ---- operators ----------------------------------------------------------
    def rhs(self):
        """Actions::Adv of the current state (def-Actions.cpp:73-94)."""
        f = np.zeros(self.f_shape(), np.complex128)
        fld = np.zeros(self.fld_shape(), np.complex128)
        self._chk(self._lib.oshun_rhs(self._h, _dptr(f.view(np.float64)), _dptr(fld.view(np.float64))))
        return f, fld

    def rk4_step(self, h):
        self._chk(self._lib.oshun_rk4_step(self._h, float(h)))

    def collide_ee(self, tnew):
        self._chk(self._lib.oshun_collide_ee(self._h, float(tnew)))

    def collide_flm(self, tnew, which):
        self._chk(self._lib.oshun_collide_flm(self._h, float(tnew), int(which)))

    def step(self, h, nsteps=1):
        self._chk(self._lib.oshun_step(self._h, float(h), int(nsteps)))

    def set_sources(self, src, x, y, xglob):
        """Laser parameters (a dict of oshun_sources' fields) and the positions of the interior cells
        (x, y) and of the global x axis; oshun_step then runs the enabled sources every step."""
        s = Sources(**{name: (int(src.get(name, 0)) if typ is ctypes.c_int else float(src.get(name, 0.0)))
                       for name, typ in Sources._fields_})
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        xg = np.ascontiguousarray(xglob, np.float64)
        gd = self.grid.nbc if (self.grid.bndx or self.grid.bndy) else 0  # mirror: every cell, guards included
        if x.size != self.nx + 2 * gd or y.size != self.ny + 2 * gd:
            raise ValueError(f"set_sources: {self.nx + 2 * gd} x and {self.ny + 2 * gd} y positions expected")
        self._chk(self._lib.oshun_set_sources(self._h, ctypes.byref(s), _dptr(x), _dptr(y), _dptr(xg), int(xg.size)))

    def laser_source(self, tnew):
        """PLaserSource::Laser_Source (def-PLSource.cpp:138-252)."""
        self._chk(self._lib.oshun_laser_source(self._h, float(tnew)))

    def inverse_bremsstrahlung(self, tnew):
        """InverseBremsstrahlung::loop (def-PLSource.cpp:780-861)."""
        self._chk(self._lib.oshun_inverse_bremsstrahlung(self._h, float(tnew)))

    def laser_efield(self, tnew):
        """ExternalEfield::Laser_Efield (def-PLSource.cpp:368-418)."""
        self._chk(self._lib.oshun_laser_efield(self._h, float(tnew)))

    def step_implicit_1d(self, h, nsteps=1, es=0):
        """main.cpp:79-107 with implicit_E and if_implicit1D (Runge_Kutta_2_Imp +
        Implicit_E_Field); 1D, m0 = 0, l0 >= 3, FP64."""
        self._chk(self._lib.oshun_step_implicit_1d(self._h, float(h), int(nsteps), int(es)))

    def step_implicit_2d(self, h, nsteps=1, es=0):
        """main.cpp:79-107 with implicit_E, not if_implicit1D (Runge_Kutta_2_Imp +
        Implicit_E_Field with the 3x3 conductivity solve); l0 >= 3, m0 >= 2, FP64.
        Returns the cells left with a zero determinant in the last step."""
        self._chk(self._lib.oshun_step_implicit_2d(self._h, float(h), int(nsteps), int(es)))
        return self._lib.oshun_implicit_zero_dets(self._h)

    def set_implicit_es_current(self, jn, nbc=0):
        """JN of the electrostatic 2D implicit field: complex [3][ny+2nbc][nx+2nbc]."""
        shp = self.fld_shape(nbc)
        jn = np.ascontiguousarray(jn, np.complex128).reshape((3,) + tuple(shp[1:]))
        self._chk(self._lib.oshun_set_implicit_es_current(self._h, _dptr(jn.view(np.float64)), int(nbc)))

    def fill_guards(self):
        """Parallel_Environment::Neighbor_Communications of one node (mirror-boundary contexts)."""
        self._chk(self._lib.oshun_fill_guards(self._h))

    def synchronize(self):
        self._chk(self._lib.oshun_synchronize(self._h))

    def stream(self):
        s = ctypes.c_void_p()
        self._chk(self._lib.oshun_get_stream(self._h, ctypes.byref(s)))
        return s.value

    def plane_info(self):
        n = ctypes.c_int()
        pe = ctypes.c_longlong()
        self._chk(self._lib.oshun_plane_info(self._h, ctypes.byref(n), ctypes.byref(pe), None, None, None))
        L = (ctypes.c_int * n.value)()
        M = (ctypes.c_int * n.value)()
        IM = (ctypes.c_int * n.value)()
        self._chk(self._lib.oshun_plane_info(self._h, ctypes.byref(n), ctypes.byref(pe), L, M, IM))
        return dict(ncomp=n.value, plane=pe.value, l=list(L), m=list(M), im=list(IM))

    def state_bytes(self):
        """Bytes of the stored distribution function (the S_f of the roofline)."""
        info = self.plane_info()
        esz = 4 if self.precision == FP32_SPLIT else 8
        return info["ncomp"] * info["plane"] * esz

    def moments(self):
        """The moments of MOMENT_NAMES per (y, x) in FP64 (def-Parallel.cpp:1677-2253)."""
        out = np.zeros((len(MOMENT_NAMES), self.ny, self.nx))
        self._chk(self._lib.oshun_moments(self._h, _dptr(out)))
        return out

    def moments_noalloc(self, out_ptr):
        """Moments into the device buffer out_ptr (len(MOMENT_NAMES) x ny x nx doubles) on the
        context's stream, allocating nothing (static scratch; nx*ny <= 104857)."""
        self._chk(self._lib.oshun_moments_noalloc(self._h, out_ptr))

    # ---- device-resident state and the adjoint (device pointers, e.g. torch data_ptr())
    def copy_device_state(self, f_ptr, fld_ptr, into_ctx):
        """Copy the device state to (into_ctx False) or from (True) device buffers laid
        out as oshun_device_state: f = ncomp planes of ny*np*nx, fields 6*nx*ny."""
        self._chk(self._lib.oshun_copy_device_state(self._h, f_ptr, fld_ptr, 1 if into_ctx else 0))

    def adjoint_step(self, h, lam_f_ptr, lam_fld_ptr):
        """Cotangent of the step output -> of its input, in place (device pointers)."""
        self._chk(self._lib.oshun_adjoint_step(self._h, float(h), lam_f_ptr, lam_fld_ptr))

    def adjoint_forward(self, h, nsteps, every):
        """Record a checkpointed run; returns its generation number."""
        gen = ctypes.c_longlong(0)
        self._chk(self._lib.oshun_adjoint_forward(self._h, float(h), int(nsteps), int(every), ctypes.byref(gen)))
        return gen.value

    def adjoint_backward(self, lam_f_ptr, lam_fld_ptr, generation=0):
        """Cotangent of the recorded run's final state -> of its initial state, in place.
        generation > 0: fail unless that run is still the recorded one."""
        self._chk(self._lib.oshun_adjoint_backward(self._h, int(generation), lam_f_ptr, lam_fld_ptr))

    def diag_copy(self, nbytes):
        """Plain device copy of nbytes between scratch buffers (counter calibration)."""
        self._chk(self._lib.oshun_diag_copy(self._h, int(nbytes)))

    def stage_kernel(self, set=-1):
        """Which kernel runs the Vlasov stage: "1d" (m0 = 0, one y row), "spec" (compile-time
        specialised general kernel), "general" (term-list interpreter) or "lm" (l-streaming
        general kernel, m0 = 1..4).  set=1 forces the interpreter, set=0 restores the default,
        set=-1 only queries."""
        w = ctypes.c_int(-1)
        self._chk(self._lib.oshun_stage_kernel(self._h, int(set), ctypes.byref(w)))
        return ("1d", "spec", "general", "lm")[w.value]

    def pair_rows(self, set=0):
        """p rows of the 1D stage-pair kernels: 6 or 10 (tile kernels), -4 .. -32 (marched kernels, rows
        per wavefront); set chooses, 0 only queries."""
        r = ctypes.c_int(0)
        self._chk(self._lib.oshun_pair_rows(self._h, int(set), ctypes.byref(r)))
        return r.value

    def profile(self, on=True):
        self._chk(self._lib.oshun_profile_enable(self._h, 1 if on else 0))

    def profile_read(self):
        ms = (ctypes.c_double * 4)()
        n = (ctypes.c_longlong * 4)()
        self._chk(self._lib.oshun_profile_read(self._h, ms, n, 4))
        names = ["vlasov_stage", "field_stage", "collide", "diag_copy"]
        return {names[k]: (ms[k], n[k]) for k in range(4)}


class Runge_Kutta_4:
    """Explicit_Method facade (decl-RK.h:48-79): advance_p1() is one RK4 step of size h."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = float(h)

    def advance_p1(self):
        self.ctx.rk4_step(self.h)
        return self

    def advance_p2(self):
        return self

    def time(self):
        return self.ctx.time()


class Explicit_EC_FP:
    """e-e collisions of f00 (decl-fokkerplanck.h:140-163)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def loop(self, tnew):
        self.ctx.collide_ee(tnew)


class Euler_Backward:
    """Implicit l >= 1 collisions (decl-fokkerplanck.h:289-311)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def advance_1(self, tnew):
        self.ctx.collide_flm(tnew, 1)

    def advance_lm(self, tnew):
        self.ctx.collide_flm(tnew, 2)


class Batch:
    """Independent 1D runs (m0 = 0) of uneven shape advanced together, one
    hipGraph launch per step (oshun_batch_* in include/oshun_hip.h).

    ``runs`` is a list of keyword dicts for Context (l0, nump, numx, pmax, dx,
    precision, density_np, small_dt, nb_algorithms, collisions, ...); ``h`` the
    RK4 step of each run.  ``batch.run(i)`` is a Context view of run i for
    upload / download / moments; ``batch.step(n)`` advances every run n steps.
    """

    def __init__(self, runs, h):
        lib = _load()
        n = len(runs)
        grids = (Grid * n)()
        phys = (Phys * n)()
        for i, r in enumerate(runs):
            nump = r["nump"]
            pmax = r.get("pmax", 0.3)
            pmin = r.get("pmin", deck_pmin(pmax, nump))
            grids[i] = Grid(r["l0"], r.get("m0", 0), nump, r["numx"], 1, pmin, pmax, r.get("dx", 1.0), 1.0,
                            r.get("precision", FP64))
            phys[i] = Phys(r.get("density_np", 1e21), r.get("zeta", 1.0), r.get("small_dt", 1.0),
                           r.get("nb_algorithms", 4), r.get("collisions", COLL_EE | COLL_FLM),
                           r.get("ref_quirks", 1))
        hs = np.ascontiguousarray(h, np.float64)
        hb = ctypes.c_void_p()
        rc = lib.oshun_batch_create(grids, phys, _dptr(hs), n, ctypes.byref(hb))
        if rc != 0:
            raise OshunError(f"oshun_batch_create failed ({rc})")
        self._h = hb
        self._lib = lib
        self.grids, self.phys = grids, phys
        self.h = hs
        self._runs = [Context._view(lib, lib.oshun_batch_run(hb, i), grids[i], phys[i], self) for i in range(n)]

    def _chk(self, rc):
        if rc != 0:
            msg = self._lib.oshun_batch_last_error(self._h)
            raise OshunError(f"oshun batch call failed ({rc}): {msg.decode() if msg else ''}")

    def __len__(self):
        return len(self._runs)

    def run(self, i):
        return self._runs[i]

    def step(self, nsteps=1):
        self._chk(self._lib.oshun_batch_step(self._h, int(nsteps)))

    def synchronize(self):
        self._chk(self._lib.oshun_batch_synchronize(self._h))

    def autotune(self, steps=3):
        """Times each stage-pair kernel choice on the batch (state restored); returns (p rows, ms/step).

        p rows as pair_rows reports them: positive for the tile kernels, negative
        (rows per wavefront) for the marched kernels."""
        rows = ctypes.c_int()
        ms = ctypes.c_double()
        self._chk(self._lib.oshun_batch_autotune(self._h, int(steps), ctypes.byref(rows), ctypes.byref(ms)))
        return rows.value, ms.value

    def pair_rows(self, set=0):
        """The batch's stage-pair kernels: 6, 10 (14 in FP32) tile rows, or -4 .. -32
        marched rows per wavefront; 0 only queries.  Returns the setting."""
        r = ctypes.c_int()
        self._chk(self._lib.oshun_batch_pair_rows(self._h, int(set), ctypes.byref(r)))
        return r.value

    def close(self):
        if getattr(self, "_h", None):
            for r in self._runs:
                r._h = None
            self._lib.oshun_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
