"""The C-ABI library loads and exports every entry point its header declares (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "oshun_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(oshun_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("oshun_create", "oshun_destroy", "oshun_rhs", "oshun_rk4_step", "oshun_collide_ee",
                 "oshun_collide_flm", "oshun_step", "oshun_upload_harmonics", "oshun_download_harmonics"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import oshun_amd
    path = oshun_amd.library_path()
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (python -m oshun_amd._build)")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_null_arguments_are_rejected_without_touching_the_gpu():
    import oshun_amd
    lib = ctypes.CDLL(oshun_amd.library_path())
    lib.oshun_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    assert lib.oshun_create(None, None, None) == -1
    lib.oshun_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
    assert lib.oshun_step(None, 0.1, 1) == -1
    # invalid grid (l0 = 0) is rejected before any HIP call
    g = oshun_amd.Grid(0, 0, 64, 64, 1, 0.001, 0.3, 1.0, 1.0, 0)
    ph = oshun_amd.Phys(1e21, 1.0, 1.0, 4, 3, 1)
    h = ctypes.c_void_p()
    lib.oshun_create.argtypes = [ctypes.POINTER(oshun_amd.Grid), ctypes.POINTER(oshun_amd.Phys),
                                 ctypes.POINTER(ctypes.c_void_p)]
    assert lib.oshun_create(ctypes.byref(g), ctypes.byref(ph), ctypes.byref(h)) == -1


def _plan(l0, m0, np_, nx, ny, precision=0):
    import oshun_amd
    lib = ctypes.CDLL(oshun_amd.library_path())
    lib.oshun_kernel_plan.argtypes = [ctypes.POINTER(oshun_amd.Grid), ctypes.POINTER(oshun_amd.Phys),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    pmax = 0.3
    g = oshun_amd.Grid(l0, m0, np_, nx, ny, pmax / (2 * np_), pmax, 1.0, 1.0, precision)
    ph = oshun_amd.Phys(1e21, 1.0, 1.0, 4, 3, 1)
    st, co = ctypes.c_int(-9), ctypes.c_int(-9)
    assert lib.oshun_kernel_plan(ctypes.byref(g), ctypes.byref(ph), ctypes.byref(st), ctypes.byref(co)) == 0
    return st.value, co.value


def test_kernel_plan_keeps_32bit_offsets_in_range():
    """The stage-pair kernels address the whole state with 32-bit byte offsets, collide_rows one plane
    (their out-of-range sentinel is 0x7ffffff0): larger 1D states are planned on the 64-bit kernels
    (decided on the host, before any allocation or launch)."""
    assert _plan(8, 0, 256, 4096, 1) == (0, 0)               # the target shape: stage pairs + lane layout
    # 9 planes x 256 p x nx x 8 B: just below and at 2 GiB
    nx_ok = (2 ** 31 - 1) // (9 * 256 * 8)
    assert _plan(8, 0, 256, nx_ok, 1) == (0, 0)
    assert _plan(8, 0, 256, nx_ok + 1, 1) == (1, 0)           # 64-bit single-stage kernel
    assert _plan(8, 0, 256, nx_ok + 1, 1, precision=1) == (0, 0)  # float state: half the bytes
    # one plane above the collide_rows sentinel: one wavefront per column (64-bit)
    nx_big = 0x7ffffff0 // (256 * 8) + 1
    assert _plan(8, 0, 256, nx_big, 1) == (1, 1)
    assert _plan(4, 4, 128, 256, 256) == (2, 0)               # C3: specialised general kernel
    assert _plan(5, 4, 64, 16, 16)[0] == 4                    # no specialisation: l-streaming kernel
    assert _plan(48, 2, 288, 640, 2)[0] == 4                  # a production deck's shape
    assert _plan(6, 5, 64, 16, 16)[0] == 3                    # m0 > 4: interpreter
