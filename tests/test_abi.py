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
