"""The drop-in boundary compiles and links against the reference (no GPU needed).

oshun_amd/shim/oshun_shim.cpp is compiled with the reference's own headers
and linked ahead of eight unmodified reference translation units and the
reference classes' test driver (oracle/Makefile, target `shim`).  The link
must succeed, the replaced member functions must exist once in the binary
(the shim's), and they must call the library's C ABI.  main.cpp itself is not
linked: it needs MPI (def-Parallel.cpp), an external library.  The GPU side of
the same binary is tests/test_gpu_shim.py.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/OSHUN-1"
SHIM = os.path.join(ROOT, "oracle", "_ref", "oshun_ref_shim")
LIB = os.path.join(ROOT, "oshun_amd", "liboshun_hip.so")

REPLACED = ["Runge_Kutta_4::Runge_Kutta_4(Stat&, int)", "Runge_Kutta_4::advance_p1()",
            "Runge_Kutta_2_Imp::Runge_Kutta_2_Imp(Stat&, int)", "Runge_Kutta_2_Imp::advance_p1()",
            "Runge_Kutta_2_Imp::advance_p2()", "Explicit_EC_FP::loop(double const&)",
            "Euler_Backward::advance_1(double const&)", "Euler_Backward::advance_lm(double const&)",
            "Electric_Field_Methods::Implicit_E_Field::advance(Explicit_Method*, Euler_Backward*)",
            "PLaserSource::Laser_Source(double const&)", "InverseBremsstrahlung::loop(double const&)"]
ABI = ["oshun_create", "oshun_upload_harmonics", "oshun_download_harmonics", "oshun_set_time", "oshun_rk4_step",
       "oshun_collide_ee", "oshun_collide_flm", "oshun_step_implicit_1d", "oshun_set_sources", "oshun_laser_source",
       "oshun_inverse_bremsstrahlung"]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree absent (the link needs its sources)")
@pytest.mark.skipif(not os.path.exists(LIB), reason="liboshun_hip.so not built")
def test_shim_links_with_the_reference_translation_units():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "shim"], check=True, capture_output=True)
    assert os.path.exists(SHIM)
    syms = subprocess.run(["nm", "-C", "--defined-only", SHIM], check=True, capture_output=True,
                          text=True).stdout.splitlines()
    addr = {}
    for ln in syms:
        parts = ln.split(" ", 2)
        if len(parts) == 3:
            addr.setdefault(parts[2], set()).add(parts[0])
    for fn in REPLACED:  # one body (a constructor's complete/base variants alias it)
        assert len(addr.get(fn, ())) == 1, (fn, addr.get(fn))
    dyn = subprocess.run(["nm", "-D", "--undefined-only", SHIM], check=True, capture_output=True,
                         text=True).stdout
    for fn in ABI:
        assert fn in dyn, fn
    # every replaced body that does work calls into the library
    dis = subprocess.run(["objdump", "-d", "-C", "--no-show-raw-insn", SHIM], check=True, capture_output=True,
                         text=True).stdout
    for fn, call in (("Runge_Kutta_4::advance_p1()", "oshun_rk4_step"),
                     ("Runge_Kutta_2_Imp::advance_p1()", "oshun_step_implicit_1d"),
                     ("Explicit_EC_FP::loop(double const&)", "oshun_collide_ee"),
                     ("Euler_Backward::advance_1(double const&)", "oshun_collide_flm"),
                     ("Euler_Backward::advance_lm(double const&)", "oshun_collide_flm"),
                     ("PLaserSource::Laser_Source(double const&)", "oshun_laser_source"),
                     ("InverseBremsstrahlung::loop(double const&)", "oshun_inverse_bremsstrahlung")):
        body = dis.split(f"<{fn}>:\n", 1)[1].split("\n\n", 1)[0]
        assert f"<{call}@plt>" in body, fn
