"""Mirror boundaries on the GPU (oshun_grid bndx / bndy = 1): the device keeps
the reference's whole arrays, guards included, advances every cell with the
reference's flattened-array differences and refreshes the guards at the end
of each step as Neighbor_Communications does (sameNode_mirror_X/Y,
def-Parallel.cpp:1160-1307, with its loop bounds: l0 across x, l = 0 and l0
across y are not refreshed and evolve in their guards).

Against the fixtures mirror_x_2d (x mirror, 3 full steps), mirror_y_2d (y
mirror, 3 full steps), mirror_xy_rk4 (both, 4 RK4 steps) and the implicit-E
steps mirror_imp_x / mirror_imp_x_es / mirror_imp_x_src (x mirror, RKLevel-2
guards; the last with laser heating and inverse bremsstrahlung), written by the
compiled reference's classes with the driver's restatement of the fill: the
whole arrays after the steps, guards included, at <= 1e-12 normwise, and the
interior download.  Also: an interior-only upload fills the guards exactly as
the driver did before the first step.
"""
import numpy as np
import pytest

from test_gpu_parity import TOL, rel

pytestmark = pytest.mark.gpu

CASES = ["mirror_x_2d", "mirror_y_2d", "mirror_xy_rk4"]
IMP_CASES = ["mirror_imp_x", "mirror_imp_x_es", "mirror_imp_x_src"]
FLOOR_MULT = 5.0


def _ctx(g):
    import oshun_amd
    c = g["cfg"]
    coll = oshun_amd.COLL_EE | oshun_amd.COLL_FLM if int(c["op"]) in (1, 9) else 0  # op 2: RK4 alone
    return oshun_amd.Context(l0=int(c["l0"]), m0=int(c["m0"]), nump=int(c["np"]), numx=int(c["nx"]),
                             numy=int(c["ny"]), pmax=float(c["pmax"]), dx=float(g["dx"]), dy=float(g["dy"]),
                             density_np=float(c["density_np"]), zeta=float(c["zeta"]),
                             small_dt=float(c["small_dt"]), nb_algorithms=int(c["nb"]), collisions=coll,
                             bndx=int(c["bndX"]), bndy=int(c["bndY"]), nbc=int(c["nbc"]))


@pytest.mark.parametrize("name", CASES)
def test_mirror_steps_vs_reference(golden, name):
    g = golden(name)
    c = g["cfg"]
    nb = int(c["nbc"])
    print(f"case mirror {name}", flush=True)
    ctx = _ctx(g)
    ctx.upload(g["f_ext0"], g["fld_ext0"], nbc=nb)
    ctx.step(float(c["h"]), int(c["nsteps"]))
    f, fld = ctx.download(nbc=nb)
    ef, ee = rel(f, g["f_ext"]), rel(fld.real, g["fld_ext"].real)
    fi, fldi = ctx.download()
    print(f"  whole arrays: f {ef:.2e} fields {ee:.2e}; interior f {rel(fi, g['f_out']):.2e}", flush=True)
    assert ef <= TOL and ee <= TOL
    assert rel(fi, g["f_out"]) <= TOL and rel(fldi.real, g["fld_out"].real) <= TOL
    assert np.array_equal(fi, f[:, nb:-nb, nb:-nb])


@pytest.mark.parametrize("name", CASES)
def test_interior_upload_fills_the_guards(golden, name):
    g = golden(name)
    nb = int(g["cfg"]["nbc"])
    print(f"case mirror fill {name}", flush=True)
    ctx = _ctx(g)
    ctx.upload(g["f_in"], g["fld_in"])  # interior only: guards zero, then refreshed on the device
    f, fld = ctx.download(nbc=nb)
    assert np.array_equal(f, g["f_ext0"])
    assert np.array_equal(fld, g["fld_ext0"])


@pytest.mark.parametrize("name", IMP_CASES)
def test_mirror_implicit_steps_vs_reference(golden, name):
    """implicit-E steps without if_implicit1D with an x mirror and RKLevel-2 guards
    (the production deck's boundary): whole arrays within FLOOR_MULT times the
    reference's own one-ulp noise floor of its field solve (stored in the fixture)"""
    g = golden(name)
    c = g["cfg"]
    nb = int(c["nbc"])
    print(f"case mirror implicit {name}", flush=True)
    ctx = _ctx(g)
    ctx.upload(g["f_ext0"], g["fld_ext0"], nbc=nb)
    if int(c["es"]):
        ctx.set_implicit_es_current(np.stack([g["fld_ext0"][0]] * 3), nbc=nb)
    if "src_I_0" in c:  # the laser on every cell of the whole arrays, inverse bremsstrahlung on the interior
        from oracle.oracle import SOURCE_KEYS, deck_axis
        x, xg = deck_axis(int(c["nx"]), float(c["xmin"]), float(c["xmax"]), guards=nb, whole=True)
        y, _ = deck_axis(int(c["ny"]), float(c["ymin"]), float(c["ymax"]), guards=nb, whole=True)
        ctx.set_sources({k: float(c["src_" + k]) for k in SOURCE_KEYS}, x, y, xg)
    z = ctx.step_implicit_2d(float(c["h"]), int(c["nsteps"]), int(c["es"]))
    f, fld = ctx.download(nbc=nb)
    ef = rel(f, g["f_ext"])
    ee = [rel(fld[k].real, g["fld_ext"][k].real) if np.any(g["fld_ext"][k]) else float(np.abs(fld[k]).max())
          for k in range(6)]
    wf, we = float(g["floor_f"]), g["floor_fld"]
    print(f"  f {ef:.2e} (floor {wf:.1e}), fields {['%.1e' % v for v in ee]} (floors {['%.1e' % v for v in we]})",
          flush=True)
    assert z == 0
    assert wf < 1e-10 and np.all(we < 1e-4)
    assert ef <= max(TOL, FLOOR_MULT * wf)
    for k in range(6):
        assert ee[k] <= max(TOL, FLOOR_MULT * we[k]), k


def test_mirror_context_rejects_what_it_does_not_run():
    import oshun_amd
    ctx = oshun_amd.Context(l0=3, m0=2, nump=12, numx=5, numy=5, pmax=0.25, bndx=1, nbc=4)
    with pytest.raises(oshun_amd.OshunError):
        ctx.moments()
    with pytest.raises(oshun_amd.OshunError):
        ctx.upload(np.zeros(ctx.f_shape(2), np.complex128), np.zeros(ctx.fld_shape(2), np.complex128), nbc=2)
    with pytest.raises(oshun_amd.OshunError):
        oshun_amd.Context(l0=3, m0=2, nump=12, numx=3, numy=5, pmax=0.25, bndx=1, nbc=4)  # numx < nbc
