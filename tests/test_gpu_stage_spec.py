"""The compile-time specialised stage kernel of the general system
(vlasov_stage_spec) against the term-list interpreter (vlasov_stage) it
replaces for the shapes it is built for.  Both evaluate Actions::Adv
(def-Actions.cpp:73-94) with every target's terms in the same order and the
same operations, so the slopes and the RK4 steps agree bit for bit; the
parity of both with the reference is covered by the golden and oracle tests
of the step."""
import numpy as np
import pytest

from test_gpu_parity import _maxwellian_state

pytestmark = pytest.mark.gpu

# (l0, m0, Np, Nx, Ny): every shape with a specialised kernel, ragged Np and Nx
SHAPES = [(4, 4, 40, 70, 5), (4, 3, 16, 9, 4), (6, 3, 24, 12, 3), (7, 3, 10, 5, 4), (8, 2, 36, 20, 1),
          (3, 2, 64, 10, 1), (4, 4, 10, 9, 3)]  # the last: p rows past Np in the last workgroup


def _ctx(l0, m0, np_, nx, ny, prec):
    import oshun_amd
    return oshun_amd.Context(l0, m0, np_, nx, ny, pmax=0.3, dx=0.4, dy=0.5, density_np=1e22, small_dt=0.1,
                             precision=prec)


@pytest.mark.parametrize("l0,m0,np_,nx,ny", SHAPES)
def test_spec_slope_equals_interpreter(l0, m0, np_, nx, ny):
    import oshun_amd
    print(f"case stage spec slope l0={l0} m0={m0} np={np_} nx={nx} ny={ny}", flush=True)
    f, fld = _maxwellian_state(l0, m0, np_, nx, ny, 0.3, seed=l0 * 10 + m0, amp=1e-2)
    fld[3] = 2e-3
    fld[4] = -1e-3 * np.cos(np.arange(nx) / nx * 2 * np.pi)[None, :]
    out = {}
    for force in (0, 1):
        ctx = _ctx(l0, m0, np_, nx, ny, oshun_amd.FP64)
        assert ctx.stage_kernel(force) == ("spec", "general")[force]
        ctx.upload(f, fld)
        out[force] = ctx.rhs()
        ctx.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.abs(out[0][0]).max() > 0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("l0,m0,np_,nx,ny", [SHAPES[0], SHAPES[2], SHAPES[4], SHAPES[6]])
def test_spec_steps_equal_interpreter(l0, m0, np_, nx, ny, prec):
    import oshun_amd
    print(f"case stage spec steps l0={l0} m0={m0} np={np_} nx={nx} ny={ny} {prec}", flush=True)
    f, fld = _maxwellian_state(l0, m0, np_, nx, ny, 0.3, seed=l0 * 10 + m0 + 1)
    fld[5] = 1e-3
    p = oshun_amd.FP64 if prec == "fp64" else oshun_amd.FP32_SPLIT
    out = {}
    for force in (0, 1):
        ctx = _ctx(l0, m0, np_, nx, ny, p)
        ctx.stage_kernel(force)
        ctx.upload(f, fld)
        ctx.step(0.2, 3)
        out[force] = ctx.download()
        ctx.close()
    assert np.all(np.isfinite(out[0][0]))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_shapes_without_spec_use_the_lm_kernel():
    import oshun_amd
    print("case stage spec dispatch", flush=True)
    for shape, want in (((5, 1, 16, 8, 1), "lm"), ((3, 1, 16, 8, 3), "lm"), ((4, 4, 16, 8, 2), "spec"),
                        ((4, 0, 16, 8, 1), "1d"), ((6, 5, 16, 8, 2), "general")):
        ctx = _ctx(*shape, oshun_amd.FP64)
        assert ctx.stage_kernel() == want, shape
        ctx.close()


def test_kernel_knobs_reject_bad_values():
    import oshun_amd
    print("case stage knobs", flush=True)
    ctx = _ctx(4, 4, 16, 8, 2, oshun_amd.FP64)
    with pytest.raises(oshun_amd.OshunError):
        ctx.stage_kernel(5)
    with pytest.raises(oshun_amd.OshunError):
        ctx.pair_rows(10)  # no stage-pair kernels on a 2D context
    assert ctx.pair_rows() == 0
    ctx.close()
    ctx = _ctx(3, 0, 64, 64, 1, oshun_amd.FP64)
    assert ctx.pair_rows() == 6  # a small grid
    assert ctx.pair_rows(10) == 10
    with pytest.raises(oshun_amd.OshunError):
        ctx.pair_rows(7)
    ctx.close()
    ctx = _ctx(10, 0, 64, 64, 1, oshun_amd.FP64)
    assert ctx.pair_rows() == 10  # l0 > 8 keeps 10 rows
    with pytest.raises(oshun_amd.OshunError):
        ctx.pair_rows(6)
    ctx.close()
