"""The l-streaming stage kernel of the general system (vlasov_stage_lm, any l0
at m0 = 1..4) against the term-list interpreter (vlasov_stage) it replaces.
Both evaluate the coupling program of Actions::Adv (def-Actions.cpp:73-94) --
and of the implicit step's Adv1 / Adv2 (:118-160) -- with every target's terms
in the same order and the same operations, so the slopes, the RK4 steps, the
implicit-E steps and the mirror-boundary steps agree bit for bit.  The parity
of both with the reference is covered by the golden and oracle tests of each
step."""
import numpy as np
import pytest

from test_gpu_parity import _maxwellian_state

pytestmark = pytest.mark.gpu

# (l0, m0, Np, Nx, Ny): 2D and 1D, every m0, ragged Np and Nx, l0 from 1 past the interior
# range (l >= m0 + 1) to a production-sized harmonic count
SHAPES = [(5, 4, 40, 70, 5), (16, 3, 20, 12, 3), (48, 2, 12, 16, 2), (9, 1, 24, 10, 1), (6, 2, 33, 7, 1),
          (1, 1, 16, 9, 2), (2, 2, 10, 9, 3), (4, 4, 10, 9, 1)]


def _ctx(l0, m0, np_, nx, ny, prec, **kw):
    import oshun_amd
    return oshun_amd.Context(l0, m0, np_, nx, ny, pmax=0.3, dx=0.4, dy=0.5, density_np=1e22, small_dt=0.1,
                             precision=prec, **kw)


def _state(l0, m0, np_, nx, ny, seed):
    f, fld = _maxwellian_state(l0, m0, np_, nx, ny, 0.3, seed=seed, amp=1e-2)
    x = np.arange(nx) / nx
    fld[1] = 2e-5 * np.cos(2 * np.pi * x)[None, :]
    fld[2] = -1e-5 * np.sin(4 * np.pi * x)[None, :]
    fld[3] = 2e-3
    fld[4] = -1e-3 * np.cos(2 * np.pi * x)[None, :]
    return f, fld


@pytest.mark.parametrize("l0,m0,np_,nx,ny", SHAPES)
def test_lm_slope_equals_interpreter(l0, m0, np_, nx, ny):
    import oshun_amd
    print(f"case stage lm slope l0={l0} m0={m0} np={np_} nx={nx} ny={ny}", flush=True)
    f, fld = _state(l0, m0, np_, nx, ny, seed=l0 * 10 + m0)
    out = {}
    for force in (0, 1):
        ctx = _ctx(l0, m0, np_, nx, ny, oshun_amd.FP64)
        if force == 0:
            assert ctx.stage_kernel() == "lm"
        ctx.stage_kernel(force)
        ctx.upload(f, fld)
        out[force] = ctx.rhs()
        ctx.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.abs(out[0][0]).max() > 0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("l0,m0,np_,nx,ny", [SHAPES[0], SHAPES[1], SHAPES[3], SHAPES[5]])
def test_lm_steps_equal_interpreter(l0, m0, np_, nx, ny, prec):
    import oshun_amd
    print(f"case stage lm steps l0={l0} m0={m0} np={np_} nx={nx} ny={ny} {prec}", flush=True)
    f, fld = _state(l0, m0, np_, nx, ny, seed=l0 * 10 + m0 + 1)
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


@pytest.mark.parametrize("l0,m0,np_,nx,ny,es", [(16, 3, 24, 12, 6, 0), (12, 2, 20, 16, 2, 1), (5, 4, 16, 9, 5, 0)])
def test_lm_implicit_2d_equals_interpreter(l0, m0, np_, nx, ny, es):
    """Runge_Kutta_2_Imp's stages (modes 5, 6) over the Adv1 (with and without the B
    rotation) and Adv2 programs; the probes stay on the interpreter."""
    import oshun_amd
    print(f"case stage lm implicit 2D l0={l0} m0={m0} np={np_} nx={nx} ny={ny} es={es}", flush=True)
    f, fld = _state(l0, m0, np_, nx, ny, seed=l0 + 3)
    jn = np.zeros((3, ny, nx), np.complex128)
    jn[0] = 0.3 * fld[0]
    out = {}
    for force in (0, 1):
        ctx = _ctx(l0, m0, np_, nx, ny, oshun_amd.FP64)
        ctx.stage_kernel(force)
        ctx.upload(f, fld)
        if es:
            ctx.set_implicit_es_current(jn)
        assert ctx.step_implicit_2d(0.5, 2, es) == 0
        out[force] = ctx.download()
        ctx.close()
    assert np.all(np.isfinite(out[0][0]))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("bnd,implicit", [((1, 0), 0), ((0, 1), 0), ((1, 1), 0), ((1, 0), 1)])
def test_lm_mirror_equals_interpreter(bnd, implicit):
    """Whole arrays with guard cells (mirror boundaries): the flattened-array x / y
    differences of the reference's matrices, RK4 or the implicit step."""
    import oshun_amd
    l0, m0, np_, nx, ny, nbc = 10, 2, 16, 9, 6, 2
    print(f"case stage lm mirror bnd={bnd} implicit={implicit}", flush=True)
    f, fld = _state(l0, m0, np_, nx, ny, seed=17)
    out = {}
    for force in (0, 1):
        ctx = _ctx(l0, m0, np_, nx, ny, oshun_amd.FP64, bndx=bnd[0], bndy=bnd[1], nbc=nbc)
        if force == 0:
            assert ctx.stage_kernel() == "lm"
        ctx.stage_kernel(force)
        ctx.upload(f, fld)
        if implicit:
            ctx.step_implicit_2d(0.5, 2, 0)
        else:
            ctx.step(0.2, 2)
        out[force] = ctx.download(nbc)
        ctx.close()
    assert np.all(np.isfinite(out[0][0]))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
