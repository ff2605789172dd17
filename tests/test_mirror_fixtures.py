"""The mirror-boundary fixtures (tests/golden/make_golden.py mirror_*) and the
guard fill the driver restates from sameNode_mirror_X/Y (def-Parallel.cpp:
1160-1307, which lives with the reference's MPI code and is not built here).

The fill is pinned by what a reflection about the cell face means rather than
by its index formulas: in the initial whole arrays, the guard cell at distance
d outside a face holds the interior cell at distance d inside it, times
(-1)^(l+m) across x and (-1)^m conj across y; Ey, Ez, Bx change sign across x
and Ex, Ez, By across y; the harmonics the reference's loops leave out (l = l0
across x; l = 0 and l = l0 across y) keep zero guards; and a periodic side
holds the periodic images.  Each check also runs on a copy with one planted
slip (the image taken one cell off), which it must reject.
"""
import numpy as np
import pytest

from oracle.oracle import hind

CASES = ["mirror_x_2d", "mirror_y_2d", "mirror_xy_rk4"]


def _harmonics(c):
    l0, m0 = int(c["l0"]), int(c["m0"])
    return [(l, m, hind(l, m, m0)) for l in range(l0 + 1) for m in range(min(l, m0) + 1)]


def x_guard_errors(c, f, fld, shift=0, l0_zero=True):
    """max |guard - expected image| over the x guards of the interior rows (with an x mirror the
    l = l0 guards are expected zero, or skipped when l0_zero is False)"""
    g, l0 = int(c["nbc"]), int(c["l0"])
    nxg = f.shape[2]
    nx = nxg - 2 * g
    rows = slice(g, f.shape[1] - g)
    err = 0.0
    for d in range(1, g + 1):
        lg, li = g - d, g + d - 1 + shift            # left guard, interior cell at distance d
        rg, ri = g + nx - 1 + d, g + nx - d - shift  # right guard, interior cell at distance d
        for l, m, h in _harmonics(c):
            if int(c["bndX"]) and l == l0 and not l0_zero:
                continue
            if int(c["bndX"]):
                s = (-1.0) ** (l + m)
                el = np.zeros_like(f[h, rows, lg]) if l == l0 else s * f[h, rows, li]
                er = np.zeros_like(f[h, rows, rg]) if l == l0 else s * f[h, rows, ri]
            else:
                el, er = f[h, rows, g + nx - d], f[h, rows, g + d - 1]
            err = max(err, np.abs(f[h, rows, lg] - el).max(), np.abs(f[h, rows, rg] - er).max())
        for k in range(6):
            s = -1.0 if (int(c["bndX"]) and k in (1, 2, 3)) else 1.0
            if int(c["bndX"]):
                el, er = s * fld[k, rows, li], s * fld[k, rows, ri]
            else:
                el, er = fld[k, rows, g + nx - d], fld[k, rows, g + d - 1]
            err = max(err, np.abs(fld[k, rows, lg] - el).max(), np.abs(fld[k, rows, rg] - er).max())
    return err


def y_guard_errors(c, f, fld, shift=0):
    """max |guard - expected image| over the y guards of every column"""
    g, l0 = int(c["nbc"]), int(c["l0"])
    nyg = f.shape[1]
    ny = nyg - 2 * g
    err = 0.0
    for d in range(1, g + 1):
        lg, li = g - d, g + d - 1 + shift
        rg, ri = g + ny - 1 + d, g + ny - d - shift
        for l, m, h in _harmonics(c):
            if int(c["bndY"]):
                if l == 0 or l == l0:
                    continue  # left as they were: the x pass (or zero) -- not a y image
                el, er = (-1.0) ** m * np.conj(f[h, li]), (-1.0) ** m * np.conj(f[h, ri])
            else:
                el, er = f[h, g + ny - d], f[h, g + d - 1]
            err = max(err, np.abs(f[h, lg] - el).max(), np.abs(f[h, rg] - er).max())
        for k in range(6):
            s = -1.0 if (int(c["bndY"]) and k in (0, 2, 4)) else 1.0
            if int(c["bndY"]):
                el, er = s * fld[k, li], s * fld[k, ri]
            else:
                el, er = fld[k, g + ny - d], fld[k, g + d - 1]
            err = max(err, np.abs(fld[k, lg] - el).max(), np.abs(fld[k, rg] - er).max())
    return err


@pytest.mark.parametrize("name", CASES)
def test_initial_guards_are_the_images(golden, name):
    g = golden(name)
    c = g["cfg"]
    f, fld = g["f_ext0"], g["fld_ext0"]
    nb = int(c["nbc"])
    # the interior of the whole arrays is the input state
    assert np.array_equal(f[:, nb:-nb, nb:-nb], g["f_in"])
    assert np.array_equal(fld[:, nb:-nb, nb:-nb], g["fld_in"])
    assert x_guard_errors(c, f, fld) == 0.0
    assert y_guard_errors(c, f, fld) == 0.0
    # a slip of one cell in the image is seen
    if int(c["bndX"]):
        assert x_guard_errors(c, f, fld, shift=1) > 0.0
    if int(c["bndY"]):
        assert y_guard_errors(c, f, fld, shift=1) > 0.0
    # the harmonics the mirror leaves out keep zero guards (never written)
    l0, m0 = int(c["l0"]), int(c["m0"])
    if int(c["bndX"]):
        for m in range(min(l0, m0) + 1):
            h = hind(l0, m, m0)
            assert not np.any(f[h, nb:-nb, :nb]) and not np.any(f[h, nb:-nb, -nb:])


@pytest.mark.parametrize("name", CASES)
def test_final_state_keeps_the_fill_and_moves(golden, name):
    """After the steps the guards are again the images of the final interior (the
    fill runs at the end of every step), and the interior moved."""
    g = golden(name)
    c = g["cfg"]
    f, fld = g["f_ext"], g["fld_ext"]
    nb = int(c["nbc"])
    assert np.array_equal(f[:, nb:-nb, nb:-nb], g["f_out"])
    # with an x mirror the l0 guards evolve with the stages instead of being refreshed
    assert x_guard_errors(c, f, fld, l0_zero=False) == 0.0
    assert y_guard_errors(c, f, fld) == 0.0
    if int(c["bndX"]):
        l0, m0 = int(c["l0"]), int(c["m0"])
        assert np.any(f[hind(l0, 0, m0), nb:-nb, :nb])
    assert np.abs(g["f_out"] - g["f_in"]).max() > 1e-6 * np.abs(g["f_in"]).max()
