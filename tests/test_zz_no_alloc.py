"""oshun_moments_noalloc (the allocation-free variant of oshun_moments) against oshun_moments,
through the C ABI: two calls in a row, the second on a smaller grid, and the size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(l0, m0, np_, nx, ny, seed):
    from test_gpu_parity import _maxwellian_state
    return _maxwellian_state(l0, m0, np_, nx, ny, 0.3, seed=seed, amp=1e-2)


def _both(ctx):
    import torch
    ref = ctx.moments()
    out = torch.full(ref.shape, float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.moments_noalloc(out.data_ptr())
    ctx.synchronize()
    return ref, out.cpu().numpy()


def test_noalloc_moments_match_the_allocating_entry_point():
    import oshun_amd
    print("case noalloc moments: l0=3 m0=2 np=96 nx=48 ny=6, then l0=2 m0=1 np=64 nx=20 ny=3", flush=True)
    ctx1 = oshun_amd.Context(3, 2, 96, 48, 6, pmax=0.3, dx=0.5, dy=0.7)
    f, fld = _state(3, 2, 96, 48, 6, seed=11)
    ctx1.upload(f, fld)
    ref1, got1 = _both(ctx1)
    assert np.array_equal(ref1, got1)                    # same chunking: bitwise
    ctx2 = oshun_amd.Context(2, 1, 64, 20, 3, pmax=0.3, dx=0.5, dy=0.7)
    f, fld = _state(2, 1, 64, 20, 3, seed=12)
    ctx2.upload(f, fld)
    ref2, got2 = _both(ctx2)                             # second call, smaller input, same scratch
    assert np.array_equal(ref2, got2)
    ref1b, got1b = _both(ctx1)                           # and the first context again
    assert np.array_equal(ref1b, got1b) and np.array_equal(got1b, got1)


def test_noalloc_moments_large_grid_and_size_limit():
    import oshun_amd
    print("case noalloc moments: nx*ny = 65536 (one chunk), nx*ny = 131072 (refused)", flush=True)
    ctx = oshun_amd.Context(1, 0, 128, 256, 256, pmax=0.3, dx=0.5, dy=0.7)
    # smooth T(x, y) with gradients along both axes, so that Kx and Ky are well conditioned
    p = np.arange(128) * ((0.3 - 0.3 / 256) / 127) + 0.3 / 256
    gx, gy = np.meshgrid(2 * np.pi * (np.arange(256) + 0.3) / 256, 2 * np.pi * (np.arange(256) + 0.7) / 256)
    tt = (1.0 + 0.1 * np.cos(gx) + 0.1 * np.sin(gy)) / 511.0
    f = np.zeros((2, 256, 256, 128), np.complex128)
    f[0] = np.exp(-p ** 2 / (2 * tt[..., None])) / (2 * np.pi * tt[..., None]) ** 1.5
    f[1] = 1e-3 * np.sin(gx)[..., None] * (p / np.sqrt(tt[..., None])) * f[0]
    ctx.upload(f, np.zeros((6, 256, 256), np.complex128))
    ref, got = _both(ctx)
    assert np.array_equal(ref, got)  # the chunking is capped by the scratch in both entry points
    big = oshun_amd.Context(1, 0, 4, 512, 256, pmax=0.3)
    import torch
    out = torch.zeros((len(oshun_amd.MOMENT_NAMES), 256, 512), dtype=torch.float64, device="cuda")
    with pytest.raises(oshun_amd.OshunError, match="exceeds"):
        big.moments_noalloc(out.data_ptr())
