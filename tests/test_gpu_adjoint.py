"""Adjoint of the step (oshun_adjoint_*), 1D, m0 = 0, FP64.

* One step, operator by operator (collisions off, e-e only, implicit l >= 1
  only, all on): <adjoint(lam), d> against the central difference
  <lam, (step(y + eps d) - step(y - eps d)) / (2 eps)> for seeded random lam, d,
  <= 1e-8 relative.
* Against the FP64 tangent-linear model of the oracle (forward-mode dual
  numbers, tests/test_oracle_tlm.py pins it): <adjoint(lam_N), d_0> =
  <lam_N, TLM(d_0)> over six checkpointed steps, <= 1e-10 relative.
* C5, the adjoint sensitivity configuration: Nx = 512, Np = 192, l0 = 3,
  200 RK4 + FP steps checkpointed every 20; objective J = sum_x |q_x| dx at
  the last step; gradient with respect to the 512 initial temperatures T(x)
  (T0 = 1 keV, seeded 8-mode Fourier perturbation <= 5 %) through the
  torch.autograd wrapper, against FP64 central differences (h = 1e-6) along
  16 seeded random directions, <= 1e-6 relative.
The heat flux q_x of the objective is written in torch from the moment
formulas (Parallel_Output::nonrelativistic_moments, def-Parallel.cpp:1699-2134)
and checked against the library's oshun_moments first.
"""
import math

import numpy as np
import pytest

from test_gpu_parity import _maxwellian_state

pytestmark = pytest.mark.gpu

DENSITY = 1e21


def _torch():
    import torch
    return torch


def _to_dev(ctx, f, fld):
    torch = _torch()
    ctx.upload(f, fld)
    F = torch.empty(ctx.device_shape(), dtype=torch.float64, device="cuda")
    G = torch.empty(ctx.fld_device_shape(), dtype=torch.float64, device="cuda")
    ctx.copy_device_state(F.data_ptr(), G.data_ptr(), False)
    return F, G


def _step_from(ctx, F, G, h, nsteps=1):
    torch = _torch()
    F = F.contiguous()
    G = G.contiguous()
    torch.cuda.synchronize()
    ctx.copy_device_state(F.data_ptr(), G.data_ptr(), True)
    ctx.set_time(0.0)
    ctx.step(h, nsteps)
    F2 = torch.empty_like(F)
    G2 = torch.empty_like(G)
    ctx.copy_device_state(F2.data_ptr(), G2.data_ptr(), False)
    return F2, G2


STEP_CASES = [("vlasov", 0), ("ee", 1), ("flm", 2), ("all", 3)]


@pytest.mark.parametrize("name,coll", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_adjoint_step_matches_central_differences(name, coll):
    import oshun_amd
    torch = _torch()
    print(f"case adjoint step {name}: l0=3 np=48 nx=24", flush=True)
    l0, np_, nx, h = 3, 48, 24, 0.2
    kw = dict(pmax=0.3, dx=0.4, density_np=1e22, small_dt=0.1, nb_algorithms=4, collisions=coll)
    ctx = oshun_amd.Context(l0, 0, np_, nx, 1, **kw)
    f, fld = _maxwellian_state(l0, 0, np_, nx, 1, 0.3, seed=77, amp=1e-2)
    x = np.arange(nx) / nx
    fld[0] = 2e-3 * np.sin(2 * np.pi * x)[None, :]
    fld[1] = 1e-3 * np.cos(2 * np.pi * x)[None, :]
    fld[5] = 5e-4 * np.sin(4 * np.pi * x)[None, :]
    F, G = _to_dev(ctx, f, fld)
    gen = torch.Generator().manual_seed(1234)
    # every plane and field weighted by its own size, so that each operator counts;
    # the direction is relative to f00 point by point (the Maxwellian tail is tiny,
    # and an absolute perturbation there would leave the linear regime)
    rms = F.pow(2).mean(dim=(1, 2, 3), keepdim=True).sqrt()
    lam = (torch.randn(F.shape, generator=gen, dtype=torch.float64) / rms.cpu()).cuda()
    lamf = torch.randn(G.shape, generator=gen, dtype=torch.float64).cuda() * 1e3
    scale = torch.full((F.shape[0], 1, 1, 1), 1e-2, dtype=torch.float64)
    scale[0] = 1.0
    d = (torch.randn(F.shape, generator=gen, dtype=torch.float64) * scale).cuda() * F[0:1].abs()
    dG = torch.randn(G.shape, generator=gen, dtype=torch.float64).cuda() * 1e-3
    eps = 1e-5
    Fp, Gp = _step_from(ctx, F + eps * d, G + eps * dG, h)
    Fm, Gm = _step_from(ctx, F - eps * d, G - eps * dG, h)
    fd = float(((lam * (Fp - Fm)).sum() + (lamf * (Gp - Gm)).sum()) / (2 * eps))
    torch.cuda.synchronize()
    ctx.copy_device_state(F.data_ptr(), G.data_ptr(), True)
    ctx.set_time(0.0)
    L = lam.clone()
    Lf = lamf.clone()
    torch.cuda.synchronize()
    ctx.adjoint_step(h, L.data_ptr(), Lf.data_ptr())
    torch.cuda.synchronize()
    ad = float((L * d).sum() + (Lf * dG).sum())
    print(f"{name}: adjoint {ad:.12e} central difference {fd:.12e} rel {abs(ad - fd) / abs(fd):.2e}")
    assert math.isfinite(ad) and abs(ad - fd) <= 1e-8 * abs(fd)
    # the state is left as it was
    F2 = torch.empty_like(F)
    ctx.copy_device_state(F2.data_ptr(), None, False)
    assert torch.equal(F2, F)


# ---- C5 ----------------------------------------------------------------------
def nu_ei_over_wp(T_kev, Z=1.0, lnl=10.0):
    nu = 2.91e-6 * Z * DENSITY * lnl * (T_kev * 1e3) ** -1.5
    return nu / (5.64e4 * math.sqrt(DENSITY))


def lambda_ei(T_kev):
    return math.sqrt(T_kev / 511.0) / nu_ei_over_wp(T_kev)


def p_axis(np_, pmax):
    """Axis<double>(np, pmin, pmax) with the decks' pmin = pmax/(2 np) (matrices.h:860-866)."""
    pmin = pmax / (2.0 * np_)
    d = (pmax - pmin) / (np_ - 1)
    return np.arange(np_) * d + pmin


def heat_flux(f00, f10, f20, pr):
    """q_x per x from f00, f10, f20 [p][x] (def-Parallel.cpp:1699-2134, m0 = 0, Vy = 0)."""
    torch = _torch()
    N = pr.shape[0]
    dp = float(pr[1] - pr[0])
    ig = 1.0 / np.sqrt(1.0 + pr * pr)
    r = pr[0] ** 2 / pr[1] ** 2
    inv = 1.0 / (1.0 - r)

    def trap(f, p0, w):
        w = torch.as_tensor(w, dtype=torch.float64, device=f.device)
        s = (f[0] * w[0] if p0 == 0 else 0.0) + f[N - 1] * w[N - 1]
        s = s * 0.5
        return s + (w[1:N - 1, None] * f[1:N - 1]).sum(dim=0)

    w2, w3, w4, w5 = pr ** 2, pr ** 3 * ig, pr ** 4 * ig ** 2, pr ** 5 * ig ** 3
    n = trap(f00, 0, w2)
    c00 = f00[1] / 5.0 * r + (f00[0] - f00[1] * r) * inv * (1.0 / 3.0 - 1.0 / 5.0 * r)
    n = (n + c00 * (pr[0] * (pr[0] / dp) * pr[0])) * (4.0 * math.pi * dp)
    vx = (trap(f10, 0, w3) + f10[1] * (pr[0] ** 5 / (dp * pr[1] * 5.0))) * ((4.0 * math.pi / 3.0) * dp) / n
    c01 = f00[1] / 7.0 * r + (f00[0] - f00[1] * r) * inv * (1.0 / 5.0 - 1.0 / 7.0 * r)
    m1 = (trap(f00, 0, w4) + c01 * (pr[0] ** 4 * (pr[0] / dp))) * (4.0 * math.pi / 3.0 * dp)
    t20 = trap(f20, 1, w4)
    m2 = t20 * (8.0 * math.pi / 15.0 * dp)
    vxvx = (m1 + m2) / n
    m1b = m1 - 0.5 * m2
    t22 = t20 * (16.0 * math.pi / 5.0 * dp)
    vyvy = (t22 + m1b) / n
    vzvz = (m1b - t22) / n
    q = -vx * (vxvx + vyvy + vzvz) - 2.0 * (vxvx * vx) + 2.0 * (vx * vx) * vx
    q = q * n
    m5 = (trap(f10, 0, w5) + f10[1] * (pr[0] ** 7 / (dp * pr[1] * 7.0))) * ((4.0 * math.pi / 3.0) * dp)
    return 0.5 * (q + m5)


def test_adjoint_c5_gradient_vs_finite_differences():
    import oshun_amd
    from oshun_amd.autograd import run_steps
    torch = _torch()
    print("case C5 adjoint: Nx=512 Np=192 l0=3, 200 steps, checkpoints every 20", flush=True)
    nx, np_, l0, pmax, h, steps, every = 512, 192, 3, 0.25, 0.1, 200, 20
    L = 100.0 * lambda_ei(1.0)
    dx = L / nx
    x = np.arange(nx) * dx
    rng = np.random.default_rng(1234)
    pert = np.zeros(nx)
    for m in range(1, 9):
        pert += rng.uniform(-1, 1) * np.cos(2 * np.pi * m * x / L + rng.uniform(0, 2 * np.pi))
    T0 = 1.0 * (1.0 + 0.05 * pert / np.abs(pert).max())
    dirs = rng.standard_normal((16, nx))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ctx = oshun_amd.Context(l0, 0, np_, nx, 1, pmax=pmax, dx=dx, density_np=DENSITY, small_dt=h, nb_algorithms=4)
    pr = p_axis(np_, pmax)
    prt = torch.as_tensor(pr, dtype=torch.float64, device="cuda")
    nc = ctx.device_shape()[0]

    def initial(T):  # Setup_Y::Gaussian_p (def-Setup.cpp:500-533), n = 1
        pt2 = T / 511.0
        f00 = torch.exp(-prt[:, None] ** 2 / (2.0 * pt2[None, :])) / (math.sqrt(2 * math.pi) * 2 * math.pi * pt2 ** 1.5)
        f = torch.cat([f00[None], torch.zeros((nc - 1, np_, nx), dtype=torch.float64, device="cuda")])
        return f.reshape(nc, 1, np_, nx)

    def objective(f):
        q = heat_flux(f[0, 0], f[1, 0], f[2, 0], pr)
        return q.abs().sum() * dx, q

    fld0 = torch.zeros((6, 1, nx), dtype=torch.float64, device="cuda")
    # the torch heat flux against the library's moments
    f_init = initial(torch.as_tensor(T0, device="cuda"))
    F1, _ = _step_from(ctx, f_init, fld0, h, 3)
    _, q = objective(F1)
    mom = ctx.moments()
    assert np.abs(q.cpu().numpy() - mom[6, 0]).max() <= 1e-12 * np.abs(mom[6, 0]).max()
    # gradient by the adjoint
    T = torch.as_tensor(T0, device="cuda").clone().requires_grad_(True)
    ctx.set_time(0.0)
    fN, _ = run_steps(ctx, initial(T), fld0, h, steps, every)
    J, _ = objective(fN)
    J.backward()
    grad = T.grad.detach().cpu().numpy()
    print(f"J = {float(J.detach()):.12e}, |dJ/dT| = {np.linalg.norm(grad):.6e}")

    def J_of(Tv):
        Fe, _ = _step_from(ctx, initial(torch.as_tensor(Tv, device="cuda")), fld0, h, steps)
        return float(objective(Fe)[0])

    eps = 1e-6
    worst = 0.0
    for k, dvec in enumerate(dirs):
        fd = (J_of(T0 + eps * dvec) - J_of(T0 - eps * dvec)) / (2 * eps)
        ad = float(grad @ dvec)
        rel = abs(ad - fd) / abs(fd)
        worst = max(worst, rel)
        print(f"direction {k:2d}: adjoint {ad:+.10e} central difference {fd:+.10e} rel {rel:.2e}", flush=True)
    assert worst <= 1e-6


def test_adjoint_matches_the_tangent_linear_oracle():
    """<adjoint(lam_N), d_0> = <lam_N, TLM(d_0)> over 6 steps (checkpoints every 2),
    the tangent from the oracle's forward-mode model (Step1D<Dual>), all FP64."""
    import oshun_amd
    from oracle.oracle import Config, Oracle
    torch = _torch()
    print("case adjoint vs tangent-linear oracle: l0=3 np=48 nx=24, 6 steps", flush=True)
    l0, np_, nx, h, steps, every = 3, 48, 24, 0.2, 6, 2
    kw = dict(pmax=0.3, dx=0.4, density_np=1e22, small_dt=0.1)
    ctx = oshun_amd.Context(l0, 0, np_, nx, 1, nb_algorithms=4, **kw)
    f, fld = _maxwellian_state(l0, 0, np_, nx, 1, 0.3, seed=91, amp=1e-2)
    x = np.arange(nx) / nx
    fld[0] = 2e-3 * np.sin(2 * np.pi * x)[None, :]
    fld[1] = 1e-3 * np.cos(2 * np.pi * x)[None, :]
    fld[5] = 5e-4 * np.sin(4 * np.pi * x)[None, :]
    F, G = _to_dev(ctx, f, fld)
    ctx.set_time(0.0)
    ctx.adjoint_forward(h, steps, every)
    FN = torch.empty_like(F)
    GN = torch.empty_like(G)
    ctx.copy_device_state(FN.data_ptr(), GN.data_ptr(), False)
    gen = np.random.default_rng(2024)
    f0 = F.cpu().numpy()[:, 0].transpose(0, 2, 1).copy()  # [l][x][p]
    g0 = G.cpu().numpy()[:, 0].copy()
    rms = np.sqrt((f0 ** 2).mean(axis=(1, 2)))[:, None, None]
    lam = gen.standard_normal(f0.shape) / rms
    lamf = gen.standard_normal(g0.shape) * 1e3
    d = gen.standard_normal(f0.shape) * np.abs(f0[0])[None] * np.array([1.0] + [1e-2] * l0)[:, None, None]
    dg = gen.standard_normal(g0.shape) * 1e-3
    # GPU adjoint of the 6 steps
    L = torch.as_tensor(lam.transpose(0, 2, 1)[:, None].copy(), device="cuda")
    Lf = torch.as_tensor(lamf[:, None].copy(), device="cuda")
    torch.cuda.synchronize()
    ctx.adjoint_backward(L.data_ptr(), Lf.data_ptr())
    torch.cuda.synchronize()
    lam0 = L.cpu().numpy()[:, 0].transpose(0, 2, 1)
    lamf0 = Lf.cpu().numpy()[:, 0]
    # oracle tangent-linear model of the same 6 steps
    o = Oracle(Config(l0=l0, m0=0, np=np_, nx=nx, nb=4, **kw))
    fo, go, dfo, dgo = o.tlm_1d(f0, g0, d, dg, h=h, nsteps=steps)
    fgpu = FN.cpu().numpy()[:, 0].transpose(0, 2, 1)
    assert np.abs(fgpu - fo).max() <= 1e-12 * np.abs(fo).max()
    lhs = float((lam0 * d).sum() + (lamf0 * dg).sum())
    rhs = float((lam * dfo).sum() + (lamf * dgo).sum())
    print(f"<adjoint, d> {lhs:.14e}  <lam, TLM d> {rhs:.14e}  rel {abs(lhs - rhs) / abs(rhs):.2e}")
    assert abs(lhs - rhs) <= 1e-10 * abs(rhs)


def test_run_steps_rejects_a_replaced_recording_and_bad_tensors():
    """A second run_steps on one context replaces the first one's checkpoints: the first segment's
    backward then raises instead of differentiating the wrong trajectory; wrong tensors raise early."""
    import oshun_amd
    from oshun_amd.autograd import run_steps
    torch = _torch()
    print("case adjoint generation check: l0=2 np=32 nx=16", flush=True)
    l0, np_, nx = 2, 32, 16
    ctx = oshun_amd.Context(l0, 0, np_, nx, 1, pmax=0.3, dx=0.4, density_np=1e22, small_dt=0.1, nb_algorithms=4)
    f, fld = _maxwellian_state(l0, 0, np_, nx, 1, 0.3, seed=5)
    F, G = _to_dev(ctx, f, fld)
    F.requires_grad_(True)
    f1, g1 = run_steps(ctx, F, G, 0.1, 2, 1)
    f2, g2 = run_steps(ctx, f1, g1, 0.1, 2, 1)   # replaces the first recording
    with pytest.raises(oshun_amd.OshunError, match="replaced"):
        f2.sum().backward()
    # a single segment differentiates normally
    F.grad = None
    f1, _ = run_steps(ctx, F, G, 0.1, 2, 1)
    f1.sum().backward()
    assert torch.isfinite(F.grad).all()
    with pytest.raises(ValueError, match="elements"):
        run_steps(ctx, F.detach()[:, :, :-1].contiguous(), G, 0.1, 1, 1)
    with pytest.raises(ValueError, match="float64"):
        run_steps(ctx, F.detach(), G.float(), 0.1, 1, 1)
    with pytest.raises(ValueError, match="GPU"):
        run_steps(ctx, F.detach(), G.cpu(), 0.1, 1, 1)
