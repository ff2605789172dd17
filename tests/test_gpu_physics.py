"""Physics checks of the HIP path on the benchmark configurations.

C2 (collisionless Langmuir wave): Nx = 256 over one wavelength, Np = 256 to
pmax = 0.3, l0 = 8, m0 = 0, Maxwellian at 1 keV with dn/n = 1e-3 sin(kx),
k lambda_D = 0.3, 2000 RK4 steps of h = 0.01/omega_p, collisions off.
The E_x mode must oscillate at the real part and decay at the imaginary part
of the root of the Maxwellian dispersion relation 1 + (1 + zeta Z(zeta))/(k lambda_D)^2 = 0
(textbook Landau damping; the root is computed here with scipy's Faddeeva
function).  The FP32-split path is compared with FP64 on the same run and its
error per field is printed and bounded (f00 is stored in float relative to
its FP64 x-mean, so float rounding acts on the 1e-3 deviation only).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KLD = 0.3


def landau_root(kld=KLD):
    from scipy.optimize import fsolve
    from scipy.special import wofz

    def disp(w):
        z = complex(w[0], w[1]) / (np.sqrt(2.0) * kld)
        Z = 1j * np.sqrt(np.pi) * wofz(z)
        d = 1.0 + (1.0 + z * Z) / kld ** 2
        return [d.real, d.imag]

    wr, wi = fsolve(disp, [1.16, -0.01])
    return wr, wi


def langmuir_state(nx=256, np_=256, l0=8, pmax=0.3, dn=1e-3):
    pt = np.sqrt(1.0 / 511.0)
    k = KLD / pt
    L = 2 * np.pi / k
    dx = L / nx
    pmin = pmax / (2.0 * np_)
    p = np.arange(np_) * ((pmax - pmin) / (np_ - 1)) + pmin
    x = np.arange(nx) * dx
    f = np.zeros((l0 + 1, 1, nx, np_), np.complex128)
    f[0, 0] = (1.0 + dn * np.sin(k * x))[:, None] * np.exp(-p[None, :] ** 2 / (2 * pt ** 2)) / (
        np.sqrt(2 * np.pi) * 2 * np.pi * pt ** 3)
    fld = np.zeros((6, 1, nx), np.complex128)
    fld[0, 0] = dn / k * np.cos(k * x)  # Gauss's law for the electron perturbation
    return f, fld, dx, k, x


def run_langmuir(precision, steps=2000, h=0.01):
    import oshun_amd
    f, fld, dx, k, x = langmuir_state()
    ctx = oshun_amd.Context(8, 0, 256, 256, 1, pmax=0.3, dx=dx, precision=precision, collisions=0)
    ctx.upload(f, fld)
    amp = np.zeros(steps + 1)
    c, s = np.cos(k * x), np.sin(k * x)
    E = ctx.fields()[0, 0]
    amp[0] = 2.0 / len(x) * (E * c).sum()
    for n in range(steps):
        ctx.step(h, 1)
        E = ctx.fields()[0, 0]
        amp[n + 1] = 2.0 / len(x) * (E * c).sum()
    fo, fldo = ctx.download()
    return amp, fo, fldo


def fit_damped(t, a, wr0, wi0):
    from scipy.optimize import curve_fit

    def model(tt, c0, A, g, w, ph):
        return c0 + A * np.exp(g * tt) * np.cos(w * tt + ph)

    p0 = [0.0, np.abs(a).max(), wi0, wr0, 0.0]
    popt, _ = curve_fit(model, t, a, p0=p0, maxfev=20000)
    return popt


def test_langmuir_wave_frequency_and_landau_damping():
    import oshun_amd
    print("case C2 Langmuir wave, 2000 steps, FP64 and FP32 split", flush=True)
    wr, wi = landau_root()
    h = 0.01
    a64, f64, fl64 = run_langmuir(oshun_amd.FP64)
    t = np.arange(len(a64)) * h
    sel = t > 2.0  # skip the ballistic transient
    c0, A, g, w, ph = fit_damped(t[sel], a64[sel], wr, wi)
    print(f"theory omega={wr:.5f} gamma={wi:.5f}; FP64 fit omega={w:.5f} gamma={g:.5f}")
    assert abs(w - wr) / wr < 0.03
    assert abs(g - wi) / abs(wi) < 0.25
    # FP32 split on the same run: error per field against FP64
    a32, f32, fl32 = run_langmuir(oshun_amd.FP32_SPLIT)
    e = {
        "f00": np.linalg.norm(f32[0] - f64[0]) / np.linalg.norm(f64[0]),
        "f10": np.linalg.norm(f32[1] - f64[1]) / np.linalg.norm(f64[1]),
        "Ex": np.linalg.norm(fl32[0] - fl64[0]) / np.linalg.norm(fl64[0]),
    }
    print("FP32-split relative L2 error after 2000 steps:", {k: f"{v:.3e}" for k, v in e.items()})
    _, _, g32, w32, _ = fit_damped(t[sel], a32[sel], wr, wi)
    print(f"FP32 split fit omega={w32:.5f} gamma={g32:.5f}")
    # target <= 1e-5 in f00 and E; the background-relative FP32 storage of
    # f00 gives about 1e-9 and 1e-7 here
    assert e["f00"] < 1e-7
    assert e["Ex"] < 1e-6
    assert e["f10"] < 1e-4
    assert abs(g32 - g) / abs(g) < 0.01
