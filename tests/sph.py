"""Projection of a phase-space function onto OSHUN's spherical harmonics (test helper).

Convention (the one the reference's moment formulas imply, def-Parallel.cpp:
1731-2013): x is the polar axis, mu = p_x/|p|, phi measured from y towards z,
    f(p, mu, phi) = sum_l f_l0(p) P_l(mu) + sum_{m>=1} 2 Re(f_lm(p) e^{i m phi}) P_l^m(mu)
with P_l^m without the Condon-Shortley phase.  Hence
    f_lm = (2l+1)(l-m)! / (4 pi (l+m)!) * integral f P_l^m e^{-i m phi} dOmega.
Checks of that convention against closed forms are in test_oracle_moments.py
(Vy of a y-drift from (8 pi/3) Re f11, VxVz from -(8 pi/5) Im f21, ...).
"""
from math import factorial

import numpy as np
from scipy.special import lpmv


def project(fun, p, l0, m0, nmu=48, nphi=48):
    """fun(px, py, pz) -> values; returns complex128[nh][len(p)] in ind(l, m) order."""
    mu, wmu = np.polynomial.legendre.leggauss(nmu)
    phi = np.arange(nphi) * (2 * np.pi / nphi)
    P, MU, PHI = np.meshgrid(p, mu, phi, indexing="ij")
    s = np.sqrt(1.0 - MU ** 2)
    F = fun(P * MU, P * s * np.cos(PHI), P * s * np.sin(PHI))
    nh = ((m0 + 1) * (2 * l0 - m0 + 2)) // 2
    res = np.zeros((nh, len(p)), np.complex128)
    for l in range(l0 + 1):
        for m in range(min(l, m0) + 1):
            plm = (-1) ** m * lpmv(m, l, mu)  # no Condon-Shortley phase
            ang = plm[None, :, None] * np.exp(-1j * m * phi)[None, None, :]
            integ = (F * ang).sum(axis=2) * (2 * np.pi / nphi)  # over phi
            integ = (integ * wmu[None, :]).sum(axis=1)           # over mu
            norm = (2 * l + 1) * factorial(l - m) / (4 * np.pi * factorial(l + m))
            i = l * (l + 1) // 2 + m if l < m0 + 1 else l * (m0 + 1) - m0 * (m0 + 1) // 2 + m
            res[i] = norm * integ
    return res


def maxwellian3(n, pt, u=(0.0, 0.0, 0.0)):
    """n (2 pi pt^2)^{-3/2} exp(-|p - u|^2 / (2 pt^2)) as a function of (px, py, pz)."""
    ux, uy, uz = u

    def f(px, py, pz):
        r2 = (px - ux) ** 2 + (py - uy) ** 2 + (pz - uz) ** 2
        return n * np.exp(-r2 / (2 * pt * pt)) / (2 * np.pi * pt * pt) ** 1.5

    return f
