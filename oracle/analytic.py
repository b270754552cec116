"""TEST INFRASTRUCTURE — closed-form acoustic solutions used to pin the oracle.

Only tests/ may import this module.

* ``incoming_wave``                 src/analytical/3d_scattering.jl:4-12
* ``plane_wave_scattering_sphere``  src/analytical/3d_scattering.jl:20-44 (series restated
  with scipy's spherical Bessel functions: (-i)^n (2n+1) [j_n(kr) - j_n'(ka)/h_n'(ka)
  h_n(kr)] P_n(cos theta), summed until the relative increment drops below ``ieps``).
* ``pulsating_sphere_interior``     examples/3d_pulsating_sphere.jl:35
* ``pulsating_sphere_exterior``     textbook monopole: p(r) = Z0 v a/r * ika/(ika-1) e^{ik(r-a)}
* ``sphere_row_sum_G/F``            integrals of G and dG/dn (inward normal) over a sphere
  of radius a for a point on it: (e^{2ika}-1)/(2ik) and -e^{2ika}/2 + (e^{2ika}-1)/(2ika).
"""
from __future__ import annotations

import numpy as np
from scipy import special


def incoming_wave(angles, P0, xyz, k):
    A = np.sin(angles[0]) * np.cos(angles[1])
    B = np.sin(angles[0]) * np.sin(angles[1])
    C = np.cos(angles[0])
    dps = (A * xyz[0] + B * xyz[1] + C * xyz[2]) / np.sqrt(A ** 2 + B ** 2 + C ** 2)
    return P0 * np.exp(-1j * (k * dps))


def _h(n, z):
    return special.spherical_jn(n, z) + 1j * special.spherical_yn(n, z)


def _dh(n, z):
    return special.spherical_jn(n, z, True) + 1j * special.spherical_yn(n, z, True)


def plane_wave_scattering_sphere(k, a, r, theta, ieps=1e-6, nmax=200):
    """Total pressure for the plane wave exp(-i k r cos theta) on a rigid sphere."""
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    ka, kr = k * a, k * r
    ptot = np.zeros(theta.shape, dtype=np.complex128)
    for n in range(nmax):
        apri = special.spherical_jn(n, ka, True) / _dh(n, ka)
        it = (-1j) ** n * (2 * n + 1) * (special.spherical_jn(n, kr) - apri * _h(n, kr)) * \
            special.eval_legendre(n, np.cos(theta))
        ptot = ptot + it
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = np.nanmin(np.nanmax(np.abs(it / ptot)))
        if n > 2 and rel < ieps:
            break
    return ptot


def pulsating_sphere_interior(k, a, r, Z0, vr):
    r = np.asarray(r, dtype=np.float64)
    return 1j * Z0 * vr * (a / r) * (k * a * np.sin(k * r)) / (k * a * np.cos(k * a) - np.sin(k * a))


def pulsating_sphere_exterior(k, a, r, Z0, vr):
    r = np.asarray(r, dtype=np.float64)
    ika = 1j * k * a
    return Z0 * vr * (a / r) * ika / (ika - 1.0) * np.exp(1j * k * (r - a))


def sphere_row_sum_G(k, a):
    k = complex(k)
    if k == 0:
        return complex(a)
    return (np.exp(2j * k * a) - 1.0) / (2j * k)


def sphere_row_sum_F(k, a):
    k = complex(k)
    if k == 0:
        return 0.5
    e = np.exp(2j * k * a)
    return -e / 2.0 + (e - 1.0) / (2j * k * a)
