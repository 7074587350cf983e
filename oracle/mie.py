"""Retarded Mie theory for a homogeneous sphere (TEST INFRASTRUCTURE ONLY).

Restates Mie/@mieret: miecoefficients.m:14-31 (Bohren-Huffman a_l, b_l with
mu_r = 1), riccatibessel.m:17-34, extinction.m:9-29, scattering.m:9-29 and
lmax = 20 (private/init.m:23-24).  Spherical Bessel functions come from
scipy.special instead of Matlab's besselj/bessely.
"""
import numpy as np
from scipy.special import spherical_jn, spherical_yn


def _riccati(z, lmax):
    """Spherical j_l, h_l (l=1..lmax) and derivatives [z j_l(z)]', [z h_l(z)]'."""
    l = np.arange(1, lmax + 1)
    j = spherical_jn(l, z)
    y = spherical_yn(l, z)
    j0 = np.sin(z) / z
    y0 = -np.cos(z) / z
    h = j + 1j * y
    h0 = j0 + 1j * y0
    zjp = z * np.concatenate([[j0], j[:-1]]) - l * j
    zhp = z * np.concatenate([[h0], h[:-1]]) - l * h
    return j, h, zjp, zhp


def mie_coefficients(k, diameter, epsr, lmax=20):
    """a_l, b_l for relative permittivity epsr and mu_r = 1 (miecoefficients.m:42-52)."""
    nr = np.sqrt(complex(epsr))
    j1, _, zjp1, _ = _riccati(nr * k * diameter / 2, lmax)
    j2, h2, zjp2, zhp2 = _riccati(k * diameter / 2 + 0j, lmax)
    a = (nr ** 2 * j1 * zjp2 - j2 * zjp1) / (nr ** 2 * j1 * zhp2 - h2 * zjp1)
    b = (j1 * zjp2 - j2 * zjp1) / (j1 * zhp2 - h2 * zjp1)
    return a, b


def mie_cross_sections(enei, eps_in, eps_out, diameter, lmax=20):
    """(ext, sca, abs) in nm^2 for each wavelength (eps_in per wavelength)."""
    enei = np.atleast_1d(np.asarray(enei, dtype=float))
    eps_in = np.broadcast_to(np.asarray(eps_in, dtype=complex), enei.shape)
    nb = np.sqrt(complex(eps_out))
    ext = np.zeros(len(enei))
    sca = np.zeros(len(enei))
    l = np.arange(1, lmax + 1)
    for i, (lam, ein) in enumerate(zip(enei, eps_in)):
        k = (2 * np.pi / lam * nb).real
        a, b = mie_coefficients(k, diameter, ein / eps_out, lmax)
        ext[i] = 2 * np.pi / k ** 2 * np.sum((2 * l + 1) * np.real(a + b))
        sca[i] = 2 * np.pi / k ** 2 * np.sum((2 * l + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2))
    return ext, sca, ext - sca


def miestat_cross_sections(enei, eps_in, eps_out, diameter):
    """Quasistatic sphere (Mie/@miestat/scattering.m:13-27, absorption.m,
    extinction.m): alpha = (eps_z - 1)/(eps_z + 2) a^3, sca = 8 pi/3 k^4 |alpha|^2,
    abs = 4 pi k Im(alpha)."""
    enei = np.atleast_1d(np.asarray(enei, dtype=float))
    eps_in = np.broadcast_to(np.asarray(eps_in, dtype=complex), enei.shape)
    k = (2 * np.pi / enei * np.sqrt(complex(eps_out))).real
    epsz = eps_in / eps_out
    alpha = (epsz - 1) / (epsz + 2) * (diameter / 2) ** 3
    sca = 8 * np.pi / 3 * k ** 4 * np.abs(alpha) ** 2
    ab = 4 * np.pi * k * np.imag(alpha)
    return sca + ab, sca, ab
