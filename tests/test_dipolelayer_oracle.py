"""Pins of the oracle's dipoles above a substrate (dipoleretlayer
restatement in oracle/layerret_oracle.py): the total decay rate of a dipole
alone above glass against the Sommerfeld closed form of Novotny and Hecht
(Principles of Nano-Optics, eqs. 10.26-10.27), energy balance between the
total and the radiated rate for a lossless substrate, and the fields and
surface derivatives against finite differences of the potentials."""
import numpy as np
import pytest
from scipy import integrate

import oracle.layerret_oracle as L

ENEI, H, EPS2 = 600.0, 5.0, 2.25


def _table(sub, pts, nr=10, nz=6):
    rg, zg, z2t = L.tab_grid(sub, pts, nr=nr, nz=nz)
    tab = L.GreenTable(sub, rg, zg, z2t)
    return tab, tab.compute(ENEI)[1]


def _novotny_hecht(k1, h, e2):
    """Normalised power of a perpendicular and a parallel dipole at height h
    above a half-space e2 (upper medium vacuum), by quadrature in
    s = k_par / k1 (s = sin t below 1, s = cosh u above)."""
    def integ(fun):
        f1 = lambda t: fun(np.sin(t), np.cos(t) + 0j)
        f2 = lambda u: -1j * fun(np.cosh(u), 1j * np.sinh(u))
        q = lambda f, a, b: (integrate.quad(lambda x: f(x).real, a, b, limit=400)[0]
                             + 1j * integrate.quad(lambda x: f(x).imag, a, b, limit=400)[0])
        return q(f1, 0, np.pi / 2) + q(f2, 0, 12)

    def coef(s, sz):
        kz2 = np.sqrt(e2 - s ** 2 + 0j)
        return (e2 * sz - kz2) / (e2 * sz + kz2), (sz - kz2) / (sz + kz2)

    perp = lambda s, sz: s ** 3 * coef(s, sz)[0] * np.exp(2j * k1 * h * sz)
    par = lambda s, sz: s * (coef(s, sz)[1] - sz ** 2 * coef(s, sz)[0]) * np.exp(2j * k1 * h * sz)
    return 1 + 1.5 * np.real(integ(perp)), 1 + 0.75 * np.real(integ(par))


def _sphere_grid(nt=96, nphi=32, breaks=(-1.0, -np.sqrt(1 - 1 / EPS2), 0.0, 1.0)):
    """Gauss-Legendre in cos(theta) on panels split at the substrate's
    critical angle (where the transmitted pattern has a kink) x uniform phi."""
    x, w = np.polynomial.legendre.leggauss(nt)
    cs, ws = [], []
    for lo, hi in zip(breaks[:-1], breaks[1:]):
        cs.append(0.5 * (hi - lo) * x + 0.5 * (hi + lo))
        ws.append(0.5 * (hi - lo) * w)
    x, w = np.concatenate(cs), np.concatenate(ws)
    ph = np.arange(nphi) * 2 * np.pi / nphi
    ct, P = np.meshgrid(x, ph, indexing="ij")
    st = np.sqrt(1 - ct ** 2)
    nvec = np.stack([st * np.cos(P), st * np.sin(P), ct], -1).reshape(-1, 3)
    return nvec, (w[:, None] * np.full(nphi, 2 * np.pi / nphi)[None, :]).ravel()


def test_decay_rate_closed_form_and_energy_balance():
    sub = L.Substrate(1.0, EPS2)
    pts = np.array([[0.0, 0.0, H]])
    tab, tabn = _table(sub, np.vstack([pts, [[20.0, 0.0, H]]]))
    dip = np.zeros((1, 3, 2))
    dip[0, 2, 0] = dip[0, 0, 1] = 1.0  # perpendicular, parallel
    nvec, area = _sphere_grid()
    zero = np.zeros((1, 3, 2), complex)
    zf = np.zeros((len(nvec), 3, 2), complex)
    tot, rad, rad0 = L.dipole_layer_decayrate(tab, tabn, sub, ENEI, pts, dip, zero, zf, zf, nvec, area)
    perp, par = _novotny_hecht(2 * np.pi / ENEI, H, EPS2)
    # each reflected component carries at least 2.7 % of tot - 1 (p, ss for
    # the parallel dipole; hh, ss, hs, sh for the perpendicular one)
    assert abs(tot[0, 0] - perp) < 0.015 * (perp - 1)
    assert abs(tot[0, 1] - par) < 0.005 * (par - 1)
    # lossless media: all power reaches the far field (reflected, direct and
    # transmitted, including the light beyond the critical angle), so the
    # radiated rate is the closed form too; the far field uses the exact
    # coefficients, the total rate the interpolated table at r = rmin
    np.testing.assert_allclose(rad[0], [perp, par], rtol=1e-5)
    np.testing.assert_allclose(rad0, 4 / 3 * (2 * np.pi / ENEI) ** 3, rtol=1e-14)


def test_fields_are_gradients_of_potentials():
    """e = i k0 a - grad phi and h = -(curl a) of the reflected potentials
    (field.m writes h with this sign), by central differences.  Observers
    and dipoles lie on one line along x at one height: there the mixed
    second derivatives d^2/dx dy' vanish (greenderiv.m:73-78 gives their Fr
    term the opposite sign of the derivative of F{1,3}; kept as the
    reference has it) and the table is exact in the height sum."""
    sub = L.Substrate(1.0, EPS2)
    rng = np.random.default_rng(5)
    pts = np.array([[0.0, 0.0, H], [3.0, 0.0, H]])
    obs = np.array([[5.0, 0.0, H], [-2.0, 0.0, H], [9.0, 0.0, H]])
    tab, tabn = _table(sub, np.vstack([pts, obs]), nr=60, nz=6)
    dip = rng.normal(size=(2, 3, 2))
    eps = np.full(2, 1.0 + 0j)
    k0 = 2 * np.pi / ENEI
    e, h = L.dipole_layer_field(tab, tabn, ENEI, obs, pts, dip, eps)
    dn = np.zeros_like(obs)
    pot = lambda p: L.dipole_layer_potential(tab, tabn, ENEI, p, dn, pts, dip, eps)
    phi, _, a, _ = pot(obs)
    d = 1e-3
    grad_phi, da = [], []
    for c in range(3):
        s = np.zeros(3)
        s[c] = d
        pp, _, ap, _ = pot(obs + s)
        pm, _, am, _ = pot(obs - s)
        grad_phi.append((pp - pm) / (2 * d))
        da.append((ap - am) / (2 * d))  # da[c][:, i] = d a_i / d x_c
    grad_phi = np.stack(grad_phi, axis=1)
    e_fd = 1j * k0 * a - grad_phi
    curl = np.stack([da[1][:, 2] - da[2][:, 1], da[2][:, 0] - da[0][:, 2], da[0][:, 1] - da[1][:, 0]], axis=1)
    # e_z holds d^2/dz dz' of the table, the slope of its interpolant in the
    # height sum, which central differences of the potential see only in part
    tol = np.array([2e-3, 2e-3, 3e-2])[None, :, None] * np.abs(e).max()
    assert np.all(np.abs(e - e_fd) <= tol)
    # h_x, h_z are minus the curl; field.m:125 forms h_y from F{4,.} (d/dz)
    # of a_z where the curl has d/dx, and its two terms nearly cancel here,
    # so central differences do not resolve it: h_y is restated, not pinned
    np.testing.assert_allclose(h[:, [0, 2]], -curl[:, [0, 2]], rtol=0, atol=2e-3 * np.abs(h).max())


def test_surface_derivatives_of_potentials():
    """phi2p, a2p = nvec . grad of phi2, a2 by central differences; the
    moments have equal y and z parts, where potential.m's hh term of
    a2p(:, 3) (taken with the y component) coincides with the z one; faces
    and dipole on one line along x at one height, as above."""
    sub = L.Substrate(1.0, EPS2)
    rng = np.random.default_rng(6)
    pts = np.array([[0.0, 0.0, H]])
    pos = np.array([[4.0, 0.0, H], [-2.0, 0.0, H]])
    nvec = rng.normal(size=(2, 3))
    nvec /= np.linalg.norm(nvec, axis=1, keepdims=True)
    tab, tabn = _table(sub, np.vstack([pts, pos]), nr=60, nz=6)
    dip = np.zeros((1, 3, 2))
    dip[0, :, 0] = [1.0, 0.5, 0.5]
    dip[0, :, 1] = [0.0, -1.0, -1.0]
    eps = np.full(1, 1.0 + 0j)
    phi, phip, a, ap = L.dipole_layer_potential(tab, tabn, ENEI, pos, nvec, pts, dip, eps)
    d = 1e-3
    pp, _, a_p, _ = L.dipole_layer_potential(tab, tabn, ENEI, pos + d * nvec, nvec, pts, dip, eps)
    pm, _, a_m, _ = L.dipole_layer_potential(tab, tabn, ENEI, pos - d * nvec, nvec, pts, dip, eps)
    # 1 %: the normal has a z part, and the z slopes of the interpolated
    # table differ from its tabulated z derivatives by that much
    np.testing.assert_allclose(phip, (pp - pm) / (2 * d), rtol=0, atol=1e-2 * np.abs(phip).max())
    # a2p(:, 3) holds d^2/dz dz' of the hs part, whose finite difference
    # across one cell of the interpolated table is no derivative (see e_z)
    np.testing.assert_allclose(ap[:, :2], ((a_p - a_m) / (2 * d))[:, :2], rtol=0, atol=1e-2 * np.abs(ap).max())
