"""CPU restatement of MNPBEM's retarded dipole excitation (TEST INFRASTRUCTURE ONLY).

- dipole potentials phi, phi', A, A' on the particle surface:
  Simulation/retarded/@dipoleret/potential.m:64-122 (Jackson Eq. 9.16);
- Green function between points and the surface with near-field refinement:
  Greenfun/+green/refinematrix.m:21-68 (a point takes the radius of the
  face), Greenfun/@greenret/private/init.m:96-177 (rule-18 Taylor
  coefficients, deriv 'cart'), private/eval1.m:84-107 (Gp);
- induced fields at points: Greenfun/@compgreenret/field.m:14-40;
- dipole far fields and decay rates: @dipoleret/farfield.m:22-58,
  scattering.m:11-13, decayrate.m:13-62, Simulation/retarded/scattering.m:14-31;
- Mie decay rates: Mie/@mieret/decayrate.m:16-58 (Kim et al., Surf. Sci. 195, 1 (1988));
- quasistatic dipoles: Simulation/static/@dipolestat/field.m:50-78,
  potential.m:17, decayrate.m:25-56 (the field of the charges is the point
  field above at k = 0, Greenfun/@compgreenstat/field.m:14-25 with the
  refinement of @greenstat/private/init.m), pinned against the quasistatic
  image series of a sphere (sphere_stat_decayrate).

Dipole columns are ordered (point fastest, then orientation): column
q = ipt + npt * idip, as the reference's (n, npt, ndip) arrays flattened.
"""
import numpy as np

from oracle.mie import _riccati
from oracle.mnpbem_oracle import FACT, quad_flat


def dipole_potential(pos, nvec, pts, dip, enei, eps):
    """phi, phip (n, npol) and a, ap (n, 3, npol) of dipoles ``dip`` (npt, 3, ndip)
    at ``pts`` (npt, 3) in a medium of permittivity eps (potential.m:64-122)."""
    eps = complex(eps)
    k0 = 2 * np.pi / enei
    k = k0 * np.sqrt(eps)
    pts = np.atleast_2d(pts)
    dip = np.asarray(dip, dtype=float).reshape(len(pts), 3, -1)
    ndip = dip.shape[2]
    x = pos[:, None, :] - pts[None, :, :]                 # (n, npt, 3)
    r = np.sqrt(np.sum(x * x, axis=2))
    u = x / r[:, :, None]
    G = np.exp(1j * k * r) / r
    F = (1j * k - 1 / r) * G
    en = np.einsum("ic,ipc->ip", nvec, u)
    n, npt = r.shape
    phi = np.zeros((n, npt, ndip), complex)
    phip = np.zeros_like(phi)
    a = np.zeros((n, 3, npt, ndip), complex)
    ap = np.zeros_like(a)
    for i in range(ndip):
        d = dip[:, :, i]                                   # (npt, 3)
        ep = np.einsum("ipc,pc->ip", u, d)
        np_ = nvec @ d.T
        phi[:, :, i] = -ep * F / eps
        phip[:, :, i] = ((np_ - 3 * en * ep) / r ** 2 * (1 - 1j * k * r) * G / eps
                         + k ** 2 * ep * en * G / eps)
        for c in range(3):
            a[:, c, :, i] = -1j * k0 * d[None, :, c] * G
            ap[:, c, :, i] = -1j * k0 * d[None, :, c] * en * F
    return (phi.reshape(n, -1, order="F"), phip.reshape(n, -1, order="F"),
            a.reshape(n, 3, -1, order="F"), ap.reshape(n, 3, -1, order="F"))


def point_green(green, pts, k, relcutoff=3.0, abscutoff=0.0):
    """G (npt, n) and Gp (npt, 3, n) from faces to points, with the refined
    Taylor form on near pairs (d - rad_j < abscutoff or (d - rad_j)/rad_j < relcutoff)."""
    pts = np.atleast_2d(pts)
    pos, area, rad, order = green.pos, green.area, green.rad, green.order
    x = pts[:, None, :] - pos[None, :, :]                 # (npt, n, 3)
    d = np.sqrt(np.sum(x * x, axis=2))
    e = np.exp(1j * k * d)
    G = e / d * area[None, :]
    f = (1j * k - 1 / d) / d ** 2
    Gp = np.transpose(x * (f * area[None, :] * e)[:, :, None], (0, 2, 1))
    d2 = d - rad[None, :]
    near = ((d2 < abscutoff) | (d2 / rad[None, :] < relcutoff))
    ip, jf = np.nonzero(near)
    if len(ip):
        q, w = quad_flat(green.verts, green.faces, jf)
        xv = pts[ip][:, None, :] - q
        r = np.sqrt(np.sum(xv * xv, axis=2))
        r0 = d[ip, jf]
        dr = r - r0[:, None]
        ikp = (1j * k) ** np.arange(order + 1)
        g = np.stack([np.sum((dr ** o / r) / FACT[o] * w, axis=1) for o in range(order + 1)], axis=1)
        fo = [-np.sum(xv * (w / r ** 3)[:, :, None], axis=1)]
        for o in range(1, order + 1):
            fac = -dr ** o / (r ** 3 * FACT[o]) + dr ** (o - 1) / (r ** 2 * FACT[o - 1])
            fo.append(np.sum(xv * (fac * w)[:, :, None], axis=1))
        fo = np.stack(fo, axis=2)                          # (npairs, 3, order+1)
        ph = np.exp(1j * k * r0)
        G[ip, jf] = (g @ ikp) * ph
        Gp[ip, :, jf] = (fo @ ikp) * ph[:, None]
    return G, Gp


def field_points(green, pts, sig, h, k, enei, relcutoff=3.0, abscutoff=0.0):
    """e, h (npt, 3, ncol) at points from the surface charges/currents of one
    side: e = i k0 G h - Gp sig, h = Gp x h (field.m:19-35)."""
    k0 = 2 * np.pi / enei
    G, Gp = point_green(green, pts, k, relcutoff, abscutoff)
    e = 1j * k0 * np.einsum("pj,jcq->pcq", G, h) - np.einsum("pcj,jq->pcq", Gp, sig)
    hx = (np.einsum("pj,jq->pq", Gp[:, 1], h[:, 2]) - np.einsum("pj,jq->pq", Gp[:, 2], h[:, 1]))
    hy = (np.einsum("pj,jq->pq", Gp[:, 2], h[:, 0]) - np.einsum("pj,jq->pq", Gp[:, 0], h[:, 2]))
    hz = (np.einsum("pj,jq->pq", Gp[:, 0], h[:, 1]) - np.einsum("pj,jq->pq", Gp[:, 1], h[:, 0]))
    return e, np.stack([hx, hy, hz], axis=1)


def dipole_farfield(dirs, pts, dip, enei, eps):
    """Far fields e, h (ndir, 3, npol) of the dipoles themselves (farfield.m:40-55)."""
    eps = complex(eps)
    k = 2 * np.pi / enei * np.sqrt(eps)
    nb = np.sqrt(eps)
    pts = np.atleast_2d(pts)
    dip = np.asarray(dip, dtype=float).reshape(len(pts), 3, -1)
    npt, _, ndip = dip.shape
    g = np.exp(-1j * k * dirs @ pts.T)                    # (ndir, npt)
    e = np.zeros((len(dirs), 3, npt, ndip), complex)
    h = np.zeros_like(e)
    for i in range(ndip):
        for p in range(npt):
            hh = np.cross(dirs, dip[p, :, i][None, :]) * g[:, p, None]
            ee = np.cross(hh, dirs)
            e[:, :, p, i] = k ** 2 * ee / eps
            h[:, :, p, i] = k ** 2 * hh / nb
    return e.reshape(len(dirs), 3, -1, order="F"), h.reshape(len(dirs), 3, -1, order="F")


def radiated_power(e, h, pinfty_nvec, pinfty_area):
    """sum_area 0.5 Re(n . (e x h*)) (Simulation/retarded/scattering.m:18-28)."""
    s = 0.5 * np.real(np.einsum("dc,dcq->dq", pinfty_nvec, np.cross(e, np.conj(h), axis=1)))
    return pinfty_area @ s


def decayrate(green, pts, dip, enei, eps_out, sig2, h2, e_sig_far, h_sig_far,
              pinfty_nvec, pinfty_area, relcutoff=3.0, abscutoff=0.0):
    """(tot, rad, rad0) (npt, ndip) for dipoles outside the particle
    (decayrate.m:13-62); e_sig_far/h_sig_far are the far fields of sig2, h2
    on the pinfty normals (spectrumret/farfield.m)."""
    pts = np.atleast_2d(pts)
    dip = np.asarray(dip, dtype=float).reshape(len(pts), 3, -1)
    npt, _, ndip = dip.shape
    k0 = 2 * np.pi / enei
    k = k0 * np.sqrt(complex(eps_out))
    nb = np.sqrt(complex(eps_out))
    gamma = 4 / 3 * k0 ** 3
    e, _ = field_points(green, pts, sig2, h2, k, enei, relcutoff, abscutoff)
    ed, hd = dipole_farfield(pinfty_nvec, pts, dip, enei, eps_out)
    sca = radiated_power(e_sig_far + ed, h_sig_far + hd, pinfty_nvec, pinfty_area)
    tot = np.zeros((npt, ndip))
    rad = np.zeros((npt, ndip))
    for p in range(npt):
        for i in range(ndip):
            q = p + npt * i
            tot[p, i] = np.real(1 + np.imag(e[p, :, q] @ dip[p, :, i]) / (0.5 * nb * gamma))
            rad[p, i] = np.real(sca[q] / (2 * np.pi * k0) / (0.5 * nb * gamma))
    return tot, rad, np.real(nb * gamma) * np.ones((npt, ndip))


def dipolestat_field(pos, pts, dip, eps):
    """e (n, 3, npt, ndip) at pos of the quasistatic dipoles dip (npt, 3, ndip)
    at pts, each divided by eps[ipt] of its medium (@dipolestat/field.m:50-78)."""
    pts = np.atleast_2d(pts)
    dip = np.asarray(dip).reshape(len(pts), 3, -1)
    eps = np.broadcast_to(np.asarray(eps, dtype=complex), (len(pts),))
    x = pos[:, None, :] - pts[None, :, :]                  # (n, npt, 3)
    r = np.sqrt(np.sum(x * x, axis=2))
    u = x / r[:, :, None]
    e = np.zeros((len(pos), 3, len(pts), dip.shape[2]), complex)
    for i in range(dip.shape[2]):
        p = dip[:, :, i][None, :, :]                       # (1, npt, 3)
        inner = np.sum(u * p, axis=2)
        e[:, :, :, i] = np.transpose((3 * u * inner[:, :, None] - p) / (r ** 3 * eps[None, :])[:, :, None],
                                     (0, 2, 1))
    return e


def dipolestat_potential(nvec, e):
    """phi' (n, npt ndip) = -nvec . e (@dipolestat/potential.m:17), columns
    ipt + npt idip."""
    return -np.einsum("ic,icq->iq", nvec, e.reshape(len(nvec), 3, -1, order="F"))


def dipolestat_decayrate(green, pts, dip, enei, eps_out, sig, relcutoff=3.0, abscutoff=0.0):
    """(tot, rad, rad0) (npt, ndip) of @dipolestat/decayrate.m:25-56: e is the
    field -Gp sig of the surface charges at each dipole (k = 0), indip the
    induced dipole sum area pos sig."""
    pts = np.atleast_2d(pts)
    dip = np.asarray(dip, dtype=float).reshape(len(pts), 3, -1)
    npt, _, ndip = dip.shape
    nb = np.sqrt(complex(eps_out))
    gamma = 4 / 3 * (2 * np.pi / enei) ** 3
    e, _ = field_points(green, pts, sig, np.zeros((green.n, 3, sig.shape[1]), complex), 0.0, enei,
                        relcutoff, abscutoff)
    indip = (green.area[:, None] * green.pos).T @ sig    # (3, ncol)
    tot, rad = np.zeros((npt, ndip)), np.zeros((npt, ndip))
    for p in range(npt):
        for i in range(ndip):
            q = p + npt * i
            rad[p, i] = np.linalg.norm(nb ** 2 * indip[:, q] + dip[p, :, i]) ** 2
            tot[p, i] = np.real(1 + np.imag(e[p, :, q] @ dip[p, :, i]) / (0.5 * nb * gamma))
    return tot, rad, np.real(nb * gamma) * np.ones((npt, ndip))


def sphere_stat_decayrate(enei, eps_in, eps_out, radius, dist, lmax=200):
    """Quasistatic decay rates (tot, rad) of dipoles at distance ``dist`` from
    the centre of a sphere, orientations (radial, tangential), in the
    convention of @dipolestat/decayrate.m (tot = 1 + Im(E_ind . p) / (0.5 nb
    gamma), rad = |p + p_ind|^2 for |p| = 1).  Image series of the
    quasistatic sphere: alpha_l = a^(2l+1) l (e - 1) / (l e + l + 1), e =
    eps_in / eps_out; E_ind . p = sum_l c_l alpha_l / dist^(2l+4) / eps_out
    with c_l = (l+1)^2 (radial) or l (l+1) / 2 (tangential); the induced
    dipole (l = 1) is 2 alpha_1 / dist^3 (radial) or -alpha_1 / dist^3."""
    e = complex(eps_in) / complex(eps_out)
    nb = np.sqrt(complex(eps_out))
    gamma = 4 / 3 * (2 * np.pi / enei) ** 3
    l = np.arange(1, lmax + 1)
    t = radius / dist
    s = t ** (2 * l + 1) * l * (e - 1) / (l * e + l + 1) / dist ** 3   # alpha_l / dist^(2l+4)
    alpha1 = radius ** 3 * (e - 1) / (e + 2)
    eind = np.array([np.sum((l + 1) ** 2 * s), np.sum(l * (l + 1) / 2 * s)]) / complex(eps_out)
    tot = np.real(1 + np.imag(eind) / (0.5 * nb * gamma))
    a1 = alpha1 / dist ** 3
    rad = np.abs(np.array([1 + 2 * a1, 1 - a1])) ** 2
    return tot, rad


def mie_decayrate(enei, eps_in, eps_out, diameter, z, lmax=20):
    """Total and radiative decay rates (len(z), 2) for dipoles on the z axis
    at distances z from the sphere centre, orientations x and z (mieret/decayrate.m).

    The series of decayrate.m is summed in ratio form: with
    A = (j1/h1) a_l and B = (j1/h1) b_l (a_l, b_l built from logarithmic
    derivatives), the products A h(y)^2, B h(y)^2 and A h(y) are formed as
    (j1 h(y)/h1) h(y) and so on, which stay in floating-point range where j1
    underflows and h1, h(y) grow large.  The values are those of the
    reference's expressions; the form only lets lmax go beyond the
    reference's 20 (mieret/private/init.m:23-24), which a dipole a few nm
    from a large sphere needs (the terms fall off like (R/z)^(2l))."""
    k = 2 * np.pi / enei * np.sqrt(complex(eps_out))
    kin = 2 * np.pi / enei * np.sqrt(complex(eps_in))
    epsb, epsin = complex(eps_out), complex(eps_in)
    l = np.arange(1, lmax + 1)
    x, xin = 0.5 * k * diameter, 0.5 * kin * diameter
    j1, h1, zjp1, zhp1 = _riccati(x, lmax)
    j2, _, zjp2, _ = _riccati(xin, lmax)
    rho2, lj1, lh1 = zjp2 / j2, zjp1 / j1, zhp1 / h1
    a = (rho2 - lj1) / (lh1 - rho2)                            # A = (j1 / h1) a
    b = (epsb * rho2 - epsin * lj1) / (epsin * lh1 - epsb * rho2)  # B = (j1 / h1) b
    z = np.atleast_1d(z)
    tot = np.zeros((len(z), 2))
    rad = np.zeros((len(z), 2))
    for i, zz in enumerate(z):
        y = k * zz
        j, h, zjp, zhp = _riccati(y, lmax)
        s, sp = j1 * (h / h1), j1 * (zhp / h1)                 # (j1/h1) h, (j1/h1) zhp
        tot[i, 0] = 1 + 1.5 * np.real(np.sum((l + 0.5) * (b * sp * zhp / y ** 2 + a * s * h)))
        tot[i, 1] = 1 + 1.5 * np.real(np.sum((2 * l + 1) * l * (l + 1) * b * s * h / y ** 2))
        rad[i, 0] = 0.75 * np.sum((2 * l + 1) * (np.abs(j + a * s) ** 2 + np.abs((zjp + b * sp) / y) ** 2))
        rad[i, 1] = 1.5 * np.sum((2 * l + 1) * l * (l + 1) * np.abs((j + b * s) / y) ** 2)
    if not (np.all(np.isfinite(tot)) and np.all(np.isfinite(rad))):
        raise OverflowError(f"mie_decayrate: lmax {lmax} beyond the range of the Bessel functions")
    return tot, rad
