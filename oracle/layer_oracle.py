"""CPU restatement of MNPBEM's quasistatic solver and plane wave above a
substrate (one interface) (TEST INFRASTRUCTURE ONLY: imported by tests/,
never by the product).

- layer structure: Particles/@layerstructure/layerstructure.m (eps of the
  upper and lower medium, interface at z0), indlayer.m;
- image particle and reflected Green function: Greenfun/@compgreenstatlayer/
  init.m:18-29 (p2r = shift(flip(shift(p2, -z0), 3), z0)), eval.m:12-38
  (f2 = -(eps2 - eps1)/(eps2 + eps1); for a particle above the interface
  F = F_direct + f2 F(p, p2r)), H1 = F + 2 pi, H2 = F - 2 pi (eval.m:63-67);
- solver: BEM/@bemstatlayer/subsref.m:28-32 (mat = -inv(eps1 H1 - eps2 H2)
  (eps1 - eps2)), mldivide.m:13;
- plane wave: Simulation/static/@planewavestatlayer/decompose.m, fresnel.m,
  field.m, potential.m, extinction.m, scattering.m, absorption.m;
  Simulation/static/@spectrumstatlayer/efarfield.m, scattering.m;
- dipoles above the substrate: Simulation/static/@dipolestatlayer/init.m,
  field.m, field2.m, decayrate.m, decayrate0.m.

Pinned by tests/test_layer_oracle.py (see there): equal media reduce every
function to the free-space one, Fresnel coefficients conserve energy and
reduce to the textbook ones at normal incidence, the reflected F equals the
collocation formula at the image centroids away from the interface, and a
small sphere above a substrate follows the coupled image-dipole model.
"""
import numpy as np

from . import mnpbem_oracle as O

MATLAB_EPS = np.finfo(float).eps


def image_mesh(verts, faces, z0):
    """[particle; image], the image flipped at z = z0 with reversed faces."""
    v = np.asarray(verts, float)
    vi = v.copy()
    vi[:, 2] = 2 * z0 - vi[:, 2]
    f = np.asarray(faces)
    return np.vstack([v, vi]), np.vstack([f, f[:, ::-1] + len(v)])


def reflected_F(verts, faces, z0, **kw):
    """F(p, p2r): rows of the particle, columns of its image (k = 0)."""
    V, Fc = image_mesh(verts, faces, z0)
    g = O.GreenRet(V, Fc, closed=False, **kw)
    n = len(faces)
    return g.eval(0.0, ("F",))["F"][:n, n:]


class BemStatLayer:
    def __init__(self, green, Frefl):
        self.g, self.F0, self.Fr = green, green.eval(0.0, ("F",))["F"], Frefl

    def initmat(self, eps1, eps2, eps_up, eps_down):
        f2 = -(eps_down - eps_up) / (eps_down + eps_up)
        F = self.F0 + f2 * self.Fr
        n = self.g.n
        H1, H2 = F + 2 * np.pi * np.eye(n), F - 2 * np.pi * np.eye(n)
        self.mat = -np.linalg.inv(eps1 * H1 - eps2 * H2) * (eps1 - eps2)
        return self

    def solve(self, phip):
        return self.mat @ phip


def decompose(pol, dirs):
    """TE and TM parts of the polarisations (decompose.m)."""
    pol, dirs = np.atleast_2d(pol).astype(float), np.atleast_2d(dirs).astype(float)
    ex, ey = pol[:, 0], pol[:, 1]
    kx, ky = dirs[:, 0], dirs[:, 1]
    kt = np.maximum(np.sqrt(kx ** 2 + ky ** 2), MATLAB_EPS)
    nvec = np.column_stack([ky / kt, -kx / kt, 0 * kx])
    te = nvec * ((ky * ex - kx * ey) / kt)[:, None]
    return te, pol - te


def fresnel(dirs, enei, eps1, eps2):
    """fresnel.m: eps1, eps2 of the upper and lower medium."""
    k0 = 2 * np.pi / enei
    k1, k2 = k0 * np.sqrt(complex(eps1)), k0 * np.sqrt(complex(eps2))
    dirs = np.atleast_2d(dirs).astype(float)
    m = len(dirs)
    out = {key: np.zeros(m, complex) for key in ("re", "te", "rm", "tm")}
    ki, kr, kt = (np.zeros((m, 3), complex) for _ in range(3))
    for i, d in enumerate(dirs):
        if d[2] < 0:
            k1z = -k1 * d[2]
            k2z = np.sqrt(k2 ** 2 - k1 ** 2 + k1z ** 2)
            if k2z.imag < 0:
                k2z = np.conj(k2z)
            out["re"][i] = (k1z - k2z) / (k1z + k2z)
            out["te"][i] = 2 * k1z / (k1z + k2z)
            out["rm"][i] = (eps2 * k1z - eps1 * k2z) / (eps2 * k1z + eps1 * k2z)
            out["tm"][i] = 2 * eps2 * k1z / (eps2 * k1z + eps1 * k2z)
            ki[i] = k1 * d
            kr[i] = [ki[i, 0], ki[i, 1], k1z]
            kt[i] = [ki[i, 0], ki[i, 1], -k2z]
        else:
            k2z = k2 * d[2]
            k1z = np.sqrt(k1 ** 2 - k2 ** 2 + k2z ** 2)
            if k1z.imag < 0:
                k1z = np.conj(k1z)
            out["re"][i] = (k2z - k1z) / (k2z + k1z)
            out["te"][i] = 2 * k2z / (k2z + k1z)
            out["rm"][i] = (eps1 * k2z - eps2 * k1z) / (eps1 * k2z + eps2 * k1z)
            out["tm"][i] = 2 * eps1 * k2z / (eps1 * k2z + eps2 * k1z)
            ki[i] = k2 * d
            kr[i] = [ki[i, 0], ki[i, 1], -k2z]
            kt[i] = [ki[i, 0], ki[i, 1], k1z]
    out.update(ki=ki, kr=kr, kt=kt)
    return out


def field(pol, dirs, enei, eps1, eps2):
    """(e above, e below) per polarisation (field.m:25-47), each npol x 3."""
    pol = np.atleast_2d(pol).astype(float)
    k0 = 2 * np.pi / enei
    te, tm = decompose(pol, dirs)
    r = fresnel(dirs, enei, eps1, eps2)
    tm = np.cross(r["ki"], tm) / k0
    up, down = [], []
    for i, d in enumerate(np.atleast_2d(dirs)):
        er = np.cross(tm[i], r["kr"][i]) * k0 / np.sum(r["kr"][i] ** 2)
        et = np.cross(tm[i], r["kt"][i]) * k0 / np.sum(r["kt"][i] ** 2)
        refl = pol[i] + r["re"][i] * te[i] + r["rm"][i] * er
        trans = r["te"][i] * te[i] + r["tm"][i] * et
        up.append(refl if d[2] < 0 else trans)
        down.append(trans if d[2] < 0 else refl)
    return np.array(up), np.array(down)


def potential(nvec, e_up):
    """phi' = -nvec . e for faces above the interface (potential.m:13-14)."""
    return -nvec @ np.atleast_2d(e_up).T


def dipole_moment(pos, area, sig):
    """npol x 3: sum_i area_i pos_i sig_i (extinction.m:13)."""
    return ((area[:, None] * pos).T @ sig).T


def efarfield(dip, enei, dirs, eps1, eps2):
    """efarfield.m: far field of dipoles dip (npol x 3) above the interface
    in the directions dirs (ndir x 3); returns (ndir x 3 x npol, k per dir)."""
    k0 = 2 * np.pi / enei
    k1, k2 = k0 * np.sqrt(complex(eps1)), k0 * np.sqrt(complex(eps2))
    dirs = np.atleast_2d(dirs)
    dip = np.atleast_2d(dip)
    d1, d2, d3 = dip[:, 0], dip[:, 1], dip[:, 2]
    field = np.zeros((len(dirs), 3, len(dip)), complex)
    k = np.zeros(len(dirs), complex)
    k[dirs[:, 2].real > 0] = k1
    k[dirs[:, 2].real < 0] = k2
    real_dir = ~np.any(np.abs(np.imag(dirs)) > 1e-10, axis=1)

    def angles(D):
        D = np.real(D)
        phi = np.arctan2(D[:, 1], D[:, 0])
        theta = np.pi / 2 - np.arctan2(D[:, 2], np.hypot(D[:, 0], D[:, 1]))
        return np.sin(phi), np.sin(theta), np.cos(phi), np.cos(theta)

    def fill(ind, p1, p2, p3, sinp, sint, cosp, cost):
        et = np.outer(cost * p2, 1) * (np.outer(cosp, d1) + np.outer(sinp, d2)) - np.outer(sint * p1, d3)
        ep = -np.outer(p3, 1) * (np.outer(sinp, d1) - np.outer(cosp, d2))
        field[ind, 0] = et * (cost * cosp)[:, None] - ep * sinp[:, None]
        field[ind, 1] = et * (cost * sinp)[:, None] + ep * cosp[:, None]
        field[ind, 2] = -et * sint[:, None]

    ind1 = np.flatnonzero((np.real(dirs[:, 2]) >= 0) & real_dir)
    if abs(k1.imag) < 1e-10 and len(ind1):
        sinp, sint, cosp, cost = angles(dirs[ind1])
        k1z = k1 * np.real(dirs[ind1, 2])
        k2z = np.sqrt(k2 ** 2 - k1 ** 2 + k1z ** 2 + 0j)
        rte = (k1z - k2z) / (k1z + k2z)
        rtm = (eps2 * k1z - eps1 * k2z) / (eps2 * k1z + eps1 * k2z)
        fill(ind1, 1 + rtm, 1 - rtm, 1 + rte, sinp, sint, cosp, cost)
    ind2 = np.flatnonzero((np.real(dirs[:, 2]) < 0) & real_dir)
    if abs(k2.imag) < 1e-10 and len(ind2):
        sinp, sint, cosp, cost = angles(dirs[ind2])
        k2z = -k2 * np.real(dirs[ind2, 2])
        k1z = np.sqrt(k1 ** 2 - k2 ** 2 + k2z ** 2 + 0j) + 1e-10j
        k1z = k1z * np.sign(k1z.imag)
        st = k1z / k2
        tte = 2 * k1z / (k1z + k2z)
        ttm = 2 * eps2 * k1z / (eps2 * k1z + eps1 * k2z) * np.sqrt(complex(eps1) / eps2)
        p1 = np.sqrt(complex(eps2) / eps1) * cost / st * ttm
        p2 = -np.sqrt(complex(eps2) / eps1) * ttm
        p3 = cost / st * tte
        fill(ind2, -p1, -p2, -p3, sinp, sint, cosp, cost)
    return k1 ** 2 * field, k


def extinction(dip, pol, dirs, enei, eps1, eps2):
    """extinction.m (key 'full'): reflected plus transmitted parts."""
    k0 = 2 * np.pi / enei
    te, tm = decompose(pol, dirs)
    r = fresnel(dirs, enei, eps1, eps2)
    tm = np.cross(r["ki"], tm) / k0
    ext = np.zeros(len(np.atleast_2d(dirs)))
    for i in range(len(ext)):
        kr, kt = r["kr"][i], r["kt"][i]
        er = r["re"][i] * te[i] + r["rm"][i] * np.cross(tm[i], kr) * k0 / np.vdot(kr, kr)
        et = r["te"][i] * te[i] + r["tm"][i] * np.cross(tm[i], kt) * k0 / np.vdot(kt, kt)
        esr = efarfield(dip[i:i + 1], enei, (kr / np.linalg.norm(kr))[None, :], eps1, eps2)[0][0, :, 0]
        est = efarfield(dip[i:i + 1], enei, (kt / np.linalg.norm(kt))[None, :], eps1, eps2)[0][0, :, 0]
        if abs(kt[2].imag) > 1e-10:
            est = 0 * est
        ext[i] = (4 * np.pi / np.linalg.norm(kr) * np.vdot(er, esr).imag
                  + 4 * np.pi / np.linalg.norm(kr) * np.vdot(et, est).imag)
    return ext


def scattering(dip, dirs, enei, eps1, eps2, pinf_nvec, pinf_area):
    """spectrumstatlayer/scattering.m over all of pinfty, then
    planewavestatlayer/scattering.m's normalisation by 0.5 nb."""
    k0 = 2 * np.pi / enei
    f, k = efarfield(dip, enei, pinf_nvec, eps1, eps2)
    dsca = np.sum(np.abs(f) ** 2, axis=1) * (0.5 * k / k0)[:, None]
    dsca = np.where(np.abs(np.imag(dsca)) > 0, 0, dsca).real
    sca = pinf_area @ dsca
    nb = np.sqrt(np.array([eps1, eps2], dtype=complex))
    d = np.atleast_2d(dirs)[:, 2]
    sca = sca.astype(complex)
    sca[d < 0] /= 0.5 * nb[0]
    sca[d > 0] /= 0.5 * nb[-1]
    return np.real(sca)


# ------------------------------------------------ dipoles above the substrate
def image_factor(eps_up, eps_down):
    """q1 = -(eps_down - eps_up) / (eps_down + eps_up) (@dipolestatlayer/field.m:23)."""
    return -(eps_down - eps_up) / (eps_down + eps_up)


def dipolestatlayer_field(pos, pts, dip, eps_up, eps_down, z0, image_eps="reference"):
    """e (n, 3, npt, ndip) at pos above the interface: the dipoles' field plus
    q1 times the field of the image dipoles (point mirrored at z0, p_z
    negated; @dipolestatlayer/init.m:35-39, field.m:23-28).  The image point
    lies in the lower medium, so @dipolestat/field.m divides its field by
    eps_down (image_eps 'reference'); 'upper' divides by eps_up, the
    textbook image of a dipole in the upper medium."""
    from .dipole_oracle import dipolestat_field
    pts = np.atleast_2d(pts)
    dip = np.asarray(dip, dtype=float).reshape(len(pts), 3, -1)
    ipts = pts.copy()
    ipts[:, 2] = 2 * z0 - ipts[:, 2]
    idip = dip.copy()
    idip[:, 2, :] = -idip[:, 2, :]
    e_img = eps_down if image_eps == "reference" else eps_up
    return (dipolestat_field(pos, pts, dip, eps_up)
            + image_factor(eps_up, eps_down) * dipolestat_field(pos, ipts, idip, e_img))


def _layer_rad(p, enei, eps_up, eps_down, pinf_nvec, pinf_area):
    """(sca / (2 pi k0)) / (0.5 nb gamma) of the dipole p (3,) above the
    interface over all of pinfty (@spectrumstatlayer/scattering.m:13-38 with
    no medium selected; @dipolestatlayer/decayrate.m:54-57)."""
    k0 = 2 * np.pi / enei
    nb = np.sqrt(complex(eps_up))
    gamma = 4 / 3 * k0 ** 3
    f, k = efarfield(p[None, :], enei, pinf_nvec, eps_up, eps_down)
    dsca = np.sum(np.abs(f) ** 2, axis=1)[:, 0] * (0.5 * k / k0)
    dsca = np.where(np.abs(np.imag(dsca)) > 0, 0, dsca).real
    return (pinf_area @ dsca) / (2 * np.pi * k0) / (0.5 * nb * gamma)


def dipolestatlayer_decayrate(green, gimage, pts, dip, enei, eps_up, eps_down, sig, pinf_nvec, pinf_area):
    """(tot, rad, rad0) (npt, ndip) of @dipolestatlayer/decayrate.m:25-61:
    e is the field of the surface charges at each dipole, -(Gp + f2 Gp_r) sig
    (@compgreenstatlayer/field.m with its eval), gimage the Green function of
    [particle; image]; rad from the induced dipole plus dip / nb^2 radiating
    above the interface; tot = 1 + rad + Im(e . dip) / (0.5 nb gamma)."""
    from .dipole_oracle import field_points
    pts = np.atleast_2d(pts)
    dip = np.asarray(dip, dtype=float).reshape(len(pts), 3, -1)
    npt, _, ndip = dip.shape
    n, ncol = green.n, sig.shape[1]
    k0 = 2 * np.pi / enei
    nb = np.sqrt(complex(eps_up))
    gamma = 4 / 3 * k0 ** 3
    f2 = image_factor(eps_up, eps_down)
    e, _ = field_points(green, pts, sig, np.zeros((n, 3, ncol), complex), 0.0, enei)
    sig_r = np.vstack([np.zeros_like(sig), sig])
    er, _ = field_points(gimage, pts, sig_r, np.zeros((2 * n, 3, ncol), complex), 0.0, enei)
    e = e + f2 * er
    indip = (green.area[:, None] * green.pos).T @ sig
    tot, rad = np.zeros((npt, ndip)), np.zeros((npt, ndip))
    for p in range(npt):
        for i in range(ndip):
            q = p + npt * i
            rad[p, i] = np.real(_layer_rad(indip[:, q] + dip[p, :, i] / nb ** 2, enei, eps_up, eps_down,
                                           pinf_nvec, pinf_area))
            tot[p, i] = np.real(1 + rad[p, i] + np.imag(e[p, :, q] @ dip[p, :, i]) / (0.5 * nb * gamma))
    return tot, rad, np.real(nb * gamma) * np.ones((npt, ndip))


def dipolestatlayer_decayrate0(pts, dip, enei, eps_up, eps_down, z0, pinf_nvec, pinf_area, image_eps="reference"):
    """(tot, rad, rad0) without a particle (@dipolestatlayer/decayrate0.m:
    18-56): the reflected field q1 E_image at each dipole; tot = rad for real
    media, else rad + Im(e . dip) / (0.5 nb gamma)."""
    from .dipole_oracle import dipolestat_field
    pts = np.atleast_2d(pts)
    dip = np.asarray(dip, dtype=float).reshape(len(pts), 3, -1)
    npt, _, ndip = dip.shape
    k0 = 2 * np.pi / enei
    nb = np.sqrt(complex(eps_up))
    gamma = 4 / 3 * k0 ** 3
    ipts = pts.copy()
    ipts[:, 2] = 2 * z0 - ipts[:, 2]
    idip = dip.copy()
    idip[:, 2, :] = -idip[:, 2, :]
    e = image_factor(eps_up, eps_down) * dipolestat_field(
        pts, ipts, idip, eps_down if image_eps == "reference" else eps_up)
    real = np.imag(eps_up) == 0 and np.imag(eps_down) == 0
    tot, rad = np.zeros((npt, ndip)), np.zeros((npt, ndip))
    for p in range(npt):
        for i in range(ndip):
            rad[p, i] = np.real(_layer_rad(dip[p, :, i] / nb ** 2, enei, eps_up, eps_down, pinf_nvec, pinf_area))
            tot[p, i] = rad[p, i] if real else np.real(
                rad[p, i] + np.imag(e[p, :, p, i] @ dip[p, :, i]) / (0.5 * nb * gamma))
    return tot, rad, np.real(nb * gamma) * np.ones((npt, ndip))
