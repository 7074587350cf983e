"""Particle above a substrate: one planar interface between an upper and a
lower medium (MNPBEM Particles/@layerstructure), quasistatic
(BEM/@bemstatlayer, Simulation/static/@planewavestatlayer, @spectrumstatlayer)
and retarded (Greenfun/@greentablayer, @compgreentablayer, @greenretlayer,
@compgreenretlayer, BEM/@bemretlayer, Simulation/retarded/@planewaveretlayer,
@spectrumretlayer).

Quasistatic: the solver runs on the GPU (mnp_bemstatlayer_* in
include/mnpbem_hip.h); the reflected Green function is that of the particle's
image in the interface, weighted per wavelength by f2 = -(eps_down -
eps_up)/(eps_down + eps_up).  The plane wave's field at the particle (Fresnel
reflection of the incoming wave) and the cross sections from the induced
dipole moment are host-side formulas over a handful of vectors, as in the
reference.

Retarded: the Sommerfeld integrals of the reflected Green function are
tabulated on the GPU (mnp_layertab_*), interpolated into the reflected
matrices with the near-pair refinement (mnp_greenretlayer_*) and solved with
the chain of bemretlayer in complex FP64 (mnp_bemretlayer_*); the plane
wave's potentials and the far fields with the interface are GPU kernels, the
cross sections FP64 host sums over them.

Supported: one interface; particles whose faces all lie above it (the
reference also treats faces lying in the interface, which are not built).
"""
import ctypes as C

import numpy as np

from . import _lib
from . import device as dev
from .bem import bemret, bemretiter, bemstat, compstruct, dipoleret, dipolestat, _flat_normals
from .greenfun import CompGreenRet

_MATLAB_EPS = float(np.finfo(float).eps)


class layerstructure:
    """layerstructure(epstab, ind, z, **options): media epstab[ind[0]-1] above
    and epstab[ind[1]-1] below the interface z (Particles/@layerstructure/
    layerstructure.m:27-31); options ztol, rmin, zmin, semi, ratio as
    layerstructure.options (options.m:22-29)."""

    def __init__(self, epstab, ind, z, ztol=2e-2, rmin=1e-2, zmin=1e-2, semi=0.1, ratio=2.0, **_ignored):
        z = np.atleast_1d(np.asarray(z, dtype=np.float64))
        if len(ind) != 2 or len(z) != 1:
            raise NotImplementedError("layerstructure: one interface (two media) is supported")
        self.ind = list(ind)
        self.eps = [epstab[i - 1] for i in self.ind]
        self.z = z
        self.ztol, self.rmin, self.zmin = float(ztol), float(rmin), float(zmin)
        self.semi, self.ratio = float(semi), float(ratio)

    @staticmethod
    def options(**kw):
        """layerstructure.options: the defaults, overridden by kw."""
        op = dict(ztol=2e-2, rmin=1e-2, zmin=1e-2, semi=0.1, ratio=2.0)
        op.update(kw)
        return op

    def c_options(self):
        return _lib.LayerOptions(float(self.z[0]), self.ztol, self.rmin, self.zmin, self.semi, self.ratio)

    def mindist(self, z):
        return np.abs(np.asarray(z, dtype=np.float64) - self.z[0])

    def round(self, z, zmin=None):
        """round.m:13-24: points closer than zmin to the interface move to zmin."""
        zmin = self.zmin if zmin is None else zmin
        z = np.array(z, dtype=np.float64, copy=True)
        d = np.abs(z - self.z[0])
        sel = d <= zmin
        z[sel] = self.z[0] + np.sign(z[sel] - self.z[0]) * zmin
        return z

    def tabspace(self, p, pt=None, nr=30, nz=30, scale=1.05, dip=None):
        """Table grid for the particle p above the interface, and for the
        points pt (a compoint or an npt x 3 array above the interface) when
        given (tabspace2.m:44-75 with limits and adjust.m:18-33, tabspace1.m:
        24-48 with 'log' spacing, greentablayer/private/init.m:8-14): dict r,
        z1, z2.  Dipole positions dip (npt x 3) are sources as well as
        observers (tabspace(layer, {p, pt}) of the dipole demos)."""
        v = np.asarray(p.verts, dtype=np.float64)
        z0 = self.z[0]
        if np.any(v[:, 2] <= z0):
            raise NotImplementedError("tabspace: vertices at or below the interface are not supported")
        obs = v  # observers (rows) of the table: the vertices and the points
        for q in (pt, dip):
            if q is None:
                continue
            q = np.atleast_2d(np.asarray(getattr(q, "pos", q), dtype=np.float64))
            if np.any(q[:, 2] <= z0):
                raise NotImplementedError("tabspace: points at or below the interface are not supported")
            obs = np.vstack([obs, q])
        if dip is not None:  # sources (columns): the vertices and the dipoles
            v = np.vstack([v, np.atleast_2d(np.asarray(dip, dtype=np.float64))])
        xy, oxy = v[:, :2], obs[:, :2]
        rmax = 0.0
        for i in range(0, len(oxy), 2048):  # largest observer-vertex distance, in chunks
            d2 = np.sum((oxy[i:i + 2048, None, :] - xy[None, :, :]) ** 2, axis=-1)
            rmax = max(rmax, float(np.sqrt(d2.max())))
        m = 0.5 * rmax
        r = [max(m - 0.5 * scale * rmax, 0.0), m + 0.5 * scale * rmax]

        def enlarge(z):
            zl, zh = float(z.min()), float(z.max())
            if abs(zl - zh) < 1e-10:
                zr = [zl - 1e-2, zl + 1e-2]
            else:
                mz = 0.5 * (zl + zh)
                zr = [mz + 0.5 * scale * (zl - zh), mz - 0.5 * scale * (zl - zh)]
            if zr[0] <= z0:  # adjust.m:45 has <, which leaves a bound exactly at z0 (log10(0) later)
                zr[0] = z0 + 1e-10
            return zr

        za, zb = enlarge(obs[:, 2]), enlarge(v[:, 2])
        z1 = [za[0] + zb[0] - z0, za[1] + zb[1] - z0]
        rg = np.logspace(np.log10(max(r[0], self.rmin)), np.log10(r[1]), nr)
        zz = np.sort(self.round(np.array(z1), zmin=0.999 * self.zmin))
        if abs(zz[0] - zz[1]) < 1e-3:
            raise NotImplementedError("tabspace: a flat particle at the interface is not supported")
        zg = z0 + np.logspace(np.log10(zz[0] - z0), np.log10(zz[1] - z0), nz)
        return {"r": np.unique(np.maximum(self.rmin, rg)), "z1": np.unique(self.round(zg)),
                "z2": float(self.round(np.array([z0 + 1e-10]))[0])}

    @property
    def n(self):
        return len(self.z)

    def epsilon(self, enei):
        """(eps above, eps below) at the wavelength."""
        return complex(self.eps[0](enei)[0]), complex(self.eps[1](enei)[0])

    def indlayer(self, z):
        """(medium index 1 = above / 2 = below, in-interface mask) (indlayer.m)."""
        z = np.asarray(z, dtype=np.float64)
        return np.where(z > self.z[0], 1, 2), np.abs(z - self.z[0]) < self.ztol


def _image_mesh(verts, faces, z0):
    """[particle; image]: vertices mirrored at z = z0, faces reversed
    (compgreenstatlayer/init.m:21-24 with @particle/flip.m)."""
    v = np.asarray(verts, dtype=np.float64)
    vi = v.copy()
    vi[:, 2] = 2.0 * z0 - vi[:, 2]
    f = np.asarray(faces, dtype=np.int64)
    return np.vstack([v, vi]), np.vstack([f, f[:, ::-1] + len(v)])


def _check_above(p, layer):
    pos = p.green.geometry()["pos"]
    med, inside = layer.indlayer(pos[:, 2])
    if np.any(inside):
        raise NotImplementedError("substrate: faces lying in the interface are not supported")
    if np.any(med != 1):
        raise ValueError("substrate: the particle must lie above the interface")


class bemstatlayer(bemstat):
    """Quasistatic BEM solver above a substrate (BEM/@bemstatlayer), op['layer']
    a layerstructure; precision as for bemstat."""

    def __init__(self, p, op=None, precision=None):
        op = p.op if op is None else op
        self.layer = op["layer"]
        _check_above(p, self.layer)
        self.p = p
        self.precision = precision or op.get("precision", "fp64")
        if self.precision not in bemret._SOLVERS:
            raise ValueError(f"bemstatlayer: unknown precision {self.precision!r}")
        V, F = _image_mesh(p.verts, p.faces, self.layer.z[0])
        self.gimage = CompGreenRet(V, F, order=op.get("order", 5), rel_cutoff=op.get("RelCutoff", 0.0),
                                   abs_cutoff=op.get("AbsCutoff", 0.0),
                                   npol=tuple(np.broadcast_to(op.get("npol", (7, 5)), (2,))), closed=False)
        h = C.c_void_p()
        _lib.call("mnp_bemstatlayer_create", p.green.handle, self.gimage.handle, bemret._SOLVERS[self.precision],
                  C.byref(h))
        self.handle = h
        self.enei = None

    def __call__(self, enei):
        if self.enei != enei:
            self.enei = None  # stays unset if initmat fails part-way
            eu, ed = self.layer.epsilon(enei)
            _lib.call("mnp_bemstatlayer_initmat", self.handle, float(enei), _lib.MnpComplex.of(self.p.eps1(enei)),
                      _lib.MnpComplex.of(self.p.eps2(enei)), _lib.MnpComplex.of(eu), _lib.MnpComplex.of(ed))
            self.enei = enei
        return self


def _decompose(pol, dirs):
    """TE part (along z x k) and TM rest of each polarisation (decompose.m)."""
    kx, ky = dirs[:, 0], dirs[:, 1]
    kt = np.maximum(np.hypot(kx, ky), _MATLAB_EPS)
    s = (ky * pol[:, 0] - kx * pol[:, 1]) / kt
    te = np.column_stack([ky / kt * s, -kx / kt * s, np.zeros_like(s)])
    return te, pol - te


def _fresnel(dirs, k0, eu, ed):
    """Fresnel coefficients and wave vectors (fresnel.m) per direction."""
    k1, k2 = k0 * np.sqrt(eu), k0 * np.sqrt(ed)
    res = []
    for d in dirs:
        if d[2] < 0:  # incoming from above
            kiz = -k1 * d[2]
            kz = np.sqrt(k2 * k2 - k1 * k1 + kiz * kiz)
            kz = np.conj(kz) if kz.imag < 0 else kz
            ki = k1 * d
            coef = ((kiz - kz) / (kiz + kz), 2 * kiz / (kiz + kz),
                    (ed * kiz - eu * kz) / (ed * kiz + eu * kz), 2 * ed * kiz / (ed * kiz + eu * kz))
            kr, kt = np.array([ki[0], ki[1], kiz]), np.array([ki[0], ki[1], -kz])
        else:  # incoming from below
            kiz = k2 * d[2]
            kz = np.sqrt(k1 * k1 - k2 * k2 + kiz * kiz)
            kz = np.conj(kz) if kz.imag < 0 else kz
            ki = k2 * d
            coef = ((kiz - kz) / (kiz + kz), 2 * kiz / (kiz + kz),
                    (eu * kiz - ed * kz) / (eu * kiz + ed * kz), 2 * eu * kiz / (eu * kiz + ed * kz))
            kr, kt = np.array([ki[0], ki[1], -kiz]), np.array([ki[0], ki[1], kz])
        res.append((coef, ki.astype(complex), kr.astype(complex), kt.astype(complex)))
    return res


def _dipole_far(dip, dirs, k0, eu, ed):
    """Far field of the dipoles dip (npol x 3) above the interface in the
    directions dirs (ndir x 3), and the wavenumber per direction
    (spectrumstatlayer/efarfield.m)."""
    k1, k2 = k0 * np.sqrt(eu), k0 * np.sqrt(ed)
    dirs = np.asarray(dirs)
    real = ~np.any(np.abs(np.imag(dirs)) > 1e-10, axis=1)  # complex directions carry no far field
    dirs = np.real(dirs)
    dz = dirs[:, 2]
    phi = np.arctan2(dirs[:, 1], dirs[:, 0])
    theta = np.pi / 2 - np.arctan2(dz, np.hypot(dirs[:, 0], dirs[:, 1]))
    sp, cp, st, ct = np.sin(phi), np.cos(phi), np.sin(theta), np.cos(theta)
    f1 = np.zeros(len(dirs), complex)
    f2 = np.zeros(len(dirs), complex)
    f3 = np.zeros(len(dirs), complex)
    up, lo = (dz >= 0) & real, (dz < 0) & real
    if abs(k1.imag) < 1e-10 and up.any():
        k1z = k1 * dz[up]
        k2z = np.sqrt(k2 * k2 - k1 * k1 + k1z * k1z + 0j)
        rte = (k1z - k2z) / (k1z + k2z)
        rtm = (ed * k1z - eu * k2z) / (ed * k1z + eu * k2z)
        f1[up], f2[up], f3[up] = 1 + rtm, 1 - rtm, 1 + rte
    if abs(k2.imag) < 1e-10 and lo.any():
        k2z = -k2 * dz[lo]
        k1z = np.sqrt(k1 * k1 - k2 * k2 + k2z * k2z + 0j) + 1e-10j
        k1z = k1z * np.sign(k1z.imag)
        s = k1z / k2
        tte = 2 * k1z / (k1z + k2z)
        ttm = 2 * ed * k1z / (ed * k1z + eu * k2z) * np.sqrt(eu / ed)
        r = np.sqrt(ed / eu)
        f1[lo], f2[lo], f3[lo] = -r * ct[lo] / s * ttm, r * ttm, -ct[lo] / s * tte
    # e_theta, e_phi per (direction, dipole)
    eth = (np.outer(cp, dip[:, 0]) + np.outer(sp, dip[:, 1])) * (ct * f2)[:, None] - np.outer(st * f1, dip[:, 2])
    eph = -(np.outer(sp, dip[:, 0]) - np.outer(cp, dip[:, 1])) * f3[:, None]
    e = np.stack([eth * (ct * cp)[:, None] - eph * sp[:, None], eth * (ct * sp)[:, None] + eph * cp[:, None],
                  -eth * st[:, None]], axis=1)
    k = np.where(dz > 0, k1, np.where(dz < 0, k2, 0))
    return k1 * k1 * e, k


class planewavestatlayer:
    """Quasistatic plane wave above a substrate (Simulation/static/
    @planewavestatlayer): pol and dir npol x 3, op['layer'] a layerstructure;
    scattering is integrated over ``pinfty`` = (nvec, area) of the unit
    sphere (default: the 256-point sphere of trisphere(256, 2))."""

    def __init__(self, pol, dir, op=None, pinfty=None):
        self.pol = np.atleast_2d(np.asarray(pol, dtype=np.float64))
        self.dir = np.atleast_2d(np.asarray(dir, dtype=np.float64))
        if np.any(np.sum(self.pol * self.dir, axis=1) != 0):
            raise ValueError("polarisation must be perpendicular to the propagation direction")
        if len(self.pol) > 64:
            raise ValueError("planewavestatlayer: at most 64 polarisations")
        self.layer = op["layer"]
        if pinfty is None:
            from . import shapes
            pinfty = _flat_normals(*shapes.trisphere256(2.0))
        self.pinf_nvec, self.pinf_area = (np.asarray(a, dtype=np.float64) for a in pinfty)

    def fields(self, enei):
        """(e above, e below), npol x 3 each (field.m:25-47)."""
        k0 = 2 * np.pi / enei
        eu, ed = self.layer.epsilon(enei)
        te, tm = _decompose(self.pol, self.dir)
        above, below = [], []
        for i, (coef, ki, kr, kt) in enumerate(_fresnel(self.dir, k0, eu, ed)):
            re, tte, rm, ttm = coef
            m = np.cross(ki, tm[i]) / k0
            er = np.cross(m, kr) * k0 / np.sum(kr * kr)
            et = np.cross(m, kt) * k0 / np.sum(kt * kt)
            refl = self.pol[i] + re * te[i] + rm * er
            trans = tte * te[i] + ttm * et
            above.append(refl if self.dir[i, 2] < 0 else trans)
            below.append(trans if self.dir[i, 2] < 0 else refl)
        return np.array(above), np.array(below)

    def __call__(self, p, enei):
        """exc(p, enei): phi' = -nvec . e at the faces (potential.m:13-14)."""
        _check_above(p, self.layer)
        e = np.ascontiguousarray(self.fields(enei)[0], dtype=np.complex128)
        phip = dev.colmajor((p.n, len(self.pol)))
        _lib.call("mnp_planewavestatlayer_potential", p.green.handle, len(self.pol),
                  e.ctypes.data_as(C.c_void_p), dev.ptr(dev.base(phip)))
        return compstruct(p, enei, phip=phip)

    potential = __call__

    def dipole(self, sig):
        """Induced dipole moments, npol x 3 (FP64 device reduction)."""
        npol = sig.sig.shape[-1]
        dip = np.zeros((npol, 3), dtype=np.complex128)
        _lib.call("mnp_stat_dipole_moment", sig.p.green.handle, npol, dev.ptr(dev.base(sig.sig)),
                  dip.ctypes.data_as(C.c_void_p))
        return dip

    def cross_sections(self, sig):
        """(ext, sca, abs) per polarisation (extinction.m 'full', scattering.m,
        absorption.m = ext - sca)."""
        enei = sig.enei
        k0 = 2 * np.pi / enei
        eu, ed = self.layer.epsilon(enei)
        dip = self.dipole(sig)
        te, tm = _decompose(self.pol, self.dir)
        ext = np.zeros(len(self.pol))
        for i, (coef, ki, kr, kt) in enumerate(_fresnel(self.dir, k0, eu, ed)):
            re, tte, rm, ttm = coef
            m = np.cross(ki, tm[i]) / k0
            er = re * te[i] + rm * np.cross(m, kr) * k0 / np.vdot(kr, kr).real
            et = tte * te[i] + ttm * np.cross(m, kt) * k0 / np.vdot(kt, kt).real
            nr, nt = np.linalg.norm(kr), np.linalg.norm(kt)
            far, _ = _dipole_far(dip[i:i + 1], np.vstack([kr / nr, kt / nt]), k0, eu, ed)
            esr, est = far[0, :, 0], far[1, :, 0]
            if abs(kt[2].imag) > 1e-10:
                est = 0 * est
            ext[i] = 4 * np.pi / nr * (np.vdot(er, esr).imag + np.vdot(et, est).imag)
        far, k = _dipole_far(dip, self.pinf_nvec, k0, eu, ed)
        dsca = np.sum(np.abs(far) ** 2, axis=1) * (0.5 * k / k0)[:, None]
        dsca = np.where(np.imag(dsca) != 0, 0, dsca).real
        sca = self.pinf_area @ dsca
        nb = np.sqrt(np.array([eu, ed]))
        sca = np.where(self.dir[:, 2] < 0, sca / (0.5 * nb[0]), np.where(self.dir[:, 2] > 0, sca / (0.5 * nb[-1]), sca))
        sca = np.real(sca)
        return ext, sca, ext - sca

    def ext(self, sig):
        return self.cross_sections(sig)[0]

    def sca(self, sig):
        return self.cross_sections(sig)[1]

    def abs(self, sig):
        return self.cross_sections(sig)[2]

    extinction, scattering, absorption = ext, sca, abs


class dipolestatlayer(dipolestat):
    """Quasistatic dipoles above a substrate (Simulation/static/
    @dipolestatlayer); ``pt``, ``dip``, ``full`` as for dipolestat, op['layer']
    a layerstructure, every point above the interface (init.m:28).  The
    potential adds q1 = -(eps_down - eps_up)/(eps_down + eps_up) times the
    field of the image dipoles (point mirrored at the interface, p_z negated;
    init.m:35-39, field.m:23-28).

    image_eps: 'upper' (the default) divides the image dipole's field by the
    upper medium's permittivity, the image of a dipole in the upper medium
    and the convention of the reflected Green function of bemstatlayer.
    'reference' divides it by the lower medium's, as the reference does (its
    image point lies in the lower medium, field.m:28 with init.m:38-39); over
    a lossy metal substrate that makes the absorbed part of the decay rate
    negative.  Decay rates integrate the radiation over ``pinfty`` = (nvec,
    area) of the unit sphere (default: trisphere(256, 2)).
    """

    def __init__(self, pt, dip=None, op=None, medium=None, full=False, pinfty=None, image_eps="upper"):
        super().__init__(pt, dip, op, medium, full)
        self.layer = op["layer"]
        if np.any(self.pt[:, 2] <= self.layer.z[0]):
            raise ValueError("dipolestatlayer: every dipole must lie above the interface")
        if image_eps not in ("reference", "upper"):
            raise ValueError("image_eps must be 'reference' or 'upper'")
        self.image_eps = image_eps
        if pinfty is None:
            from . import shapes
            pinfty = _flat_normals(*shapes.trisphere256(2.0))
        self.pinf_nvec, self.pinf_area = (np.asarray(a, dtype=np.float64) for a in pinfty)
        self._gimg = None

    def _eps(self, enei):
        eu, ed = self.layer.epsilon(enei)
        return eu, ed, -(ed - eu) / (ed + eu)

    def __call__(self, p, enei):
        """exc(p, enei): phi' = -nvec . (E_dip + q1 E_image) (potential.m, field.m:23-28)."""
        _check_above(p, self.layer)
        eu, ed, q1 = self._eps(enei)
        npt = len(self.pt)
        w = np.full(npt, 1.0 / eu)
        wimg = np.full(npt, q1 / (ed if self.image_eps == "reference" else eu))
        return compstruct(p, enei, phip=self._potential(p, w, self.layer.z[0], wimg))

    potential = __call__

    def _image_green(self, p):
        """Green function of [particle; image] (compgreenstatlayer's p2r), kept per particle."""
        if self._gimg is None or self._gimg[0] is not p:
            V, F = _image_mesh(p.verts, p.faces, self.layer.z[0])
            op = p.op
            g = CompGreenRet(V, F, order=op.get("order", 5), rel_cutoff=op.get("RelCutoff", 0.0),
                             abs_cutoff=op.get("AbsCutoff", 0.0),
                             npol=tuple(np.broadcast_to(op.get("npol", (7, 5)), (2,))), closed=False)
            self._gimg = (p, g)
        return self._gimg[1]

    def _rad(self, dips, enei):
        """(sca / (2 pi k0)) / (0.5 nb gamma) of the dipoles dips (m x 3) above
        the interface (spectrumstatlayer/scattering.m, decayrate.m:54-57)."""
        k0 = 2 * np.pi / enei
        eu, ed, _ = self._eps(enei)
        nb = np.sqrt(eu)
        far, k = _dipole_far(dips, self.pinf_nvec, k0, eu, ed)
        dsca = np.sum(np.abs(far) ** 2, axis=1) * (0.5 * k / k0)[:, None]
        dsca = np.where(np.imag(dsca) != 0, 0, dsca).real
        return np.real((self.pinf_area @ dsca) / (2 * np.pi * k0) / (0.5 * nb * 4 / 3 * k0 ** 3))

    def decayrate(self, sig):
        """[tot, rad, rad0], each npt x ndip (decayrate.m:25-61): the charges'
        field at each dipole -(Gp + f2 Gp_image) sig, rad of the induced
        dipole plus dip / nb^2, tot = 1 + rad + Im(e . dip) / (0.5 nb gamma)."""
        import torch
        p, enei = sig.p, sig.enei
        npt, ndip = self.dip.shape[0], self.dip.shape[2]
        eu, ed, f2 = self._eps(enei)
        nb = np.sqrt(eu)
        gamma = 4 / 3 * (2 * np.pi / enei) ** 3
        ncol = sig.sig.shape[1]
        e = self.field_own(p.green, sig.sig)
        padded = dev.colmajor((2 * p.n, ncol))
        padded[:p.n].zero_()
        padded[p.n:].copy_(sig.sig)
        e = e + f2 * self.field_own(self._image_green(p), padded)
        torch.cuda.synchronize()
        indip = self.induced_dipole(sig)
        d = np.stack([self.dip[ip, :, i] for i in range(ndip) for ip in range(npt)])  # column order
        rad = self._rad(indip + d / nb ** 2, enei)
        nonrad = np.imag(np.sum(e * d, axis=1)) / (0.5 * nb * gamma)
        tot = np.real(1 + rad + nonrad)
        shape = lambda v: np.asarray(v).reshape(ndip, npt).T
        return shape(tot), shape(rad), np.full((npt, ndip), np.real(nb * gamma))

    def decayrate0(self, enei):
        """[tot, rad, rad0] without a particle (decayrate0.m:44-55): tot = rad
        for real media, else rad + Im(q1 E_image . dip) / (0.5 nb gamma)."""
        npt, ndip = self.dip.shape[0], self.dip.shape[2]
        eu, ed, q1 = self._eps(enei)
        nb = np.sqrt(eu)
        gamma = 4 / 3 * (2 * np.pi / enei) ** 3
        z0 = self.layer.z[0]
        d = np.stack([self.dip[ip, :, i] for i in range(ndip) for ip in range(npt)])
        rad = self._rad(d / nb ** 2, enei)
        pt = np.tile(self.pt, (ndip, 1))
        r = np.zeros_like(pt)
        r[:, 2] = 2 * (pt[:, 2] - z0)  # from the image point to the dipole
        rn = np.linalg.norm(r, axis=1)
        u = r / rn[:, None]
        di = d * np.array([1.0, 1.0, -1.0])
        e = q1 * (3 * u * np.sum(u * di, axis=1)[:, None] - di) / (rn ** 3 * (ed if self.image_eps == "reference"
                                                                                 else eu))[:, None]
        real = np.imag(eu) == 0 and np.imag(ed) == 0
        tot = rad if real else np.real(rad + np.imag(np.sum(e * d, axis=1)) / (0.5 * nb * gamma))
        shape = lambda v: np.asarray(v).reshape(ndip, npt).T
        return shape(tot), shape(rad), np.full((npt, ndip), np.real(nb * gamma))


# ------------------------------------------------------------ retarded
_COMPONENTS = ("p", "ss", "hs", "sh", "hh")  # order of reflectionsubs.m's fields


class compgreentablayer:
    """Tabulated reflected Green function of the substrate (Greenfun/
    @compgreentablayer, @greentablayer) on the grid tab = layer.tabspace(p);
    set(enei) precomputes the tables at those wavelengths (set.m), which
    later wavelengths interpolate linearly (eval.m:24-50); without set the
    table is computed at each new wavelength."""

    def __init__(self, layer, tab):
        self.layer = layer
        self.r = np.ascontiguousarray(tab["r"], dtype=np.float64)
        self.z1 = np.ascontiguousarray(tab["z1"], dtype=np.float64)
        self.z2 = float(tab["z2"])
        h = C.c_void_p()
        op = layer.c_options()
        _lib.call("mnp_layertab_create", C.byref(op), len(self.r), self.r.ctypes.data_as(C.c_void_p), len(self.z1),
                  self.z1.ctypes.data_as(C.c_void_p), self.z2, C.byref(h))
        self.handle = h
        self.enei = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib.can_release():
            _lib.lib().mnp_layertab_destroy(h)
            self.handle = None

    def set(self, enei, op=None):
        enei = np.ascontiguousarray(np.atleast_1d(enei), dtype=np.float64)
        eu = (_lib.MnpComplex * len(enei))(*[_lib.MnpComplex.of(self.layer.epsilon(e)[0]) for e in enei])
        ed = (_lib.MnpComplex * len(enei))(*[_lib.MnpComplex.of(self.layer.epsilon(e)[1]) for e in enei])
        _lib.call("mnp_layertab_set", self.handle, len(enei), enei.ctypes.data_as(C.c_void_p), eu, ed)
        self.enei = enei
        return self

    def ismember(self, layer, enei=None, p=None):
        """greentablayer/ismember.m: same layer, wavelengths inside the set range."""
        if self.enei is None or layer is not self.layer:
            return False
        if enei is not None and (np.min(enei) < self.enei.min() or np.max(enei) > self.enei.max()):
            return False
        return True

    def values(self, enei):
        """The normalised table at enei as host (nr, nz, 15) complex: d G,
        d^3 Fr / r, d^3 Fz / zmin of p, ss, hs, sh, hh (norm.m)."""
        eu, ed = self.layer.epsilon(enei)
        out = dev.colmajor((15 * len(self.r) * len(self.z1),))
        _lib.call("mnp_layertab_eval", self.handle, float(enei), _lib.MnpComplex.of(eu), _lib.MnpComplex.of(ed),
                  dev.ptr(dev.base(out)))
        return dev.to_host(out).reshape(len(self.r), len(self.z1), 15)


class greenretlayer:
    """Reflected Green function of the particle with itself on a table
    (Greenfun/@greenretlayer); eval(enei) returns the device G and F = nvec .
    Gp of the components p, ss, hs, sh, hh, each 5 x n x n."""

    def __init__(self, p, tab):
        h = C.c_void_p()
        _lib.call("mnp_greenretlayer_create", p.green.handle, tab.handle, C.byref(h))
        self.handle, self.p, self.tab = h, p, tab

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib.can_release():
            _lib.lib().mnp_greenretlayer_destroy(h)
            self.handle = None

    def refined(self):
        npairs, ndiag = C.c_int64(), C.c_int64()
        _lib.call("mnp_greenretlayer_refined", self.handle, C.byref(npairs), C.byref(ndiag))
        return npairs.value, ndiag.value

    def eval(self, enei):
        n = self.p.n
        eu, ed = self.tab.layer.epsilon(enei)
        G = dev.colmajor((n, n, 5))
        F = dev.colmajor((n, n, 5))
        _lib.call("mnp_greenretlayer_eval", self.handle, float(enei), _lib.MnpComplex.of(eu), _lib.MnpComplex.of(ed),
                  dev.ptr(dev.base(G)), dev.ptr(dev.base(F)))
        return G, F


class compgreenretlayer_points:
    """compgreenretlayer(pt, p, op) for a compoint pt above the interface and
    a particle p (Greenfun/@compgreenretlayer/potential.m, field.m): the
    direct part of compgreenret(pt, p) plus, for the points in the upper
    medium, the reflected part from the charges and currents facing it
    (private/matmul2.m, matmul3.m), on the GPU (mnp_pointlayer_*).
    op['greentab'] must cover the point-face pairs (layer.tabspace(p, pt));
    without it that table is made here."""

    def __init__(self, pt, p, op):
        from .points import compgreenret_points
        self.layer = op["layer"]
        _check_above(p, self.layer)
        up = self.layer.ind[0]
        self.pt, self.p = pt, p
        self.direct = compgreenret_points(pt, p)
        self.sel = np.nonzero(np.asarray(pt.inout) == up)[0]
        pos = np.asarray(pt.pos, dtype=np.float64)[self.sel]
        if np.any(pos[:, 2] <= self.layer.z[0]):
            raise NotImplementedError("compgreenretlayer: points below the interface (transmitted fields) "
                                      "are not supported")
        tab = op.get("greentab")
        if tab is None:
            tab = compgreentablayer(self.layer, self.layer.tabspace(p, pos))
        elif tab.layer is not self.layer:
            raise ValueError("greentab: tabulated for another layer structure")
        self.tab, self.handle = tab, None
        if len(self.sel):
            h = C.c_void_p()
            pts = np.asfortranarray(pos)
            _lib.call("mnp_pointlayer_create", p.green.handle, tab.handle, len(pos), pts.ctypes.data_as(C.c_void_p),
                      C.byref(h))
            self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib.can_release():
            _lib.lib().mnp_pointlayer_destroy(h)
            self.handle = None

    def reflected(self, sig):
        """The reflected phi (nsel x ncol), a, e, h (nsel x 3 x ncol) at the
        points of the upper medium (host arrays)."""
        enei, ncol, m = sig.enei, sig.sig2.shape[-1], len(self.sel)
        s, h = self.direct._charges(sig, self.layer.ind[0])
        eu, ed = self.layer.epsilon(enei)
        out = [dev.colmajor((m, ncol))] + [dev.colmajor((m, 3, ncol)) for _ in range(3)]
        _lib.call("mnp_pointlayer_fields", self.handle, float(enei), _lib.MnpComplex.of(eu), _lib.MnpComplex.of(ed),
                  ncol, dev.ptr(dev.base(s)), dev.ptr(dev.base(h)), *[dev.ptr(dev.base(t)) for t in out])
        return [dev.to_host(t) for t in out]

    def potential(self, sig, inout=1):
        """phi1 (npt x ncol), a1 (npt x 3 x ncol) at the points, direct plus
        reflected (potential.m; points carry no normals, so phi1p, a1p are
        not formed)."""
        pot = self.direct.potential(sig, inout)
        if len(self.sel):
            phi, a, _, _ = self.reflected(sig)
            pot["phi1"][self.sel] += phi
            pot["a1"][self.sel] += a
        return pot

    def field(self, sig, inout=1):
        """e, h (npt x 3 x ncol) at the points, direct plus reflected
        (field.m)."""
        f = self.direct.field(sig, inout)
        if len(self.sel):
            _, _, e, h = self.reflected(sig)
            f["e"][self.sel] += e
            f["h"][self.sel] += h
        return f


class dipoleretlayer(dipoleret):
    """Dipoles above the substrate (Simulation/retarded/@dipoleretlayer):
    pt, dip, full and pinfty as dipoleret, op['layer'] the layerstructure,
    op['greentab'] a compgreentablayer whose grid has the dipoles as sources
    and observers (layer.tabspace(p, dip=pt)); without it that table is made
    at the first call.  The dipoles lie in the upper medium, outside the
    particle; potentials, fields and decay rates on the GPU
    (mnp_dipolelayer_*)."""

    def __init__(self, pt, dip=None, op=None, pinfty=None, full=False):
        super().__init__(pt, dip, op, None, pinfty, full)
        self.layer = op["layer"]
        if np.any(self.pt[:, 2] <= self.layer.z[0]):
            raise NotImplementedError("dipoleretlayer: dipoles at or below the interface are not supported")
        self.tab = op.get("greentab")
        if self.tab is not None and self.tab.layer is not self.layer:
            raise ValueError("greentab: tabulated for another layer structure")
        self._points = None

    def _table(self, p):
        if self.tab is None:
            self.tab = compgreentablayer(self.layer, self.layer.tabspace(p, dip=self.pt))
        return self.tab

    def _eps(self, enei):
        return (_lib.MnpComplex.of(e) for e in self.layer.epsilon(enei))

    def _check_outside(self, p):
        from .points import compoint
        cp = compoint(p, self.pt)
        if len(cp.ind) != len(self.pt) or np.any(np.asarray(cp.inout) != self.layer.ind[0]):
            raise NotImplementedError("dipoleretlayer: the dipoles must lie outside the particle, in the upper medium")

    def __call__(self, p, enei):
        """exc(p, enei): phi2, phi2p, a2, a2p on the faces, direct (dipoleret)
        plus reflected (potential.m:15-96)."""
        _check_above(p, self.layer)
        if getattr(self, "_checked", None) is not p:
            self._check_outside(p)
            self._checked = p
        exc = super().__call__(p, enei)
        pts, dip = self._host()
        _lib.call("mnp_dipolelayer_potential", p.green.handle, self._table(p).handle, float(enei), *self._eps(enei),
                  len(pts), pts.ctypes.data_as(C.c_void_p), dip.shape[2], dip.ctypes.data_as(C.c_void_p),
                  *[dev.ptr(dev.base(getattr(exc, k))) for k in ("phi2", "phi2p", "a2", "a2p")])
        return exc

    potential = __call__

    def field_refl(self, obs, enei):
        """Reflected e, h (host nobs x 3 x npol) of the dipoles at the points
        obs (field.m with 'refl'; obs = the dipole positions gives their
        self-interaction)."""
        obs = np.asfortranarray(np.atleast_2d(obs), dtype=np.float64)
        pts, dip = self._host()
        npol = self.npol
        e = dev.colmajor((len(obs), 3, npol))
        h = dev.colmajor((len(obs), 3, npol))
        if self.tab is None:
            raise ValueError("dipoleretlayer: no table yet (pass op['greentab'] or call the excitation first)")
        _lib.call("mnp_dipolelayer_field", self.tab.handle, float(enei), *self._eps(enei), len(obs),
                  obs.ctypes.data_as(C.c_void_p), len(pts), pts.ctypes.data_as(C.c_void_p), dip.shape[2],
                  dip.ctypes.data_as(C.c_void_p), dev.ptr(dev.base(e)), dev.ptr(dev.base(h)))
        return dev.to_host(e), dev.to_host(h)

    def farfield(self, dirs, enei):
        """e, h (host ndir x 3 x npol) of the dipoles in the directions dirs
        (farfield.m:21-177)."""
        ddir = dev.to_device_colmajor(np.atleast_2d(dirs), dev.torch.float64)
        pts, dip = self._host()
        nd = ddir.shape[0]
        e = dev.colmajor((nd, 3, self.npol))
        h = dev.colmajor((nd, 3, self.npol))
        op = self.layer.c_options()
        _lib.call("mnp_dipolelayer_farfield", C.byref(op), float(enei), *self._eps(enei), nd, dev.ptr(dev.base(ddir)),
                  len(pts), pts.ctypes.data_as(C.c_void_p), dip.shape[2], dip.ctypes.data_as(C.c_void_p),
                  dev.ptr(dev.base(e)), dev.ptr(dev.base(h)))
        return dev.to_host(e), dev.to_host(h)

    def field(self, sig):
        """Induced e (host npt x 3 x npol) at the dipoles: compgreenretlayer's
        field of the surface charges, direct plus reflected (decayrate.m:22-23,
        first term)."""
        from .points import compoint
        p = sig.p
        e = dev.to_host(super().field(sig).e)
        if self._points is None or self._points.p is not p:
            self._points = compgreenretlayer_points(compoint(p, self.pt), p,
                                                    {"layer": self.layer, "greentab": self._table(p)})
        pl = self._points
        if len(pl.pt.ind) != len(self.pt):
            raise NotImplementedError("dipoleretlayer: a dipole lies on the particle surface")
        if len(pl.sel):
            e[pl.pt.ind[pl.sel]] += pl.reflected(sig)[2]
        return e

    def decayrate(self, sig):
        """[tot, rad, rad0] (npt x ndip) in units of the free-space rate
        (decayrate.m:14-60): tot from the induced field at the dipoles plus
        their reflected self-field, rad from the power of the far fields of
        the particle (spectrumretlayer) and of the dipoles over pinfty."""
        enei = sig.enei
        npt, ndip = self.dip.shape[0], self.dip.shape[2]
        e = self.field(sig) + self.field_refl(self.pt, enei)[0]
        nvec = dev.to_host(self.pinf_nvec)
        area = dev.to_host(self.pinf_area)
        ed, hd = self.farfield(nvec, enei)
        es, hs = planewaveretlayer.farfield(self, sig, nvec)
        s = np.cross(es + ed, np.conj(hs + hd), axisa=1, axisb=1, axisc=1)
        sca = area @ (0.5 * np.real(np.einsum("dc,dcq->dq", nvec, s)))
        k0 = 2 * np.pi / enei
        eu = complex(self.layer.epsilon(enei)[0])
        nb = np.sqrt(eu)
        gamma = 4 / 3 * k0 ** 3
        tot = np.zeros((npt, ndip))
        rad = np.zeros((npt, ndip))
        for ip in range(npt):
            for i in range(ndip):
                q = ip + npt * i
                tot[ip, i] = np.real(1 + np.imag(e[ip, :, q] @ self.dip[ip, :, i]) / (0.5 * nb * gamma))
                rad[ip, i] = np.real(sca[q] / (2 * np.pi * k0) / (0.5 * nb * gamma))
        return tot, rad, np.real(nb * gamma) * np.ones((npt, ndip))


def compgreenretlayer(p1, p2, op):
    """compgreenretlayer(pt, p, op) for a compoint pt (the particle's own
    Green function with the interface is bemretlayer's greenretlayer)."""
    from .points import compoint
    if isinstance(p1, compoint):
        return compgreenretlayer_points(p1, p2, op)
    raise NotImplementedError("compgreenretlayer: points (compoint) and a particle")


def _greentab(p, op):
    tab = op.get("greentab")
    layer = op["layer"]
    if tab is None:
        tab = compgreentablayer(layer, layer.tabspace(p))
    elif tab.layer is not layer:
        raise ValueError("greentab: tabulated for another layer structure")
    return tab


class bemretlayer:
    """Dense retarded BEM solver above a substrate (BEM/@bemretlayer):
    op['layer'] a layerstructure whose upper medium is the particle's
    outside; op['greentab'] a compgreentablayer (default: the tabspace of p,
    tabulated at each wavelength).  precision 'fp64' (the complex FP64 chain,
    the drop-in) or 'mixed' (the chain in complex FP32, refined with the FP64
    operator)."""

    _SOLVERS = {"fp64": _lib.MNP_SOLVER_FP64, "mixed": _lib.MNP_SOLVER_MIXED}

    def __init__(self, p, op=None, precision=None):
        op = p.op if op is None else op
        self.layer = op["layer"]
        _check_above(p, self.layer)
        self.precision = precision or op.get("precision", "fp64")
        if self.precision not in self._SOLVERS:
            raise ValueError(f"bemretlayer: precision must be 'fp64' or 'mixed', not {self.precision!r}")
        self.p, self.tab = p, _greentab(p, op)
        self.green = greenretlayer(p, self.tab)
        h = C.c_void_p()
        _lib.call("mnp_bemretlayer_create", p.green.handle, self.green.handle, self._SOLVERS[self.precision],
                  C.byref(h))
        self.handle = h
        self.enei = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib.can_release():
            _lib.lib().mnp_bemretlayer_destroy(h)
            self.handle = None

    def __call__(self, enei):
        if self.enei != enei:
            self.enei = None
            eu, ed = self.layer.epsilon(enei)
            if abs(complex(self.p.eps2(enei)) - eu) > 1e-12 * max(1.0, abs(eu)):
                raise ValueError("bemretlayer: the particle's outside must be the upper medium of the layer")
            _lib.call("mnp_bemretlayer_initmat", self.handle, float(enei), _lib.MnpComplex.of(self.p.eps1(enei)),
                      _lib.MnpComplex.of(eu), _lib.MnpComplex.of(ed))
            self.enei = enei
        return self

    def solve(self, exc):
        """sig = bem \\ exc (mldivide.m:31-76)."""
        self(exc.enei)
        n = self.p.n
        npol = next(v.shape[-1] for v in exc.values() if v is not None)
        args = []
        for key in ("phi1", "phi1p", "a1", "a1p", "phi2", "phi2p", "a2", "a2p"):
            v = exc.get(key)
            args.append(None if v is None else dev.ptr(dev.base(v)))
        out = [dev.colmajor((n, npol)), dev.colmajor((n, npol)), dev.colmajor((n, 3, npol)),
               dev.colmajor((n, 3, npol))]
        _lib.call("mnp_bemretlayer_solve", self.handle, npol, *args, *[dev.ptr(dev.base(t)) for t in out])
        return compstruct(self.p, exc.enei, sig1=out[0], sig2=out[1], h1=out[2], h2=out[3])

    mldivide = solve

    def stats(self):
        """{'fallbacks': LU factorisations so far redone with partial pivoting,
        'iterations', 'relres': refinement sweeps and final relative residual
        of the last solve (mixed)}."""
        fb, it, rr = C.c_int(), C.c_int(), C.c_double()
        _lib.call("mnp_bemretlayer_stats", self.handle, C.byref(fb), C.byref(it), C.byref(rr))
        return {"fallbacks": fb.value, "iterations": it.value, "relres": rr.value}


class bemretlayeriter(bemretiter):
    """Iterative retarded BEM solver above a substrate (BEM/@bemretlayeriter):
    op['layer'] the layerstructure, op['greentab'] a compgreentablayer
    covering the particle (default: its tabspace, tabulated per wavelength),
    op['iter'] the bemiter options as for bemretiter.  The outside Green
    function is five H-matrices per operator (components p, ss, hs, sh, hh:
    direct + reflected for p, ss, hh, reflected for hs, sh; Greenfun/+aca/
    @compgreenretlayer), built on the GPU by ACA of the direct plus the
    tabulated reflected kernel, with the refined near-field values in the full
    blocks; GMRES on the operator of bemretlayeriter/private/afun.m with the
    two-level preconditioner of bemretiter on the p components
    (mnp_bemretiter_create_layer)."""

    def __init__(self, p, op):
        self.layer = op["layer"]
        _check_above(p, self.layer)
        self.tab = _greentab(p, op)
        self.green = greenretlayer(p, self.tab)
        super().__init__(p, op)

    def _create(self, prec, perm, fr, lr, htol, kmax, h):
        _lib.call("mnp_bemretiter_create_layer", self.p.green.handle, self.green.handle, prec,
                  perm.ctypes.data_as(C.c_void_p), len(fr), fr.ctypes.data_as(C.c_void_p), len(lr),
                  lr.ctypes.data_as(C.c_void_p), htol, kmax, C.byref(h))

    def __call__(self, enei):
        if self.enei != enei:
            self.enei = None
            eu, ed = self.layer.epsilon(enei)
            if abs(complex(self.p.eps2(enei)) - eu) > 1e-12 * max(1.0, abs(eu)):
                raise ValueError("bemretlayeriter: the particle's outside must be the upper medium of the layer")
            _lib.call("mnp_bemretiter_initmat_layer", self.handle, float(enei),
                      _lib.MnpComplex.of(self.p.eps1(enei)), _lib.MnpComplex.of(eu), _lib.MnpComplex.of(ed))
            self.enei = enei
        return self


def _kz(k, kpar, offset=False):
    kz = np.sqrt(k * k - kpar * kpar + 0j)
    if offset:
        kz = kz + 1e-10j
        return kz * np.sign(kz.imag)
    return kz * np.sign((kz + 1e-10j).imag)


def _coef(k0, eu, ed, kpar, o, s, offset_ratio=True):
    """p, hh, sh coefficients for the observer in medium o and the source in
    medium s (1 above) at kpar times fresnel.m's k_s,z / k_o,z
    (reflectionsubs.m:27-60, fresnel.m:17-33), and kz of both media."""
    k1, k2 = k0 * np.sqrt(eu + 0j), k0 * np.sqrt(ed + 0j)
    k1z, k2z = _kz(k1, kpar), _kz(k2, kpar)
    rr = (k1z - k2z) / (k2z + k1z)
    D = (k2z + k1z) * (eu * k2z + ed * k1z)
    hh = ((k1z - k2z) * (2 * eu * k1z - ed * k1z + eu * k2z) / D,
          (k2z - k1z) * (2 * ed * k2z - eu * k2z + ed * k1z) / D)
    sh = (-2 * k0 * (ed - eu) * k1z / D, -2 * k0 * (eu - ed) * k2z / D)
    q12, q21 = k1z / k2z, k2z / k1z
    table = {(1, 1): (rr, hh[0], sh[0]), (1, 2): (1 + rr, q12 * (hh[1] + 1), -q12 * sh[1]),
             (2, 1): (1 - rr, q21 * (hh[0] + 1), q21 * sh[0]), (2, 2): (-rr, hh[1], -sh[1])}
    c = table[(o, s)]
    kzo = {1: _kz(k1, kpar, True), 2: _kz(k2, kpar, True)}
    ratio = kzo[s] / kzo[o] if offset_ratio else 1.0
    return tuple(v * ratio for v in c), (k1z, k2z)


class planewaveretlayer:
    """Plane wave above a substrate, retarded (Simulation/retarded/
    @planewaveretlayer): pol and dir npol x 3, op['layer'] a layerstructure;
    scattering over pinfty = (nvec, area) of the unit sphere (default: the
    256-point sphere of trisphere(256, 2))."""

    def __init__(self, pol, dir, op=None, pinfty=None):
        self.pol = np.atleast_2d(np.asarray(pol, dtype=np.float64))
        self.dir = np.atleast_2d(np.asarray(dir, dtype=np.float64))
        if np.any(np.sum(self.pol * self.dir, axis=1) != 0):
            raise ValueError("polarisation must be perpendicular to the propagation direction")
        if len(self.pol) > 8:
            raise ValueError("planewaveretlayer: at most 8 polarisations")
        self.layer = op["layer"]
        if pinfty is None:
            from . import shapes
            pinfty = _flat_normals(*shapes.trisphere256(2.0))
        self.pinf_nvec, self.pinf_area = (np.asarray(a, dtype=np.float64) for a in pinfty)

    def __call__(self, p, enei):
        """exc(p, enei): phi2, phi2p, a2, a2p at the faces (potential.m:20-105)."""
        _check_above(p, self.layer)
        n, npol = p.n, len(self.pol)
        eu, ed = self.layer.epsilon(enei)
        out = [dev.colmajor((n, npol)), dev.colmajor((n, npol)), dev.colmajor((n, 3, npol)),
               dev.colmajor((n, 3, npol))]
        op = self.layer.c_options()
        pol, dr = np.asfortranarray(self.pol), np.asfortranarray(self.dir)
        _lib.call("mnp_planewaveretlayer_potential", p.green.handle, C.byref(op), float(enei),
                  _lib.MnpComplex.of(eu), _lib.MnpComplex.of(ed), npol, pol.ctypes.data_as(C.c_void_p),
                  dr.ctypes.data_as(C.c_void_p), *[dev.ptr(dev.base(t)) for t in out])
        return compstruct(p, enei, phi2=out[0], phi2p=out[1], a2=out[2], a2p=out[3])

    potential = __call__

    def farfield(self, sig, dirs):
        """e, h (host ndir x 3 x npol) in the directions dirs
        (@spectrumretlayer/farfield.m:20-104)."""
        dirs = np.asfortranarray(np.atleast_2d(dirs), dtype=np.float64)
        npol = sig.sig2.shape[-1]
        eu, ed = self.layer.epsilon(sig.enei)
        ddir = dev.to_device_colmajor(dirs, dev.torch.float64)
        e = dev.colmajor((len(dirs), 3, npol))
        h = dev.colmajor((len(dirs), 3, npol))
        op = self.layer.c_options()
        _lib.call("mnp_farfield_layer", sig.p.green.handle, C.byref(op), float(sig.enei), _lib.MnpComplex.of(eu),
                  _lib.MnpComplex.of(ed), len(dirs), dev.ptr(dev.base(ddir)), npol, dev.ptr(dev.base(sig.sig2)),
                  dev.ptr(dev.base(sig.h2)), dev.ptr(dev.base(e)), dev.ptr(dev.base(h)))
        return dev.to_host(e), dev.to_host(h)

    def efresnel(self, enei):
        """Reflected and transmitted fields and wavevectors per polarisation
        (Particles/@layerstructure/efresnel.m:19-66)."""
        k0 = 2 * np.pi / enei
        eu, ed = self.layer.epsilon(enei)
        ks = (k0 * np.sqrt(eu + 0j), k0 * np.sqrt(ed + 0j))
        z0 = self.layer.z[0]
        za, zb = z0 + 1e-10, z0 - 1e-10
        out = []
        for pl, d in zip(self.pol, self.dir):
            if d[2] < 0:
                posr, post = (za, 1, za, 1), (zb, 2, za, 1)
            else:
                posr, post = (zb, 2, zb, 2), (za, 1, zb, 2)
            kpv = ks[post[3] - 1] * d[:2]
            kp2 = np.dot(kpv, kpv)
            kzr = np.sqrt(ks[posr[1] - 1] ** 2 - kp2 + 0j)
            kzr = kzr * np.sign((kzr + 1e-10j).imag)
            kzt = np.sqrt(ks[post[1] - 1] ** 2 - kp2 + 0j)
            kzt = kzt * np.sign((kzt + 1e-10j).imag)
            kr = np.array([kpv[0], kpv[1], -np.sign(d[2]) * kzr])
            kt = np.array([kpv[0], kpv[1], np.sign(d[2]) * kzt])
            kpn = np.sqrt(kp2 + 0j)
            fields = []
            for (z1, i1, z2, i2), kv in ((posr, kr), (post, kt)):
                (cp, chh, csh), kz = _coef(k0, eu, ed, kpn, i1, i2)
                g = np.exp(1j * kz[i1 - 1] * abs(z1 - z0)) * np.exp(1j * kz[i2 - 1] * abs(z2 - z0))
                ev = (np.array([cp * pl[0], cp * pl[1], chh * pl[2]]) - kv / k0 * csh * pl[2]) * g
                fields.append(ev)
            er = np.exp(-1j * kr[2] * posr[2] - 1j * kr[2] * posr[0]) * fields[0]
            et = np.exp(-1j * kr[2] * post[2] - 1j * kt[2] * post[0]) * fields[1]
            out.append((er, et, kr, kt))
        return out

    def cross_sections(self, sig):
        """(ext, sca, abs) per polarisation (extinction.m 'full', scattering.m,
        absorption.m = ext - sca), FP64 sums on the host over the GPU far fields."""
        npol = len(self.pol)
        ext = np.zeros(npol)
        for i, (er, et, kr, kt) in enumerate(self.efresnel(sig.enei)):
            knr = np.linalg.norm(kr)
            dirs = [(kr / knr).real]
            trans = abs(kt[2].imag) <= 1e-10
            if trans:
                dirs.append((kt / np.linalg.norm(kt)).real)
            e, _ = self.farfield(sig, np.array(dirs))
            esr = e[0, :, i]
            est = e[1, :, i] if trans else np.zeros(3, complex)
            ext[i] = 4 * np.pi / knr * (np.vdot(er, esr).imag + np.vdot(et, est).imag)
        e, h = self.farfield(sig, self.pinf_nvec)
        s = np.cross(e, np.conj(h), axisa=1, axisb=1, axisc=1)
        dsca = 0.5 * np.real(np.einsum("dc,dcp->dp", self.pinf_nvec, s))
        sca = self.pinf_area @ dsca
        eu, ed = self.layer.epsilon(sig.enei)
        dz = self.dir[:, 2]
        fac = np.where(dz < 0, 0.5 * np.sqrt(eu + 0j), np.where(dz > 0, 0.5 * np.sqrt(ed + 0j), 1.0))
        sca = np.real(sca / fac)
        return ext, sca, ext - sca

    def ext(self, sig):
        return self.cross_sections(sig)[0]

    def sca(self, sig):
        return self.cross_sections(sig)[1]

    def abs(self, sig):
        return self.cross_sections(sig)[2]

    extinction, scattering, absorption = ext, sca, abs
