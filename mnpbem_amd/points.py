"""Points in the dielectric environment of a particle and the Green function
between them and the particle (MNPBEM Particles/@compoint,
Greenfun/@compgreenret with a compoint as first argument).

compoint(p, pos, mindist=0, medium=None) keeps the positions farther than
``mindist`` from the surface and gives each the medium of the side of its
nearest face (Particles/@compoint/init.m:13-60): the inside medium of that
face's particle when the point lies behind the face (negative offset along
the normal), else the outside one.  The nearest face is the one of the
nearest centroid (distmin3 takes the face of the nearest quadrature point;
the two agree for points farther from the surface than a face's size).

compgreenret(pt, p).potential(sig) and .field(sig) evaluate, for the points
of each medium m, the potentials phi = G sig, a = G h
(Greenfun/@compgreenret/potential.m:30-40) and the fields
e = i k0 G h - Gp sig, h = Gp x h (field.m:19-35), with the wavenumber of m
and the charges of the faces that border m: sig1, h1 of faces whose inside
is m and sig2, h2 of faces whose outside is m (the connectivity of
Particles/@compound/connect.m).  Both are computed on the GPU
(mnp_potential_points, mnp_field_points); the side selection is a masked sum
of the two charge sets, exact because G depends only on the point, the face
and the point's wavenumber.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from . import device as dev


def face_inout(p):
    """n x 2 medium indices (1-based, inside | outside) of every face of p."""
    fio = getattr(p, "face_inout", None)
    if fio is not None:
        return np.asarray(fio)
    return np.broadcast_to(np.asarray(p.inout)[0], (p.n, 2))


class compoint:
    """Points grouped by the medium they lie in (Particles/@compoint)."""

    def __init__(self, p, pos, mindist=0.0, medium=None):
        pos = np.atleast_2d(np.asarray(pos, dtype=np.float64))
        self.pin, self.npos = p, len(pos)
        geo = p.green.geometry()
        fpos, nvec = geo["pos"], geo["nvec"]
        near = np.empty(len(pos), dtype=np.int64)
        for a in range(0, len(pos), 4096):
            d2 = np.sum((pos[a:a + 4096, None, :] - fpos[None, :, :]) ** 2, axis=2)
            near[a:a + 4096] = np.argmin(d2, axis=1)
        r = np.sum((pos - fpos[near]) * nvec[near], axis=1)
        keep = np.nonzero(np.abs(r) >= mindist)[0]
        fio = face_inout(p)
        io = np.where(r[keep] < 0, fio[near[keep], 0], fio[near[keep], 1])
        if medium is not None:
            sel = np.isin(io, np.atleast_1d(medium))
            keep, io = keep[sel], io[sel]
        self.ind = keep                 # indices into pos of the kept points
        self.pos = pos[keep]
        self.inout = io                 # medium of each kept point
        self.eps = p.eps

    @property
    def n(self):
        return len(self.pos)


class compgreenret_points:
    """compgreenret(pt, p) for a compoint pt and a particle p."""

    def __init__(self, pt, p):
        self.pt, self.p = pt, p

    def _charges(self, sig, m):
        """Charges and currents seen from medium m: sig1, h1 on faces whose
        inside is m plus sig2, h2 on faces whose outside is m."""
        fio = face_inout(self.p)
        m1 = torch.as_tensor((fio[:, 0] == m).astype(np.float64), device="cuda")
        m2 = torch.as_tensor((fio[:, 1] == m).astype(np.float64), device="cuda")
        s = dev.colmajor(tuple(sig.sig2.shape))
        h = dev.colmajor(tuple(sig.h2.shape))
        s.copy_(sig.sig1 * m1[:, None] + sig.sig2 * m2[:, None])
        h.copy_(sig.h1 * m1[:, None, None] + sig.h2 * m2[:, None, None])
        return s, h

    def _eval(self, name, sig):
        pt, p, enei = self.pt, self.p, sig.enei
        ncol = sig.sig2.shape[-1]
        shape1 = (3, ncol) if name == "mnp_field_points" else (ncol,)
        out1 = np.zeros((pt.n,) + shape1, dtype=np.complex128)
        out2 = np.zeros((pt.n, 3, ncol), dtype=np.complex128)
        for m in np.unique(pt.inout):
            sel = np.nonzero(pt.inout == m)[0]
            s, h = self._charges(sig, m)
            eps = p.eps[m - 1](enei)[0]
            o1 = dev.colmajor((len(sel),) + shape1)
            o2 = dev.colmajor((len(sel), 3, ncol))
            pts = np.asfortranarray(pt.pos[sel])
            _lib.call(name, p.green.handle, float(enei), _lib.MnpComplex.of(eps), len(sel),
                      pts.ctypes.data_as(C.c_void_p), ncol, dev.ptr(dev.base(s)), dev.ptr(dev.base(h)),
                      dev.ptr(dev.base(o1)), dev.ptr(dev.base(o2)))
            out1[sel], out2[sel] = dev.to_host(o1), dev.to_host(o2)
        return out1, out2

    def potential(self, sig, inout=1):
        """phi1 (npt x npol) and a1 (npt x 3 x npol) at the points
        (compgreenret/potential.m; points take inout = 1).  Points carry no
        normals, so the surface derivatives phi1p, a1p are not formed."""
        if inout != 1:
            raise ValueError("compgreenret(pt, p).potential: points take inout = 1")
        phi, a = self._eval("mnp_potential_points", sig)
        return {"enei": sig.enei, "phi1": phi, "a1": a}

    def field(self, sig, inout=1):
        """e, h (npt x 3 x npol) at the points (compgreenret/field.m)."""
        if inout != 1:
            raise ValueError("compgreenret(pt, p).field: points take inout = 1")
        e, h = self._eval("mnp_field_points", sig)
        return {"enei": sig.enei, "e": e, "h": h}


def compgreenret(p1, p2, op=None):
    """compgreenret(pt, p): the Green function between a compoint and a
    particle (the particle's own compgreenret(p, p) is p.green)."""
    if isinstance(p1, compoint):
        return compgreenret_points(p1, p2)
    if p1 is p2:
        return p2.green
    raise NotImplementedError("compgreenret: points (compoint) or the particle itself")
