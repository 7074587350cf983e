"""Fields on a mesh of points around a particle (MNPBEM Simulation/misc/@meshfield).

The points are split by the side of the surface they lie on and each group
takes the fields of its side: e = i k0 G h - Gp sig and h = Gp x h
(Greenfun/@compgreenret/field.m) with sig1, h1 and the inside wavenumber for
points inside, sig2, h2 and the outside one for points outside
(mnp_field_points, computed on the GPU).

The side of a point follows Misc/distmin3.m's sign: the point's offset from
the centroid of its nearest face, projected on that face's normal (distmin3
takes the face of the nearest quadrature point; the two agree for points
farther from the surface than a face's size).  Points closer to the surface
than ``mindist`` get NaN fields, as compoint leaves them out.  Quasistatic
solutions take e = -Gp sig at k = 0 (the same kernel, no currents).
"""
import ctypes as C

import numpy as np

from . import _lib
from . import device as dev


def _expand3(a, b, c):
    """meshfield/init.m expand: a, b of equal shape; c a scalar (filled to
    that shape) or a vector stacked along a new last dimension."""
    if c.size == 1:
        return a, b, np.full(a.shape, float(c.ravel()[0]))
    n2 = c.size
    a3 = np.repeat(a[..., None], n2, axis=-1)
    b3 = np.repeat(b[..., None], n2, axis=-1)
    c3 = np.broadcast_to(c.reshape((1,) * a.ndim + (n2,)), a3.shape).copy()
    return a3, b3, c3


def _expand(x, y, z):
    """meshfield/init.m:20-28: equal shapes are kept; otherwise the two of
    equal shape span the mesh and the third is expanded over it."""
    x, y, z = (np.asarray(v, dtype=np.float64) for v in (x, y, z))
    if x.shape == y.shape == z.shape:
        return x, y, z
    if x.shape == y.shape:
        return _expand3(x, y, z)
    if x.shape == z.shape:
        x, z, y = _expand3(x, z, y)
        return x, y, z
    if y.shape == z.shape:
        y, z, x = _expand3(y, z, x)
        return x, y, z
    raise ValueError("meshfield: x, y, z do not form a mesh")


class meshfield:
    """meshfield(p, x, y, z, mindist=0): ``field(sig)`` returns e, h with the
    mesh's shape followed by (3, npol)."""

    def __init__(self, p, x, y, z, mindist=0.0):
        self.p = p
        self.x, self.y, self.z = _expand(x, y, z)
        pts = np.column_stack([self.x.ravel(), self.y.ravel(), self.z.ravel()])
        geo = p.green.geometry()
        pos, nvec = geo["pos"], geo["nvec"]
        # nearest face centroid, in chunks to bound memory
        ind = np.empty(len(pts), dtype=np.int64)
        for a in range(0, len(pts), 4096):
            d2 = np.sum((pts[a:a + 4096, None, :] - pos[None, :, :]) ** 2, axis=2)
            ind[a:a + 4096] = np.argmin(d2, axis=1)
        dmin = np.sum((pts - pos[ind]) * nvec[ind], axis=1)
        keep = np.abs(dmin) >= mindist
        self.pts = pts
        self.inside = keep & (dmin < 0)
        self.outside = keep & (dmin >= 0)

    def _side(self, sel, enei, eps, sig, h):
        npt, ncol = int(sel.sum()), sig.shape[-1]
        e = dev.colmajor((npt, 3, ncol))
        hf = dev.colmajor((npt, 3, ncol))
        pts = np.asfortranarray(self.pts[sel])
        _lib.call("mnp_field_points", self.p.green.handle, float(enei), _lib.MnpComplex.of(eps), npt,
                  pts.ctypes.data_as(C.c_void_p), ncol, dev.ptr(dev.base(sig)),
                  None if h is None else dev.ptr(dev.base(h)),
                  dev.ptr(dev.base(e)), dev.ptr(dev.base(hf)))
        return dev.to_host(e), dev.to_host(hf)

    def field(self, sig):
        """(e, h): mesh shape x 3 x npol, complex128 host arrays (field1.m).
        A quasistatic solution (surface charge ``sig`` only) gives e = -Gp sig
        at k = 0 on both sides (@compgreenstat/field.m) and h = None."""
        p, enei = self.p, sig.enei
        static = "sig" in sig and "sig2" not in sig
        ncol = (sig.sig if static else sig.sig2).shape[-1]
        e = np.full((len(self.pts), 3, ncol), np.nan, dtype=np.complex128)
        h = np.full_like(e, np.nan)
        if static:
            sides = ((self.outside | self.inside, 0.0, sig.sig, None),)
        else:
            sides = ((self.outside, p.eps2(enei), sig.sig2, sig.h2), (self.inside, p.eps1(enei), sig.sig1, sig.h1))
        for sel, eps, s, hh in sides:
            if sel.any():
                e[sel], h[sel] = self._side(sel, enei, eps, s, hh)
        shape = self.x.shape + (3, ncol)
        return e.reshape(shape), (None if static else h.reshape(shape))
