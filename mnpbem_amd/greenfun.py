"""Retarded Green function of a particle (compgreenret / greenret).

Mirrors ``compgreenret(p, p, op)`` of MNPBEM (Greenfun/@compgreenret) for a
closed particle of flat triangles, or a compound of several parts with their
own media.  The object owns a device-resident
``mnp_greenret`` handle (geometry, refinement tables, sum rule) and evaluates
``G``, ``F``, ``H1``, ``H2`` on the GPU.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from . import device as dev


class CompGreenRet:
    """compgreenret(p, p, op) for a closed particle (verts in nm, 0-based faces).

    Options follow bemoptions('sim', 'ret'): order 5, RelCutoff 3, AbsCutoff 0,
    polar rule npol [7, 5] (Greenfun/@greenret/private/init.m:9-21).
    """

    def __init__(self, verts, faces, order=5, rel_cutoff=3.0, abs_cutoff=0.0, npol=(7, 5),
                 closed=True, like=None, compound=None):
        """``like``: a CompGreenRet of the same faces whose refined pairs are
        taken instead of selecting them for this mesh (mnp_greenret_create_like).
        ``compound``: (part offsets (nparts + 1), inout (nparts x 2), closed
        masks (nparts)) of several parts (mnp_greenret_create_compound)."""
        dev.ensure()
        verts = np.asfortranarray(np.asarray(verts, dtype=np.float64))
        faces = np.asfortranarray(np.asarray(faces, dtype=np.int64))
        if verts.ndim != 2 or verts.shape[1] != 3 or faces.ndim != 2 or faces.shape[1] != 3:
            raise ValueError("verts must be (m, 3) and faces (n, 3)")
        op = _lib.GreenOptions(int(order), float(rel_cutoff), float(abs_cutoff), int(npol[0]),
                               int(npol[1]), int(bool(closed)))
        h = C.c_void_p()
        if compound is not None:
            off, inout, cmask = compound
            off = np.ascontiguousarray(off, dtype=np.int64)
            io = np.asfortranarray(np.asarray(inout, dtype=np.int32))
            cm = np.ascontiguousarray(cmask, dtype=np.uint32)
            _lib.call("mnp_greenret_create_compound", verts.ctypes.data_as(C.c_void_p), verts.shape[0],
                      faces.ctypes.data_as(C.c_void_p), faces.shape[0], len(cm), off.ctypes.data_as(C.c_void_p),
                      io.ctypes.data_as(C.c_void_p), cm.ctypes.data_as(C.c_void_p), C.byref(op), C.byref(h))
        elif like is None:
            _lib.call("mnp_greenret_create", verts.ctypes.data_as(C.c_void_p), verts.shape[0],
                      faces.ctypes.data_as(C.c_void_p), faces.shape[0], C.byref(op), C.byref(h))
        else:
            _lib.call("mnp_greenret_create_like", verts.ctypes.data_as(C.c_void_p), verts.shape[0],
                      faces.ctypes.data_as(C.c_void_p), faces.shape[0], C.byref(op), like.handle, C.byref(h))
        self.handle = h
        self.order = int(order)
        n, nnz = C.c_int64(), C.c_int64()
        _lib.call("mnp_greenret_size", h, C.byref(n), C.byref(nnz))
        self.n, self.nnz = n.value, nnz.value
        self._geo = None

    def __del__(self):
        h = getattr(self, "handle", None)
        try:
            if h is not None and h.value and _lib.can_release():
                _lib.lib().mnp_greenret_destroy(h)
                self.handle = None
        except Exception:  # interpreter shutdown: module globals may be gone
            pass

    # ------------------------------------------------------------ geometry
    def geometry(self):
        """pos, nvec, tvec1, tvec2 (n, 3) and area (n,) as host arrays."""
        if self._geo is None:
            n = self.n
            arrs = [np.empty((n, 3), order="F") for _ in range(4)] + [np.empty(n)]
            _lib.call("mnp_greenret_geometry", self.handle,
                      *[a.ctypes.data_as(C.c_void_p) for a in arrs])
            self._geo = dict(zip(("pos", "nvec", "tvec1", "tvec2", "area"), arrs))
        return self._geo

    def refined(self):
        """(rows, cols, g, f): refined pairs in column-major order and their
        Taylor tables g (nnz, order+1), f (nnz, 3, order+1) (greenret.g/.f)."""
        nnz, o1 = self.nnz, self.order + 1
        rows = np.empty(nnz, np.int64)
        cols = np.empty(nnz, np.int64)
        g = np.empty((nnz, o1), order="F")
        f = np.empty((nnz, 3, o1), order="F")
        _lib.call("mnp_greenret_refined", self.handle, *[a.ctypes.data_as(C.c_void_p)
                                                         for a in (rows, cols, g, f)])
        return rows, cols, g, f

    # ---------------------------------------------------------- evaluation
    def eval(self, k, key="G", precision="fp64", out=None):
        """Dense G, F, H1 or H2 at wavenumber k as a device (n, n) tensor
        (Greenfun/@greenret/private/eval1.m)."""
        n = self.n
        dt = torch.complex128 if precision == "fp64" else torch.complex64
        if out is None:
            out = dev.colmajor((n, n), dt)
        b = dev.base(out)
        prec = _lib.MNP_FP64 if precision == "fp64" else _lib.MNP_FP32
        _lib.call("mnp_greenret_eval", self.handle, _lib.MnpComplex.of(k), _lib.MNP_KEY[key],
                  prec, dev.ptr(b), n)
        return out
