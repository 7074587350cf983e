"""Derivatives of plane-wave extinction with respect to the particle's
dielectric function and its Drude parameters.

A solution x of A(eps1) x = b (dense retarded solver, excitation from
outside) is differentiated by one extra solve with the forward factors:
A dx = -(dA/deps1) x (mnp_bemret_tangent_eps1).  The extinction is
sigma_ext = 4 pi / k Im(pol . e(dir)) with the far field e linear in x, so
f = pol . e is holomorphic in eps1 and

    d sigma_ext / d Re eps1 = 4 pi / k Im(f'),
    d sigma_ext / d Im eps1 = 4 pi / k Re(f'),

both read off as the extinction of dx and of i dx.  For a parameter p of the
material, d sigma_ext / dp = 4 pi / k Im(f' d eps1 / dp).

The derivatives with respect to the vertex positions come from one
transposed solve with the forward factors, lam = A^{-T} c for J = c^T x
(mnp_bemret_extinction_gradient): dJ = dc^T x + lam^T (db - dA x), with dA
taken element by element where the Green function's elements are evaluated
(closed form, refined near-field and self integrals, closed-surface sum
rule).  The same lam gives d sigma_ext / d eps1 as lam^T (-dA/deps1 x).
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from . import device as dev
from .bem import EV2NM, bemret, compstruct, epsdrude


def eps1_tangent(bem, sig):
    """d sig / d eps1 (complex derivative) of the solution ``sig`` of ``bem`` at
    its current wavelength, as a compstruct with sig1, sig2, h1, h2."""
    if not isinstance(bem, bemret):
        raise TypeError("eps1_tangent needs the dense retarded solver (bemret)")
    bem(sig.enei)
    n, npol = bem.p.n, sig.sig2.shape[-1]
    out = dict(sig1=dev.colmajor((n, npol)), sig2=dev.colmajor((n, npol)),
               h1=dev.colmajor((n, 3, npol)), h2=dev.colmajor((n, 3, npol)))
    _lib.call("mnp_bemret_tangent_eps1", bem.handle, npol,
              *[dev.ptr(dev.base(sig[k])) for k in ("sig1", "sig2", "h1", "h2")],
              *[dev.ptr(dev.base(out[k])) for k in ("sig1", "sig2", "h1", "h2")])
    return compstruct(bem.p, sig.enei, **out)


def _times_i(t):
    out = dev.colmajor(tuple(t.shape))
    out.copy_(t)
    return out.mul_(1j)


def extinction_eps1_derivative(exc, bem, sig):
    """(d ext / d Re eps1, d ext / d Im eps1) per polarisation, FP64."""
    d = eps1_tangent(bem, sig)
    dr = exc.cross_sections(compstruct(bem.p, sig.enei, sig2=d.sig2, h2=d.h2))[0]
    di = exc.cross_sections(compstruct(bem.p, sig.enei, sig2=_times_i(d.sig2), h2=_times_i(d.h2)))[0]
    return np.asarray(dr), np.asarray(di)


def drude_derivatives(mat, enei):
    """d eps / d (eps0, wp, gammad) of an epsdrude at the wavelength
    (Material/@epsdrude/subsref.m:24: eps = eps0 - wp^2 / (w (w + i gammad)))."""
    w = EV2NM / enei
    den = w * (w + 1j * mat.gammad)
    return np.array([1.0, -2 * mat.wp / den, 1j * mat.wp ** 2 * w / den ** 2])


def drude_extinction_gradient(exc, bem, enei):
    """Extinction (npol) and its gradient with respect to the Drude
    parameters (eps0, wp, gammad) of the particle's inside material,
    npol x 3, for the dense solver ``bem``."""
    p = bem.p
    mat = p.eps[p.inout[0, 0] - 1]
    if not isinstance(mat, epsdrude):
        raise TypeError("the particle's inside material must be an epsdrude")
    sig = bem.solve(exc(p, enei))
    ext = np.asarray(exc.cross_sections(sig)[0])
    er, ei = extinction_eps1_derivative(exc, bem, sig)
    de = drude_derivatives(mat, enei)
    return ext, np.outer(er, de.real) + np.outer(ei, de.imag)


class DrudeExtinction(torch.autograd.Function):
    """ext = DrudeExtinction.apply(params, bem, exc, enei): extinction per
    polarisation as a function of params = (eps0, wp, gammad) of the
    particle's inside Drude material (a float64 tensor of 3), differentiable
    with respect to params.  The material of ``bem.p`` is set to params."""

    @staticmethod
    def forward(ctx, params, bem, exc, enei):
        p = bem.p
        idx = p.inout[0, 0] - 1
        e0, wp, ga = (float(v) for v in params.detach().cpu().tolist())
        p.eps[idx] = epsdrude(e0, wp, ga)
        bem.enei = None  # the material changed: initialise again
        ext, grad = drude_extinction_gradient(exc, bem, enei)
        ctx.save_for_backward(torch.as_tensor(grad, dtype=torch.float64))
        return torch.as_tensor(ext, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g.to(grad) @ grad, None, None, None


def solve_transposed(bem, c):
    """lam = A^{-T} c with the factors of ``bem`` at its current wavelength;
    c, lam: complex128 device arrays n x 8 npol, c laid out as the solution
    [sig1 | h1 | sig2 | h2] and lam as the excitation [phi | a | De | alpha]
    (mnp_bemret_solve_transposed)."""
    if not isinstance(bem, bemret):
        raise TypeError("solve_transposed needs the dense retarded solver (bemret)")
    n = bem.p.n
    assert c.shape[0] == n and c.shape[1] % 8 == 0
    lam = dev.colmajor(tuple(c.shape))
    _lib.call("mnp_bemret_solve_transposed", bem.handle, c.shape[1] // 8, dev.ptr(dev.base(c)),
              dev.ptr(dev.base(lam)))
    return lam


def extinction_gradient(exc, bem, sig, weights=None, eps1=True):
    """Gradient of sum_q w_q ext_q (w = 1 by default) for the solution ``sig``
    of the plane wave ``exc`` at its wavelength: (d/d verts, nverts x 3;
    (d/d Re eps1, d/d Im eps1) or None)."""
    if not isinstance(bem, bemret):
        raise TypeError("extinction_gradient needs the dense retarded solver (bemret)")
    bem(sig.enei)
    p = bem.p
    npol = len(exc.pol)
    w = np.ones(npol) if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    pol, dr = np.asfortranarray(exc.pol), np.asfortranarray(exc.dir)
    dv = np.zeros((3, len(p.verts)))  # column-major nverts x 3
    de = np.zeros(2)
    _lib.call("mnp_bemret_extinction_gradient", bem.handle, npol, pol.ctypes.data_as(C.c_void_p),
              dr.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
              *[dev.ptr(dev.base(sig[k])) for k in ("sig1", "sig2", "h1", "h2")],
              dv.ctypes.data_as(C.c_void_p), de.ctypes.data_as(C.c_void_p) if eps1 else None)
    return dv.T.copy(), (de if eps1 else None)


def extinction_jacobian(exc, bem, enei):
    """Extinction per polarisation (npol) and its Jacobians with respect to the
    vertices (npol x nverts x 3) and to the inside Drude parameters
    (eps0, wp, gammad) (npol x 3), one transposed solve per polarisation."""
    p = bem.p
    mat = p.eps[p.inout[0, 0] - 1]
    sig = bem.solve(exc(p, enei))
    ext = np.asarray(exc.cross_sections(sig)[0])
    npol = len(exc.pol)
    jv = np.zeros((npol, len(p.verts), 3))
    jd = np.zeros((npol, 3))
    de_dp = drude_derivatives(mat, enei) if isinstance(mat, epsdrude) else None
    for q in range(npol):
        w = np.zeros(npol)
        w[q] = 1.0
        jv[q], de = extinction_gradient(exc, bem, sig, w)
        if de_dp is not None:
            jd[q] = de[0] * de_dp.real + de[1] * de_dp.imag
    return ext, jv, jd


class ExtinctionAdjoint(torch.autograd.Function):
    """ext = ExtinctionAdjoint.apply(verts, params, faces, pol, dirs, enei,
    eps_out, precision): extinction per polarisation of the closed particle
    (verts: float64 tensor nverts x 3 in nm, faces: 0-based nfaces x 3) with
    inside material epsdrude(*params) (float64 tensor of 3: eps0, wp, gammad)
    in a medium of permittivity eps_out, differentiable with respect to verts
    and params.  The forward solve builds the Green function of the mesh and
    the dense solver of ``precision`` ('fp64', 'mixed' or 'mixed_bf16'); its
    backward contracts the Jacobians of one transposed solve per
    polarisation (reusing the forward factors)."""

    @staticmethod
    def forward(ctx, verts, params, faces, pol, dirs, enei, eps_out=1.0, precision="fp64"):
        from .bem import bemoptions, bemsolver, comparticle, epsconst, planewave
        v = verts.detach().cpu().numpy().astype(np.float64)
        e0, wp, ga = (float(t) for t in params.detach().cpu().tolist())
        op = bemoptions(sim="ret")
        p = comparticle([epsconst(eps_out), epsdrude(e0, wp, ga)], [(v, np.asarray(faces))], [2, 1], 1, op)
        bem = bemsolver(p, op, precision=precision)
        exc = planewave(pol, dirs, op)
        ext, jv, jd = extinction_jacobian(exc, bem, enei)
        ctx.save_for_backward(torch.as_tensor(jv), torch.as_tensor(jd))
        return torch.as_tensor(ext, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        jv, jd = ctx.saved_tensors
        g = g.detach().cpu().to(torch.float64)
        return (torch.einsum("q,qvc->vc", g, jv), g @ jd, None, None, None, None, None, None)
