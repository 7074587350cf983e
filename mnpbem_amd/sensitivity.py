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

The derivatives with respect to vertex positions are not provided.
"""
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
