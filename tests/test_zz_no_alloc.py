"""mnp_planewave_cross_sections_noalloc (scratch in a file-scope device array,
no device allocation) against mnp_planewave_cross_sections on the same
solution: two calls in a row, the second with fewer polarisations and fewer
scattering directions, FP64 to 1e-12 relative; and its size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _xsec(name, p, exc, sig, npol, nvec, area):
    from mnpbem_amd import _lib, device as dev
    out = [np.zeros(npol) for _ in range(3)]
    pol = np.asfortranarray(exc.pol[:npol])
    dr = np.asfortranarray(exc.dir[:npol])
    sig2 = dev.base(sig.sig2).contiguous()
    h2 = dev.base(sig.h2).contiguous()
    _lib.call(name, p.green.handle, float(sig.enei), _lib.MnpComplex.of(exc._eps_med(p, sig.enei)), npol,
              pol.ctypes.data_as(C.c_void_p), dr.ctypes.data_as(C.c_void_p), area.numel(),
              dev.ptr(dev.base(nvec)), dev.ptr(area), dev.ptr(sig2), dev.ptr(h2),
              *[o.ctypes.data_as(C.c_void_p) for o in out])
    return np.array(out)


def test_noalloc_cross_sections_match_allocating_entry_point():
    import torch

    from mnpbem_amd import _lib, bem as B, device as dev, shapes

    dev.init(0)
    v, f = shapes.icosphere(2, 10.0)
    op = B.bemoptions(sim="ret")
    p = B.comparticle([B.epsconst(1.0), B.epsdrude(10.0, 9.065, 0.0708)], [(v, f)], [2, 1], 1, op)
    exc = B.planewave([[1.0, 0, 0], [0, 1.0, 0]], [[0, 0, 1.0], [0, 0, 1.0]], op)
    sig = B.bemsolver(p, op, precision="fp64").solve(exc(p, 520.0))
    nvec, area = exc.pinf_nvec, exc.pinf_area
    print(f"[no_alloc] n={p.n} npinf={area.numel()}", flush=True)
    # call 1: both polarisations, all scattering directions
    want = _xsec("mnp_planewave_cross_sections", p, exc, sig, 2, nvec, area)
    got = _xsec("mnp_planewave_cross_sections_noalloc", p, exc, sig, 2, nvec, area)
    print(f"[no_alloc] call 1 ext/sca/abs {got.tolist()} vs {want.tolist()}", flush=True)
    assert np.all(np.abs(got - want) <= 1e-12 * np.abs(want).max())
    # call 2, smaller: the first polarisation and the first 200 directions
    m = 200
    nv2 = dev.to_device_colmajor(dev.to_host(nvec)[:m], dtype=torch.float64)
    ar2 = area[:m].contiguous()
    sig1 = B.compstruct(p, sig.enei, sig2=sig.sig2[:, :1], h2=sig.h2[:, :, :1])
    want2 = _xsec("mnp_planewave_cross_sections", p, exc, sig1, 1, nv2, ar2)
    got2 = _xsec("mnp_planewave_cross_sections_noalloc", p, exc, sig1, 1, nv2, ar2)
    print(f"[no_alloc] call 2 ext/sca/abs {got2.tolist()} vs {want2.tolist()}", flush=True)
    assert np.all(np.abs(got2 - want2) <= 1e-12 * np.abs(want2).max())
    assert abs(got2[0, 0] - got[0, 0]) <= 1e-12 * abs(got[0, 0])  # extinction needs no pinfty
    # beyond the scratch's size: rejected on the host before any GPU work
    big = torch.zeros(4096, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.MnpError, match="at most 4096"):
        _xsec("mnp_planewave_cross_sections_noalloc", p, exc, sig, 2,
              dev.to_device_colmajor(np.zeros((4096, 3)), dtype=torch.float64), big)
