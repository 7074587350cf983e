"""GPU parity of fields on a mesh of points (meshfield) through
mnp_field_points: points inside take sig1, h1 with the inside wavenumber,
points outside sig2, h2 with the outside one, against the oracle's point
fields; points within mindist of the surface are left out."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import dipole_oracle as D
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)


def test_meshfield_inside_and_outside_match_oracle():
    from mnpbem_amd import bem as B
    from mnpbem_amd.meshfield import meshfield
    v, f = shapes.ellipsoid(3, (10.0, 10.0, 15.0))
    op = B.bemoptions(sim="ret")
    eps_b = 1.33
    p = B.comparticle([B.epsconst(eps_b), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, op)
    exc = B.planewave([[1.0, 0, 0], [0, 1.0, 0]], [[0, 0, 1.0], [0, 0, 1.0]], op)
    bem = B.bemsolver(p, op, precision="fp64")
    enei = 600.0
    sig = bem.solve(exc(p, enei))
    x, z = np.meshgrid(np.linspace(-20, 20, 9), np.linspace(-25, 25, 11))
    mf = meshfield(p, x, 0.0, z, mindist=0.5)
    e, h = mf.field(sig)
    assert e.shape == x.shape + (3, 2)
    g = O.GreenRet(v, f)
    k0 = 2 * np.pi / enei
    from mnpbem_amd import device as dev
    hs = {k: dev.to_host(sig[k]) for k in ("sig1", "sig2", "h1", "h2")}
    eo = e.reshape(-1, 3, 2)
    for sel, eps, s, hh in ((mf.outside, eps_b, hs["sig2"], hs["h2"]),
                            (mf.inside, O.eps_drude(enei, *DRUDE), hs["sig1"], hs["h1"])):
        assert sel.sum() > 5
        want, _ = D.field_points(g, mf.pts[sel], s, hh, k0 * np.sqrt(complex(eps)), enei)
        err = np.abs(eo[sel] - want).max() / np.abs(want).max()
        print(f"[meshfield] {int(sel.sum())} points: rel err {err:.2e}", flush=True)
        assert err < 1e-10
    gone = ~(mf.inside | mf.outside)
    assert np.all(np.isnan(eo[gone]))
    # a point on the z axis far above the particle is outside, the centre inside
    assert mf.outside.reshape(x.shape)[-1, 4] and mf.inside.reshape(x.shape)[5, 4]


def test_meshfield_quasistatic_matches_oracle():
    from mnpbem_amd import bem as B
    from mnpbem_amd.meshfield import meshfield
    v, f = shapes.ellipsoid(3, (10.0, 10.0, 15.0))
    op = B.bemoptions(sim="stat")
    p = B.comparticle([B.epsconst(1.33), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, op)
    exc = B.planewave([[1.0, 0, 0], [0, 0, 1.0]], op=op)
    sig = B.bemsolver(p, op, precision="fp64").solve(exc(p, 600.0))
    x, z = np.meshgrid(np.linspace(-20, 20, 7), np.linspace(-25, 25, 9))
    mf = meshfield(p, x, 0.0, z, mindist=0.5)
    e, h = mf.field(sig)
    assert h is None and e.shape == x.shape + (3, 2)
    from mnpbem_amd import device as dev
    g = O.GreenRet(v, f)
    s = dev.to_host(sig.sig)
    sel = mf.inside | mf.outside
    want, _ = D.field_points(g, mf.pts[sel], s, np.zeros((p.n, 3, 2), complex), 0.0, 600.0)
    err = np.abs(e.reshape(-1, 3, 2)[sel] - want).max() / np.abs(want).max()
    print(f"[meshfield stat] {int(sel.sum())} points: rel err {err:.2e}", flush=True)
    assert err < 1e-10
