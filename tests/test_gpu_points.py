"""GPU parity of potentials and fields at points (compoint,
compgreenret(pt, p).potential / .field) through mnp_potential_points and
mnp_field_points: each point takes the charges of the faces bordering its
medium and that medium's wavenumber.  Checked against the oracle's point
Green function, and the potentials against the fields through
E = i k0 A - grad phi (central differences of the GPU potentials)."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import dipole_oracle as D
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)


def _solve(enei=600.0, eps_b=1.33):
    from mnpbem_amd import bem as B
    v, f = shapes.ellipsoid(3, (10.0, 10.0, 15.0))
    op = B.bemoptions(sim="ret")
    p = B.comparticle([B.epsconst(eps_b), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, op)
    exc = B.planewave([[1.0, 0, 0], [0, 1.0, 0]], [[0, 0, 1.0], [0, 0, 1.0]], op)
    sig = B.bemsolver(p, op, precision="fp64").solve(exc(p, enei))
    return v, f, p, sig


def _points():
    rng = np.random.default_rng(7)
    out = rng.normal(size=(40, 3))
    out = out / np.linalg.norm(out, axis=1)[:, None] * rng.uniform(17, 30, size=(40, 1))
    ins = rng.uniform(-4, 4, size=(20, 3))
    near = np.array([[0.0, 0.0, 15.6], [10.4, 0.0, 0.0], [0.0, 0.0, 14.3]])  # near the surface
    return np.vstack([out, ins, near])


def test_point_potentials_match_oracle():
    from mnpbem_amd import device as dev
    from mnpbem_amd.points import compgreenret, compoint
    enei, eps_b = 600.0, 1.33
    v, f, p, sig = _solve(enei, eps_b)
    pts = _points()
    pt = compoint(p, pts, mindist=0.2)
    assert pt.n == len(pts) and set(np.unique(pt.inout)) == {1, 2}
    pot = compgreenret(pt, p).potential(sig)
    g = O.GreenRet(v, f)
    hs = {k: dev.to_host(sig[k]) for k in ("sig1", "sig2", "h1", "h2")}
    k0 = 2 * np.pi / enei
    for m, eps, s, hh in ((2, O.eps_drude(enei, *DRUDE), hs["sig1"], hs["h1"]), (1, eps_b, hs["sig2"], hs["h2"])):
        sel = pt.inout == m
        G, _ = D.point_green(g, pt.pos[sel], k0 * np.sqrt(complex(eps)))
        phi = G @ s
        a = np.einsum("pj,jcq->pcq", G, hh)
        e1 = np.abs(pot["phi1"][sel] - phi).max() / np.abs(phi).max()
        e2 = np.abs(pot["a1"][sel] - a).max() / np.abs(a).max()
        print(f"[points] medium {m}: {int(sel.sum())} points, phi rel err {e1:.2e}, a rel err {e2:.2e}", flush=True)
        assert e1 < 1e-10 and e2 < 1e-10


def test_point_potentials_give_the_fields():
    """E = i k0 A - grad phi at points away from the surface, the gradient by
    central differences (step 1e-3 nm) of the GPU potentials."""
    from mnpbem_amd.points import compgreenret, compoint
    enei = 600.0
    _, _, p, sig = _solve(enei)
    pts = _points()[:60]
    pt = compoint(p, pts, mindist=1.0)
    gf = compgreenret(pt, p)
    e = gf.field(sig)["e"]
    a = gf.potential(sig)["a1"]
    h = 1e-3
    grad = np.zeros_like(e)
    for c in range(3):
        dx = np.zeros(3)
        dx[c] = h
        up = compoint(p, pt.pos + dx)
        dn = compoint(p, pt.pos - dx)
        assert np.array_equal(up.inout, pt.inout) and np.array_equal(dn.inout, pt.inout)
        grad[:, c] = (compgreenret(up, p).potential(sig)["phi1"] - compgreenret(dn, p).potential(sig)["phi1"]) / (2 * h)
    want = 1j * (2 * np.pi / enei) * a - grad
    err = np.abs(e - want).max() / np.abs(e).max()
    print(f"[points] E vs i k0 A - grad phi: rel err {err:.2e}", flush=True)
    assert err < 1e-6
