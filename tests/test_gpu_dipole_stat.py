"""GPU parity of the quasistatic dipole excitation (dipolestat) and of
dipoles above a substrate (dipolestatlayer) through the C ABI: potentials,
the surface-charge field at the dipoles and the decay rates against the
oracle's restatement, for FP64 and mixed solvers."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import dipole_oracle as D
from oracle import layer_oracle as L
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)
PTS = np.array([[0.0, 0.0, 16.0], [4.0, 1.0, 13.5]])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dipolestat_matches_oracle(precision):
    from mnpbem_amd import bem as B
    from mnpbem_amd import device as dev
    v, f = shapes.ellipsoid(3, (10.0, 10.0, 12.0))
    op = B.bemoptions(sim="stat")
    eps_b = 1.5
    p = B.comparticle([B.epsconst(eps_b), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, op)
    exc = B.dipole(PTS, None, op)
    assert isinstance(exc, B.dipolestat)
    bem = B.bemsolver(p, op, precision=precision)
    g = O.GreenRet(v, f)
    dip = np.repeat(np.eye(3)[None, :, :], len(PTS), axis=0)
    print(f"[dipolestat {precision}] n={p.n}", flush=True)
    for enei in (520.0, 640.0):
        pot = exc(p, enei)
        phip = D.dipolestat_potential(g.nvec, D.dipolestat_field(g.pos, PTS, dip, eps_b))
        assert np.allclose(dev.to_host(pot.phip), phip, rtol=1e-13, atol=1e-15 * np.abs(phip).max())
        sig = bem.solve(pot)
        want = O.BemStat(g).initmat(enei, O.eps_drude(enei, *DRUDE), eps_b).solve(phip)
        err = np.linalg.norm(dev.to_host(sig.sig) - want) / np.linalg.norm(want)
        tot, rad, rad0 = exc.decayrate(sig)
        tw, rw, r0w = D.dipolestat_decayrate(g, PTS, dip, enei, eps_b, want)
        print(f"  enei {enei}: sig rel err {err:.2e}; tot {tot[0]} vs {tw[0]}", flush=True)
        assert err < 1e-10
        assert np.allclose(tot, tw, rtol=1e-9) and np.allclose(rad, rw, rtol=1e-9) and np.allclose(rad0, r0w)


def _layer_setup():
    from mnpbem_amd import bem as B
    from mnpbem_amd.layer import layerstructure
    v, f = shapes.ellipsoid(3, (10.0, 10.0, 15.0))
    v = v + np.array([0, 0, 17.0])
    epstab = [B.epsconst(1.0), B.epsdrude(*DRUDE), B.epsconst(-8.0 + 1.5j)]
    layer = layerstructure(epstab, [1, 3], 0.0)
    op = B.bemoptions(sim="stat", layer=layer)
    p = B.comparticle(epstab, [(v, f)], [2, 1], 1, op)
    return B, p, op, (v, f)


@pytest.mark.parametrize("image_eps", ["upper", "reference"])
def test_dipolestatlayer_matches_oracle(image_eps):
    from mnpbem_amd import device as dev
    from mnpbem_amd.layer import dipolestatlayer
    B, p, op, (v, f) = _layer_setup()
    pts = PTS + np.array([0, 0, 20.0])  # above the particle (top at z = 32)
    exc = B.dipole(pts, None, op, image_eps=image_eps)
    assert isinstance(exc, dipolestatlayer)
    bem = B.bemsolver(p, op, precision="fp64")
    g = O.GreenRet(v, f)
    V, Fc = L.image_mesh(v, f, 0.0)
    gi = O.GreenRet(V, Fc, closed=False)
    Fr = L.reflected_F(v, f, 0.0)
    dip = np.repeat(np.eye(3)[None, :, :], len(pts), axis=0)
    eu, ed = 1.0, -8.0 + 1.5j
    print(f"[dipolestatlayer {image_eps}] n={p.n}", flush=True)
    for enei in (560.0, 700.0):
        pot = exc(p, enei)
        e = L.dipolestatlayer_field(g.pos, pts, dip, eu, ed, 0.0, image_eps=image_eps)
        phip = D.dipolestat_potential(g.nvec, e)
        assert np.allclose(dev.to_host(pot.phip), phip, rtol=1e-13, atol=1e-15 * np.abs(phip).max())
        sig = bem.solve(pot)
        want = L.BemStatLayer(g, Fr).initmat(O.eps_drude(enei, *DRUDE), eu, eu, ed).solve(phip)
        err = np.linalg.norm(dev.to_host(sig.sig) - want) / np.linalg.norm(want)
        tot, rad, rad0 = exc.decayrate(sig)
        tw, rw, r0w = L.dipolestatlayer_decayrate(g, gi, pts, dip, enei, eu, ed, want, exc.pinf_nvec, exc.pinf_area)
        t0, r0, _ = exc.decayrate0(enei)
        t0w, r0ww, _ = L.dipolestatlayer_decayrate0(pts, dip, enei, eu, ed, 0.0, exc.pinf_nvec, exc.pinf_area,
                                                    image_eps=image_eps)
        print(f"  enei {enei}: sig rel err {err:.2e}; tot {tot[0]} vs {tw[0]}; rad {rad[0]}", flush=True)
        assert err < 1e-10
        assert np.allclose(tot, tw, rtol=1e-9) and np.allclose(rad, rw, rtol=1e-9) and np.allclose(rad0, r0w)
        assert np.allclose(t0, t0w, rtol=1e-12) and np.allclose(r0, r0ww, rtol=1e-12)


def test_dipolestatlayer_mixed_and_errors():
    from mnpbem_amd.layer import dipolestatlayer
    B, p, op, _ = _layer_setup()
    pts = PTS + np.array([0, 0, 20.0])
    exc = B.dipole(pts, None, op)
    b64 = B.bemsolver(p, op, precision="fp64")
    bmx = B.bemsolver(p, op, precision="mixed")
    for enei in (600.0,):
        a = np.array(exc.decayrate(b64.solve(exc(p, enei)))[:2])
        b = np.array(exc.decayrate(bmx.solve(exc(p, enei)))[:2])
        assert np.allclose(a, b, rtol=1e-9)
    with pytest.raises(ValueError):
        dipolestatlayer(np.array([[0, 0, -1.0]]), None, op)
    with pytest.raises(ValueError):
        dipolestatlayer(pts, None, op, image_eps="lower")


@pytest.mark.parametrize("sym", ["x", "y", "xy"])
def test_dipolestatmirror_equals_full_particle(sym):
    """Dipoles near a mirror-symmetric particle: the expanded mirror solution
    and its decay rates equal the oracle's dipolestat on the full particle
    and the GPU dipolestat on the full particle."""
    from mnpbem_amd import bem as B
    from mnpbem_amd import device as dev
    from oracle import mirror_oracle as M
    seg = {"xy": (0.0, np.pi / 2), "y": (0.0, np.pi), "x": (-np.pi / 2, np.pi / 2)}[sym]
    mesh = shapes.ellipsoid_segment(4 if sym == "xy" else 8, 10, (10.0, 12.0, 20.0), *seg)
    op = B.bemoptions(sim="stat", sym=sym)
    eps_b = 1.2
    p = B.comparticlemirror([B.epsconst(eps_b), B.epsdrude(*DRUDE)], [mesh], [2, 1], 1, op)
    pts = np.array([[3.0, 4.0, 24.0], [14.0, 2.0, 3.0]])
    dip = np.array([[1.0, 0, 0], [0, 0.6, 0.8]])
    exc = B.dipole(pts, dip, op)
    assert isinstance(exc, B.dipolestatmirror)
    bem = B.bemsolver(p, op, precision="fp64")
    enei = 580.0
    sig = bem.solve(exc(p, enei))
    full = exc.full(sig)
    tot, rad, _ = exc.decayrate(sig)
    V, F = M.full_mesh(*mesh, sym)
    g = O.GreenRet(V, F)
    d3 = np.repeat(dip.T[None, :, :], len(pts), axis=0)
    phip = D.dipolestat_potential(g.nvec, D.dipolestat_field(g.pos, pts, d3, eps_b))
    want = O.BemStat(g).initmat(enei, O.eps_drude(enei, *DRUDE), eps_b).solve(phip)
    err = np.linalg.norm(dev.to_host(full.sig) - want) / np.linalg.norm(want)
    tw, rw, _ = D.dipolestat_decayrate(g, pts, d3, enei, eps_b, want)
    print(f"[dipolestatmirror {sym}] part {p.n}: sig rel err {err:.2e}; tot {tot.ravel()} vs {tw.ravel()}",
          flush=True)
    assert err < 1e-9
    assert np.allclose(tot, tw, rtol=1e-9) and np.allclose(rad, rw, rtol=1e-9)


@pytest.mark.parametrize("sym", ["x", "xy"])
def test_dipoleretmirror_equals_full_particle(sym):
    """Retarded dipoles near a mirror-symmetric particle: the expanded
    charges equal the oracle's bemret solution on the full particle and the
    decay rates equal the GPU dipoleret's on the full particle."""
    from mnpbem_amd import bem as B
    from mnpbem_amd import device as dev
    from oracle import mirror_oracle as M
    seg = {"xy": (0.0, np.pi / 2), "x": (-np.pi / 2, np.pi / 2)}[sym]
    mesh = shapes.ellipsoid_segment(4 if sym == "xy" else 8, 10, (10.0, 12.0, 20.0), *seg)
    op = B.bemoptions(sim="ret", sym=sym)
    eps_b = 1.2
    p = B.comparticlemirror([B.epsconst(eps_b), B.epsdrude(*DRUDE)], [mesh], [2, 1], 1, op)
    pts = np.array([[3.0, 4.0, 24.0], [14.0, 2.0, 3.0]])
    dip = np.array([[1.0, 0, 0], [0, 0.6, 0.8]])
    exc = B.dipole(pts, dip, op)
    assert isinstance(exc, B.dipoleretmirror)
    bem = B.bemsolver(p, op, precision="fp64")
    enei = 580.0
    sig = bem.solve(exc(p, enei))
    full = exc.full(sig)
    tot, rad, _ = exc.decayrate(sig)
    V, F = M.full_mesh(*mesh, sym)
    g = O.GreenRet(V, F)
    d3 = np.repeat(dip.T[None, :, :], len(pts), axis=0)
    phi, phip, a, ap = D.dipole_potential(g.pos, g.nvec, pts, d3, enei, eps_b)
    want = O.BemRet(g).initmat(enei, O.eps_drude(enei, *DRUDE), eps_b).solve(
        {"phi2": phi, "phi2p": phip, "a2": a, "a2p": ap})
    for key in ("sig2", "h2"):
        err = np.linalg.norm(dev.to_host(full[key]) - want[key]) / np.linalg.norm(want[key])
        print(f"[dipoleretmirror {sym}] {key} rel err {err:.2e}", flush=True)
        assert err < 1e-9
    fop = B.bemoptions(sim="ret")
    pf = B.comparticle(p.eps, [(p.full().verts, p.full().faces)], [2, 1], 1, fop)
    excf = B.dipole(pts, dip, fop)
    tf, rf, _ = excf.decayrate(B.bemsolver(pf, fop, precision="fp64").solve(excf(pf, enei)))
    print(f"  tot {tot.ravel()} vs full {tf.ravel()}", flush=True)
    assert np.allclose(tot, tf, rtol=1e-8) and np.allclose(rad, rf, rtol=1e-8)
