"""GPU parity of the retarded substrate path through the C ABI against the
oracle's restatement (oracle/layerret_oracle.py): the Sommerfeld tables,
their interpolation in wavelength, the reflected Green matrices with the
near-pair and diagonal refinement, the bemretlayer solution, the plane wave's
potentials and the cross sections; with equal media the solution of the free
particle."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import layerret_oracle as L
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)
POL = np.array([[1.0, 0, 0], [0.6, 0, 0.8], [0, 1.0, 0], [0.8, 0, -0.6]])
DIR = np.array([[0, 0, -1.0], [0.8, 0, -0.6], [0.8, 0, -0.6], [0.6, 0, 0.8]])
NTAB = 10


def _setup(eps_down, level=2, radius=5.0, gap=1.0, ntab=NTAB, eps_up=1.0):
    from mnpbem_amd import bem as B
    from mnpbem_amd import layer as LY
    v, f = shapes.icosphere(level, radius)
    v = v + [0, 0, radius + gap]
    epstab = [B.epsconst(eps_up), B.epsdrude(*DRUDE), B.epsconst(eps_down)]
    layer = LY.layerstructure(epstab, [1, 3], 0.0)
    p = B.comparticle(epstab, [(v, f)], [2, 1], 1, B.bemoptions(sim="ret"))
    tab = LY.compgreentablayer(layer, layer.tabspace(p, nr=ntab, nz=ntab))
    op = B.bemoptions(sim="ret", layer=layer, greentab=tab)
    sub = L.Substrate(eps_up, eps_down, 0.0)
    otab = L.GreenTable(sub, tab.r, tab.z1, tab.z2)
    return B, LY, p, op, tab, sub, otab, (v, f)


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("eps_down", [2.25, -10 + 1.2j])
def test_tables_match_oracle(eps_down):
    """The Sommerfeld tables (both legs: the grid has r / Z from below 1e-3
    to above 2) equal the oracle's quadrature with scipy's Bessel functions."""
    B, LY, p, op, tab, sub, otab, _ = _setup(eps_down)
    rg = np.asarray(tab.r)
    Zt = sub.mindist(tab.z1) + sub.mindist(tab.z2)
    assert (rg[:, None] > 2 * Zt[None, :]).any() and (rg[:, None] <= 2 * Zt[None, :]).any()
    for lam in (450.0, 700.0):
        got = tab.values(lam)
        want = otab.compute(lam)[1]  # (3, 5, nr, nz)
        for w in range(3):
            for c in range(5):
                e = _rel(got[:, :, 5 * w + c], want[w, c])
                assert e < 1e-10, (lam, w, c, e)
    print(f"[layerret] tables eps_down={eps_down}: {got.shape} match", flush=True)


def test_tables_interpolate_in_wavelength():
    """set() tables at 500 and 600 nm; at 530 nm the table is the linear
    interpolation (greentablayer/eval.m)."""
    from mnpbem_amd import layer as LY
    B, _, p, op, tab, sub, otab, _ = _setup(2.25, ntab=6)
    t2 = LY.compgreentablayer(tab.layer, {"r": tab.r, "z1": tab.z1, "z2": tab.z2}).set([500.0, 600.0])
    otab.set([500.0, 600.0])
    got = t2.values(530.0)
    want = otab.normalised(530.0)
    assert _rel(np.moveaxis(got.reshape(len(tab.r), len(tab.z1), 3, 5), (2, 3), (0, 1)), want) < 1e-10
    with pytest.raises(Exception):
        t2.values(700.0)


def test_reflected_matrices_match_oracle():
    """Reflected G and F of the five components, with the shape-function
    refinement of near pairs and the polar rule on diagonal elements."""
    from mnpbem_amd import device as dev
    B, LY, p, op, tab, sub, otab, (v, f) = _setup(-10 + 1.2j)
    gl = LY.greenretlayer(p, tab)
    og = L.GreenRetLayer(sub, v, f, otab)
    npairs, ndiag = gl.refined()
    assert npairs == int((og.ir == 1).sum()) and ndiag == int((np.diag(og.ir) == 2).sum()) and ndiag > 0
    lam = 560.0
    G, F = gl.eval(lam)
    G, F = dev.to_host(G), dev.to_host(F)
    ref = og.eval(lam)
    for c in range(5):
        eg = _rel(G[:, :, c], ref["G"][c])
        ef = _rel(F[:, :, c], ref["F"][c])
        print(f"[layerret] component {L.COMPONENTS[c]}: G {eg:.1e} F {ef:.1e}", flush=True)
        assert eg < 1e-10 and ef < 1e-10


@pytest.mark.parametrize("eps_down", [2.25, -10 + 1.2j])
def test_solver_matches_oracle(eps_down, pinfty):
    """sig1, sig2, h1, h2 and the cross sections for light from above (normal
    and oblique TM and TE) and from below, at two wavelengths."""
    from mnpbem_amd import device as dev
    from mnpbem_amd.layer import bemretlayer, planewaveretlayer
    B, LY, p, op, tab, sub, otab, (v, f) = _setup(eps_down)
    bem = B.bemsolver(p, op)
    exc = B.planewave(POL, DIR, op)
    assert isinstance(bem, bemretlayer) and isinstance(exc, planewaveretlayer)
    g = O.GreenRet(v, f)
    ob = L.BemRetLayer(g, L.GreenRetLayer(sub, v, f, otab))
    for lam in (520.0, 640.0):
        ex = exc(p, lam)
        phi, phip, a, ap = L.planewave_potential(sub, g.pos, g.nvec, POL, DIR, lam)
        for key, want in (("phi2", phi), ("phi2p", phip), ("a2", a), ("a2p", ap)):
            assert _rel(dev.to_host(ex[key]), want) < 1e-12, key
        sig = bem.solve(ex)
        ref = ob.initmat(lam, O.eps_drude(lam, *DRUDE)).solve({"phi2": phi, "phi2p": phip, "a2": a, "a2p": ap})
        e_sig = max(_rel(dev.to_host(sig[k]), ref[k]) for k in ref)
        ext, sca, _ = exc.cross_sections(sig)
        oext = L.extinction(sub, g.pos, g.area, ref, POL, DIR, lam)
        osca = L.scattering(sub, g.pos, g.area, ref, DIR, lam, *pinfty)
        # light from below through the lossy metal has no far field: ext 0
        e_cs = max(np.max(np.abs(ext - oext)) / np.max(np.abs(oext)), np.max(np.abs(sca / osca - 1)))
        print(f"[layerret] eps_down={eps_down} {lam:.0f} nm: sig {e_sig:.1e}, cross sections {e_cs:.1e}, "
              f"ext {np.round(ext, 4)}", flush=True)
        assert e_sig < 1e-9 and e_cs < 1e-9


def test_equal_media_is_the_free_particle():
    """eps_down = eps_up: the layer solver and plane wave give the free
    particle's solution and cross sections (bemret, planewaveret)."""
    from mnpbem_amd import device as dev
    B, LY, p, op, tab, sub, otab, _ = _setup(1.77, eps_up=1.77, ntab=6)
    lam = 560.0
    sig = B.bemsolver(p, op).solve(B.planewave(POL[:3], DIR[:3], op)(p, lam))
    pf = B.comparticle([B.epsconst(1.77), B.epsdrude(*DRUDE)], [(p.verts, p.faces)], [2, 1], 1,
                       B.bemoptions(sim="ret"))
    excf = B.planewave(POL[:3], DIR[:3], B.bemoptions(sim="ret"))
    ref = B.bemret(pf, precision="fp64").solve(excf(pf, lam))
    for k in ("sig1", "sig2", "h1", "h2"):
        assert _rel(dev.to_host(sig[k]), dev.to_host(ref[k])) < 1e-10
    a = B.planewave(POL[:3], DIR[:3], op).cross_sections(sig)
    b = excf.cross_sections(ref)
    assert np.allclose(a[0], b[0], rtol=1e-10) and np.allclose(a[1], b[1], rtol=1e-10)


@pytest.mark.parametrize("eps_down", [2.25, -10 + 1.2j])
def test_mixed_matches_fp64(eps_down):
    """The complex FP32 chain refined with the FP64 operator gives the FP64
    solution: sig, h to 1e-10 and cross sections to 1e-10 in a few sweeps."""
    from mnpbem_amd import device as dev
    B, LY, p, op, tab, sub, otab, _ = _setup(eps_down, level=3)
    exc = B.planewave(POL, DIR, op)
    b64 = B.bemsolver(p, op, precision="fp64")
    bmx = B.bemsolver(p, op, precision="mixed")
    for lam in (520.0, 640.0):
        ex = exc(p, lam)
        s64, smx = b64.solve(ex), bmx.solve(ex)
        e_sig = max(_rel(dev.to_host(smx[k]), dev.to_host(s64[k])) for k in ("sig1", "sig2", "h1", "h2"))
        c64, cmx = exc.cross_sections(s64), exc.cross_sections(smx)
        e_cs = max(np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(cmx[:2], c64[:2]))
        st = bmx.stats()
        print(f"[layerret] mixed eps_down={eps_down} {lam:.0f} nm: sig {e_sig:.1e}, cross sections {e_cs:.1e}, "
              f"sweeps {st['iterations']}, relres {st['relres']:.1e}", flush=True)
        assert e_sig < 1e-10 and e_cs < 1e-10 and st["relres"] <= 1e-12 and st["iterations"] <= 6


def test_point_fields_match_oracle():
    """compgreenretlayer(pt, p): potentials and fields at points above the
    substrate (far, near the surface, near the interface below the particle
    with refined pairs, and inside the particle, which see no reflection)
    equal the oracle's direct plus reflected parts."""
    from mnpbem_amd import device as dev
    from mnpbem_amd.layer import compgreenretlayer
    from mnpbem_amd.points import compoint
    from oracle import dipole_oracle as D
    B, LY, p, op, tab, sub, otab, (v, f) = _setup(-10 + 1.2j)
    pts = np.array([[0.0, 0.0, 14.0], [9.0, 2.0, 7.0], [-3.0, 12.0, 20.0], [0.0, 0.0, 11.8], [5.6, 0.0, 6.0],
                    [3.0, 3.0, 1.5], [1.5, 0.0, 0.6], [0.0, 0.0, 0.3], [0.0, 0.0, 6.0], [1.0, -1.0, 4.0]])
    pt = compoint(p, pts, mindist=0.2)
    assert pt.n == len(pts) and set(np.unique(pt.inout)) == {1, 2}
    up = pt.inout == 1
    tabp = LY.compgreentablayer(tab.layer, tab.layer.tabspace(p, pt.pos[up], nr=NTAB, nz=NTAB))
    gf = compgreenretlayer(pt, p, dict(op, greentab=tabp))
    otp = L.GreenTable(sub, tabp.r, tabp.z1, tabp.z2)
    pg = L.PointGreenLayer(sub, pt.pos[up], v, f, otp)
    assert pg.ir.any()  # some pairs are refined
    g = O.GreenRet(v, f)
    lam = 600.0
    exc = B.planewave(POL[:2], DIR[:2], op)
    sig = B.bemsolver(p, op).solve(exc(p, lam))
    hs = {k: dev.to_host(sig[k]) for k in ("sig1", "sig2", "h1", "h2")}
    refl = L.points_reflected(pg, lam, hs)
    fld, pot = gf.field(sig), gf.potential(sig)
    k0 = 2 * np.pi / lam
    for m, eps, s, hh in ((1, 1.0, hs["sig2"], hs["h2"]), (2, O.eps_drude(lam, *DRUDE), hs["sig1"], hs["h1"])):
        sel = pt.inout == m
        k = k0 * np.sqrt(complex(eps))
        e, h = D.field_points(g, pt.pos[sel], s, hh, k, lam)
        G, _ = D.point_green(g, pt.pos[sel], k)
        phi, a = G @ s, np.einsum("pj,jcq->pcq", G, hh)
        if m == 1:
            e, h = e + refl["e"], h + refl["h"]
            phi, a = phi + refl["phi"], a + refl["a"]
        errs = [_rel(fld["e"][sel], e), _rel(fld["h"][sel], h), _rel(pot["phi1"][sel], phi), _rel(pot["a1"][sel], a)]
        print(f"[layerret] points in medium {m}: e {errs[0]:.1e} h {errs[1]:.1e} phi {errs[2]:.1e} a {errs[3]:.1e}",
              flush=True)
        assert max(errs) < 1e-10
    # the reflected part alone is a sizeable share at the points above
    assert np.abs(refl["e"]).max() > 1e-3 * np.abs(fld["e"][up]).max()


def _strict(r1, r2, d):
    """Admissibility that keeps every near pair of the 1,280-face sphere in a
    full block (the default 2.5 min(rad) < dist puts some refined pairs into
    admissible blocks, where ACA sees the unrefined kernel, as in the
    reference's H-matrices)."""
    return 6.0 * min(r1, r2) < d


@pytest.mark.parametrize("hprec,tol", [("fp64", 1e-7), ("fp32", 1e-5)])
def test_iterative_matches_dense(hprec, tol):
    """bemretlayeriter (five H-matrices per operator, GMRES) against the dense
    FP64 bemretlayer and the oracle, air above (where the dense chain's
    operator and the iterative one coincide: hs = sh) and a lossy metal
    below, 1,280 faces: sig, h and the cross sections."""
    from mnpbem_amd import device as dev
    B, LY, p, op, tab, sub, otab, (v, f) = _setup(-10 + 1.2j, level=3)
    dense = B.bemsolver(p, op, precision="fp64")
    it = B.bemsolver(p, dict(op, iter=dict(cleaf=32, htol=1e-10, tol=1e-10, maxit=300, hprecision=hprec,
                                           fadmiss=_strict)))
    assert isinstance(it, LY.bemretlayeriter)
    exc = B.planewave(POL, DIR, op)
    g = O.GreenRet(v, f)
    ob = L.BemRetLayer(g, L.GreenRetLayer(sub, v, f, otab))
    for lam in (520.0, 640.0):
        ex = exc(p, lam)
        a, b = dense.solve(ex), it.solve(ex)
        st = it.stats()
        e = max(_rel(dev.to_host(b[k]), dev.to_host(a[k])) for k in ("sig1", "sig2", "h1", "h2"))
        ca, cb = exc.cross_sections(a), exc.cross_sections(b)
        ecs = max(np.max(np.abs(x - y)) / np.max(np.abs(y)) for x, y in zip(cb[:2], ca[:2]))
        print(f"[layerret] iterative {hprec} {lam:.0f} nm: {st['iter']} iterations, flag {st['flag']}, "
              f"sig {e:.1e}, cross sections {ecs:.1e}", flush=True)
        assert st["flag"] == 0 and e < tol and ecs < tol
        if lam == 520.0 and hprec == "fp64":
            phi, phip, aa, ap = L.planewave_potential(sub, g.pos, g.nvec, POL, DIR, lam)
            ref = ob.initmat(lam, O.eps_drude(lam, *DRUDE)).solve({"phi2": phi, "phi2p": phip, "a2": aa, "a2p": ap})
            eo = max(_rel(dev.to_host(b[k]), ref[k]) for k in ref)
            print(f"[layerret] iterative vs oracle: {eo:.1e}", flush=True)
            assert eo < tol


@pytest.mark.parametrize("key", ["G", "H2"])
def test_layer_hmatrix_components(key):
    """The substrate H-matrices with admissible blocks (default admissibility,
    1,280 faces): each reflected component equals the dense reflected matrix
    of greenretlayer to the ACA tolerance, and with the direct part the
    H-matrix minus the direct H-matrix equals it too (ACA of the sum
    kernel)."""
    import torch
    from mnpbem_amd import device as dev
    from mnpbem_amd import hmatrix as H
    B, LY, p, op, tab, sub, otab, _ = _setup(-10 + 1.2j, level=3)
    gl = LY.greenretlayer(p, tab)
    lam = 560.0
    G, F = gl.eval(lam)
    D = dev.to_host(G if key == "G" else F)
    g = p.green
    tree = H.clustertree(g.geometry()["pos"], cleaf=32)
    rng = np.random.default_rng(5)
    xh = rng.normal(size=(g.n, 2)) + 1j * rng.normal(size=(g.n, 2))
    x = torch.as_tensor(xh, device="cuda")
    k = 2 * np.pi / lam
    mk = lambda: H.hmatrix(g, tree, htol=1e-10, kmax=100, precision="fp64")
    yd = dev.to_host(mk().fill(k, key).matvec(x))
    for c, nm in enumerate(L.COMPONENTS):
        want = D[:, :, c] @ xh
        hr = mk().fill_layer(k, key, nm, False, gl, lam)
        assert hr.stats()["lowrank_entries"] > 0  # admissible blocks present
        er = _rel(dev.to_host(hr.matvec(x)), want)
        ed = _rel(dev.to_host(mk().fill_layer(k, key, nm, True, gl, lam).matvec(x)) - yd, want)
        print(f"[layerret] H-matrix {key}.{nm}: reflected {er:.1e}, with direct {ed:.1e}", flush=True)
        assert er < 1e-8 and ed < 1e-7
