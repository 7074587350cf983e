"""GPU parity of the quasistatic solver and plane wave above a substrate
(bemstatlayer, planewavestatlayer) through the C ABI, against the oracle's
restatement (image particle, Fresnel field, far field of the induced dipole
above the interface)."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import layer_oracle as L
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)
TH = 0.6
POL = np.array([[1.0, 0, 0], [np.cos(TH), 0, np.sin(TH)], [0, 1.0, 0], [1.0, 0, 0]])
DIRS = np.array([[0, 0, -1.0], [np.sin(TH), 0, -np.cos(TH)], [np.sin(TH), 0, -np.cos(TH)], [0, 0, 1.0]])


def _setup(level, axes, height, eps_sub=2.25, z0=0.0):
    from mnpbem_amd import bem as B
    from mnpbem_amd.layer import layerstructure
    v, f = shapes.ellipsoid(level, axes)
    v = v + np.array([0, 0, height])
    epstab = [B.epsconst(1.0), B.epsdrude(*DRUDE), B.epsconst(eps_sub)]
    layer = layerstructure(epstab, [1, 3], z0)
    op = B.bemoptions(sim="stat", layer=layer)
    p = B.comparticle(epstab, [(v, f)], [2, 1], 1, op)
    return B, p, op, (v, f)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_substrate_matches_oracle(precision, pinfty):
    from mnpbem_amd import device as dev
    from mnpbem_amd.layer import bemstatlayer, planewavestatlayer
    B, p, op, (v, f) = _setup(3, (10.0, 10.0, 15.0), 17.0)
    print(f"[substrate {precision}] n={p.n}", flush=True)
    bem = B.bemsolver(p, op, precision=precision)
    exc = B.planewave(POL, DIRS, op)
    assert isinstance(bem, bemstatlayer) and isinstance(exc, planewavestatlayer)
    g = O.GreenRet(v, f)
    Fr = L.reflected_F(v, f, 0.0)
    for enei in (520.0, 700.0):
        pot = exc(p, enei)
        up, _ = L.field(POL, DIRS, enei, 1.0, 2.25)
        phip = L.potential(g.nvec, up)
        assert np.allclose(dev.to_host(pot.phip), phip, rtol=1e-13, atol=1e-15)
        sig = bem.solve(pot)
        ein = O.eps_drude(enei, *DRUDE)
        want = L.BemStatLayer(g, Fr).initmat(ein, 1.0, 1.0, 2.25).solve(phip)
        got = dev.to_host(sig.sig)
        err = np.abs(got - want).max() / np.abs(want).max()
        dip = L.dipole_moment(g.pos, g.area, want)
        ext_o = L.extinction(dip, POL, DIRS, enei, 1.0, 2.25)
        sca_o = L.scattering(dip, DIRS, enei, 1.0, 2.25, *pinfty)
        ext, sca, ab = exc.cross_sections(sig)
        print(f"  enei {enei}: sig rel err {err:.2e}; ext {ext} vs {ext_o}; stats {bem.stats()}", flush=True)
        assert err < 1e-10
        assert np.allclose(ext, ext_o, rtol=1e-10) and np.allclose(sca, sca_o, rtol=1e-10)
        assert np.allclose(ab, ext - sca)


def test_substrate_mixed_bf16_and_free_limit():
    """5,120 faces: the split-bf16 factorisation with refinement reproduces
    FP64; with the substrate's eps equal to the medium's the solver equals
    the free-space quasistatic solver."""
    B, p, op, (v, f) = _setup(4, (10.0, 10.0, 40.0), 45.0)
    exc = B.planewave(POL, DIRS, op)
    b64 = B.bemsolver(p, op, precision="fp64")
    bbf = B.bemsolver(p, op, precision="mixed_bf16")
    for enei in (560.0, 800.0):
        c64 = np.array(exc.cross_sections(b64.solve(exc(p, enei))))
        cbf = np.array(exc.cross_sections(bbf.solve(exc(p, enei))))
        it, rr = bbf.stats()
        print(f"[substrate mixed_bf16] enei {enei}: sweeps {it}, relres {rr:.2e}, ext {c64[0]}", flush=True)
        assert rr < 1e-12 and np.allclose(cbf, c64, rtol=1e-9)
    Bf, pf, opf, _ = _setup(3, (10.0, 10.0, 15.0), 17.0, eps_sub=1.0)
    from mnpbem_amd import device as dev
    lay = Bf.bemsolver(pf, opf, precision="fp64")
    free = Bf.bemstat(pf, Bf.bemoptions(sim="stat"), precision="fp64")
    pot = Bf.planewave(POL, DIRS, opf)(pf, 600.0)
    a, b = dev.to_host(lay.solve(pot).sig), dev.to_host(free.solve(pot).sig)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-14 * np.abs(b).max())


def test_substrate_argument_errors():
    import ctypes as C
    from mnpbem_amd import _lib
    B, p, op, (v, f) = _setup(2, (10.0, 10.0, 10.0), 5.0)
    with pytest.raises((ValueError, NotImplementedError)):
        B.bemsolver(p, op)  # the particle reaches below the interface
    B, p, op, _ = _setup(2, (10.0, 10.0, 10.0), 15.0)
    lib = _lib.lib()
    h = C.c_void_p()
    rc = lib.mnp_bemstatlayer_create(p.green.handle, p.green.handle, _lib.MNP_SOLVER_FP64, C.byref(h))
    assert rc == _lib.MNP_ERR_ARG and b"image" in lib.mnp_last_error()
