"""GPU parity of the dipole excitation, fields at points and decay rates
against the oracle and Mie theory, through the C ABI."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import dipole_oracle as D
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)


def _particle(level, radius=10.0, eps_out=1.0, **op):
    from mnpbem_amd import bem as B
    v, f = shapes.icosphere(level, radius)
    opt = B.bemoptions(sim="ret", **op)
    return B, B.comparticle([B.epsconst(eps_out), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, opt), (v, f)


def test_dipole_potential_matches_oracle():
    from mnpbem_amd import device as dev
    B, p, (v, f) = _particle(2)
    pts = np.array([[0.0, 0.0, 15.0], [3.0, -12.0, 4.0]])
    dip = np.array([[1.0, 0, 0], [0.3, 0.4, np.sqrt(0.75)]])
    exc = B.dipoleret(pts, dip)
    enei = 550.0
    got = exc(p, enei)
    g = O.GreenRet(v, f)
    want = D.dipole_potential(g.pos, g.nvec, pts, exc.dip, enei, 1.0)
    print(f"[dipole potential] n={p.n} npol={exc.npol}", flush=True)
    for key, w in zip(("phi2", "phi2p", "a2", "a2p"), want):
        gv = dev.to_host(got[key])
        err = np.abs(gv - w).max() / np.abs(w).max()
        print(f"  {key}: max rel err {err:.2e}", flush=True)
        assert err < 1e-13


@pytest.mark.parametrize("where", ["far", "near"])
def test_field_points_matches_oracle(where):
    """Fields at points from random charges/currents; 'near' points sit
    within the refinement cutoff of the surface (rule-18 Taylor form)."""
    from mnpbem_amd import device as dev
    B, p, (v, f) = _particle(2)
    g = O.GreenRet(v, f)
    pts = (np.array([[0.0, 0.0, 25.0], [24.0, 3.0, -9.0]]) if where == "far"
           else np.array([[0.0, 0.2, 10.6], [6.2, 6.1, -5.3]]))
    rng = np.random.default_rng(4)
    ncol = 3
    sig = rng.normal(size=(g.n, ncol)) + 1j * rng.normal(size=(g.n, ncol))
    h = rng.normal(size=(g.n, 3, ncol)) + 1j * rng.normal(size=(g.n, 3, ncol))
    enei, eps = 600.0, 1.0
    k = 2 * np.pi / enei
    e_ref, h_ref = D.field_points(g, pts, sig, h, k, enei)
    G, _ = D.point_green(g, pts, k)
    x = pts[:, None, :] - g.pos[None]
    d = np.linalg.norm(x, axis=2)
    nnear = int(np.count_nonzero((d - g.rad[None]) / g.rad[None] < 3.0))
    print(f"[field points] {where}: {nnear} refined point-face pairs", flush=True)
    assert (nnear > 0) == (where == "near")
    from mnpbem_amd import _lib
    import ctypes as C
    sd, hd = dev.to_device_colmajor(sig), dev.to_device_colmajor(h)
    e = dev.colmajor((len(pts), 3, ncol))
    hf = dev.colmajor((len(pts), 3, ncol))
    pf = np.asfortranarray(pts)
    _lib.call("mnp_field_points", p.green.handle, enei, _lib.MnpComplex.of(eps), len(pts),
              pf.ctypes.data_as(C.c_void_p), ncol, dev.ptr(dev.base(sd)), dev.ptr(dev.base(hd)),
              dev.ptr(dev.base(e)), dev.ptr(dev.base(hf)))
    for name, got, want in (("e", dev.to_host(e), e_ref), ("h", dev.to_host(hf), h_ref)):
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"  {name}: max rel err {err:.2e}", flush=True)
        assert err < 1e-11


def test_decayrate_dense_fp64_matches_oracle_and_mie(pinfty):
    B, p, (v, f) = _particle(3)
    enei, z = 520.0, 15.0
    pts = np.array([[0.0, 0.0, z]])
    exc = B.dipoleret(pts, np.array([[1.0, 0, 0], [0, 0, 1.0]]))
    bem = B.bemsolver(p, p.op, precision="fp64")
    sig = bem.solve(exc(p, enei))
    tot, rad, rad0 = exc.decayrate(sig)
    # oracle on the same mesh
    g = O.GreenRet(v, f)
    eps_in = O.eps_drude(enei, *DRUDE)
    phi2, phi2p, a2, a2p = D.dipole_potential(g.pos, g.nvec, pts, exc.dip, enei, 1.0)
    ob = O.BemRet(g).initmat(enei, eps_in, 1.0)
    s = ob.solve({"phi2": phi2, "phi2p": phi2p, "a2": a2, "a2p": a2p})
    pn, pa = pinfty
    ef, hf, _ = O.farfield(g.pos, g.area, s["sig2"], s["h2"], pn, enei, 1.0)
    t_ref, r_ref, r0_ref = D.decayrate(g, pts, exc.dip, enei, 1.0, s["sig2"], s["h2"], ef, hf, pn, pa)
    tm, rm = D.mie_decayrate(enei, eps_in, 1.0, 20.0, [z])
    print(f"[decayrate] tot {tot} oracle {t_ref} mie {tm}; rad {rad} oracle {r_ref} mie {rm}", flush=True)
    assert np.allclose(tot, t_ref, rtol=1e-9) and np.allclose(rad, r_ref, rtol=1e-9)
    assert np.allclose(rad0, r0_ref, rtol=1e-14)
    assert np.all(np.abs(tot / tm - 1) < 0.05) and np.all(np.abs(rad / rm - 1) < 0.05)


def test_decayrate_mixed_and_iterative_match_fp64():
    """The FP32 chain + FP64 refinement and GMRES on FP64 H-matrices give the
    decay rates of the dense FP64 solve (5,120 faces)."""
    B, p, _ = _particle(4, radius=20.0)
    enei = 560.0
    rng = np.random.default_rng(9)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    exc = B.dipoleret(u[None] * 25.0, d[None])
    pot = exc(p, enei)
    res = {}
    for name, solver in (("fp64", B.bemsolver(p, p.op, precision="fp64")),
                         ("mixed", B.bemsolver(p, p.op, precision="mixed")),
                         ("iter", B.bemretiter(p, dict(p.op, iter=dict(tol=1e-10, htol=1e-8, cleaf=32))))):
        res[name] = exc.decayrate(solver.solve(pot))
        print(f"[decayrate n={p.n}] {name}: tot {res[name][0].ravel()} rad {res[name][1].ravel()}", flush=True)
    for i in range(2):
        assert np.allclose(res["mixed"][i], res["fp64"][i], rtol=1e-6)
        # H-matrix scheme: no refinement in admissible blocks (as the reference)
        assert np.allclose(res["iter"][i], res["fp64"][i], rtol=1e-2)
