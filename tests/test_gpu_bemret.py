"""GPU parity of the dense retarded solver (FP64 drop-in and FP32 + FP64
refinement) and of the plane-wave observables against the oracle and Mie."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import mnpbem_oracle as O
from oracle.mie import mie_cross_sections

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)


def _setup(verts, faces, eps_out, precision):
    from mnpbem_amd import bem as B
    op = B.bemoptions(sim="ret")
    epstab = [B.epsconst(eps_out), B.epsdrude(*DRUDE)]
    p = B.comparticle(epstab, [(verts, faces)], [2, 1], 1, op)
    return p, B.bemsolver(p, op, precision=precision)


def _oracle(verts, faces, enei, eps_out, pol, dirs, pinfty):
    g = O.GreenRet(verts, faces)
    e1 = O.eps_drude(enei, *DRUDE)
    bem = O.BemRet(g).initmat(enei, e1, eps_out)
    a2, a2p = O.planewave_potential(g.pos, g.nvec, pol, dirs, enei, eps_out)
    sig = bem.solve({"a2": a2, "a2p": a2p})
    ext = O.extinction(g.pos, g.area, sig, pol, dirs, enei, eps_out)
    sca = O.scattering(g.pos, g.area, sig, enei, eps_out, *pinfty)
    return sig, ext, sca, (a2, a2p)


CASES = {
    "sphere_l2": (lambda: shapes.icosphere(2, 10.0), 1.0, 520.0,
                  [[1.0, 0, 0]], [[0, 0, 1.0]]),
    "ellipsoid_l2": (lambda: shapes.ellipsoid(2, (10.0, 10.0, 40.0)), 1.77, 700.0,
                     [[1.0, 0, 0], [0, 0, 1.0]], [[0, 0, 1.0], [1.0, 0, 0]]),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_bf16"])
def test_solver_matches_oracle(case, precision, pinfty):
    from mnpbem_amd import bem as B, device as dev
    mesh, eps_out, enei, pol, dirs = CASES[case]
    print(f"[bemret parity] {case} {precision}", flush=True)
    v, f = mesh()
    pol, dirs = np.array(pol), np.array(dirs)
    sig_ref, ext_ref, sca_ref, (a2_ref, a2p_ref) = _oracle(v, f, enei, eps_out, pol, dirs, pinfty)
    p, bem = _setup(v, f, eps_out, precision)
    exc = B.planewave(pol, dirs)
    pot = exc(p, enei)
    assert np.allclose(dev.to_host(pot.a2), a2_ref, rtol=1e-13, atol=1e-15)
    assert np.allclose(dev.to_host(pot.a2p), a2p_ref, rtol=1e-13, atol=1e-15)
    sig = bem.solve(pot)
    ext, sca, ab = exc.cross_sections(sig)
    if precision == "fp64":
        for key in ("sig1", "sig2", "h1", "h2"):
            got, want = dev.to_host(sig[key]), sig_ref[key]
            assert np.max(np.abs(got - want)) < 1e-9 * np.abs(want).max(), key
        tol = 1e-10
    else:
        it, rr = bem.stats()
        print(f"  refinement sweeps {it}, relres {rr:.2e}", flush=True)
        assert rr < 1e-12 and it <= 4
        tol = 1e-8
    assert np.allclose(ext, ext_ref, rtol=tol)
    assert np.allclose(sca, sca_ref, rtol=tol)
    assert np.allclose(ab, ext_ref - sca_ref, rtol=1e3 * tol)


def test_sphere_1280_faces_against_mie():
    """Config-1 sphere (1,280 faces, 20 nm): BEM spectrum vs Mie within the
    discretisation error, FP64 and mixed agree to 1e-8."""
    from mnpbem_amd import bem as B
    v, f = shapes.icosphere(3, 10.0)
    p64, bem64 = _setup(v, f, 1.0, "fp64")
    _, bemmx = _setup(v, f, 1.0, "mixed")
    exc = B.planewave([[1.0, 0, 0]], [[0, 0, 1.0]])
    enei = np.linspace(400, 800, 5)
    drude = B.epsdrude(*DRUDE)
    for lam in enei:
        print(f"[sphere vs Mie] enei {lam}", flush=True)
        e64 = exc.cross_sections(bem64.solve(exc(p64, lam)))
        emx = exc.cross_sections(bemmx.solve(exc(p64, lam)))
        me, ms, _ = mie_cross_sections(lam, drude(lam)[0], 1.0, 20.0)
        assert np.allclose(emx[0], e64[0], rtol=1e-8) and np.allclose(emx[1], e64[1], rtol=1e-8)
        if lam < 700:
            assert abs(e64[0][0] / me[0] - 1) < 0.03
            assert abs(e64[1][0] / ms[0] - 1) < 0.06


def test_mixed_bf16_matches_fp64_5120_faces():
    """Prolate ellipsoid, 5,120 faces: large enough for the chain's products to
    run on the split-bf16 kernel (checked by its timer).  The refinement brings
    the cross sections to the FP64 path's within 1e-8."""
    import ctypes as C
    from mnpbem_amd import _lib, bem as B
    v, f = shapes.ellipsoid(4, (10.0, 10.0, 40.0))
    pol, dirs = np.array([[1.0, 0, 0], [0, 0, 1.0]]), np.array([[0, 0, 1.0], [1.0, 0, 0]])
    p, bem64 = _setup(v, f, 1.77, "fp64")
    _, bembf = _setup(v, f, 1.77, "mixed_bf16")
    exc = B.planewave(pol, dirs)
    for lam in (550.0, 800.0):
        print(f"[mixed_bf16] n={len(f)} enei {lam}", flush=True)
        _lib.call("mnp_timers_enable", 1)
        sbf = bembf.solve(exc(p, lam))
        tm = _lib.Timers()
        _lib.call("mnp_timers_read", C.byref(tm))
        _lib.call("mnp_timers_enable", 0)
        assert tm.launches[6] > 0  # MNP_TIMER_GEMM_SBF16
        it, rr = bembf.stats()
        print(f"  sweeps {it}, relres {rr:.2e}, split-bf16 launches {tm.launches[6]}", flush=True)
        assert rr < 1e-12 and it <= 6
        cb = exc.cross_sections(sbf)
        c64 = exc.cross_sections(bem64.solve(exc(p, lam)))
        for a, b in zip(cb[:2], c64[:2]):
            assert np.allclose(a, b, rtol=1e-8)


def test_solve_before_initmat_and_bad_polarisation():
    from mnpbem_amd import _lib, bem as B
    v, f = shapes.icosphere(1, 5.0)
    p, bem = _setup(v, f, 1.0, "fp64")
    with pytest.raises(ValueError):
        B.planewave([[1.0, 0, 0]], [[1.0, 0, 0]])
    import ctypes as C
    import torch
    n = p.n
    z = torch.zeros((1, 3, n), dtype=torch.complex128, device="cuda")
    s = torch.zeros((1, n), dtype=torch.complex128, device="cuda")
    st = _lib.lib().mnp_bemret_solve(bem.handle, 1, None, None, None, None, None, None,
                                     C.c_void_p(z.data_ptr()), C.c_void_p(z.data_ptr()),
                                     C.c_void_p(s.data_ptr()), C.c_void_p(s.data_ptr()),
                                     C.c_void_p(z.data_ptr()), C.c_void_p(z.data_ptr()))
    assert st == 6  # MNP_ERR_STATE
