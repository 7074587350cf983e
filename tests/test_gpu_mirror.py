"""GPU parity of the retarded solver with mirror symmetry (bemretmirror) and
its plane wave (planewaveretmirror) through the C ABI: surface charges and
currents of each excitation against the oracle's restatement of
bemretmirror, cross sections against the oracle's dense solver on the full
particle, the mixed and split-bf16 variants against FP64, and the mirror
solution against the dense GPU solver on the full particle."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import mirror_oracle as M
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)
SEGMENT = {"xy": (0.0, np.pi / 2), "y": (0.0, np.pi), "x": (-np.pi / 2, np.pi / 2)}
POL = np.array([[1.0, 0, 0], [0, 1.0, 0], [0.6, 0.8, 0]])
DIRS = np.array([[0, 0, 1.0], [0, 0, -1.0], [0, 0, -1.0]])


def _mirror(sym, nphi, ntheta, axes, eps_out):
    from mnpbem_amd import bem as B
    lo, hi = SEGMENT[sym]
    mesh = shapes.ellipsoid_segment(nphi, ntheta, axes, lo, hi)
    op = B.bemoptions(sim="ret", sym=sym)
    p = B.comparticlemirror([B.epsconst(eps_out), B.epsdrude(*DRUDE)], [mesh], [2, 1], 1, op)
    return B, p, op, mesh


@pytest.mark.parametrize("sym", ["x", "y", "xy"])
def test_mirror_fp64_matches_oracle(sym, pinfty):
    from mnpbem_amd import device as dev
    nphi = 4 if sym == "xy" else 8
    B, p, op, mesh = _mirror(sym, nphi, 10, (10.0, 12.0, 20.0), 1.33)
    J = len(p.symtable)
    print(f"[mirror] sym {sym}: part {p.n} faces, full {p.full().n}", flush=True)
    enei = 610.0
    bem = B.bemsolver(p, op, precision="fp64")
    assert isinstance(bem, B.bemretmirror)
    exc = B.planewave(POL, DIRS, op)
    assert isinstance(exc, B.planewaveretmirror)
    pot = exc(p, enei)
    sig = bem.solve(pot)
    V, F = M.full_mesh(*mesh, sym)
    assert np.array_equal(F, p.full().faces) and np.allclose(V, p.full().verts, rtol=0, atol=0)
    g = O.GreenRet(V, F)
    e1 = O.eps_drude(enei, *DRUDE)
    mb = M.BemRetMirror(g, p.n, sym).initmat(enei, e1, 1.33)
    sols = []
    for i, val in enumerate(pot.val):
        pol2 = np.zeros((2, 3))
        pol2[:, i] = 1.0
        a2, a2p = O.planewave_potential(g.pos[:p.n], g.nvec[:p.n], pol2, [[0, 0, 1.0], [0, 0, -1.0]], enei, 1.33)
        assert np.allclose(dev.to_host(val.a2), a2, rtol=1e-13, atol=1e-15)
        sx = [M.symindex(sym, k) for k in M.PLANEWAVE_KEYS[sym][i]]
        assert [p.symindex(row) for row in val.symval] == sx
        want = mb.solve({"a2": a2, "a2p": a2p}, sx)
        sols.append(want)
        for key in ("sig1", "sig2", "h1", "h2"):
            got = dev.to_host(sig.val[i][key])
            err = np.linalg.norm(got - want[key]) / np.linalg.norm(want[key])
            print(f"  excitation {i} {key}: rel err {err:.2e}", flush=True)
            assert err < 1e-10
    # expanded to the full particle: cross sections against the oracle's full dense solve
    full = exc.full(sig)
    assert full["sig2"].shape == (J * p.n, len(POL))
    exp = M.full_solution(sols, M.planewave_symvals(sym), POL, DIRS)
    assert np.allclose(dev.to_host(full["h2"]), exp["h2"], rtol=1e-9, atol=1e-12 * np.abs(exp["h2"]).max())
    fa2, fa2p = O.planewave_potential(g.pos, g.nvec, POL, DIRS, enei, 1.33)
    dense = O.BemRet(g).initmat(enei, e1, 1.33).solve({"a2": fa2, "a2p": fa2p})
    ext, sca, _ = exc.cross_sections(sig)
    ext_d = O.extinction(g.pos, g.area, dense, POL, DIRS, enei, 1.33)
    sca_d = O.scattering(g.pos, g.area, dense, enei, 1.33, *pinfty)
    print(f"  ext {ext} vs full dense {ext_d}", flush=True)
    assert np.allclose(ext, ext_d, rtol=1e-9) and np.allclose(sca, sca_d, rtol=1e-9)


@pytest.mark.parametrize("precision", ["mixed", "mixed_bf16"])
def test_mirror_mixed_matches_fp64_and_full_solver(precision):
    """xy symmetry, 1,280-face part of a 5,120-face ellipsoid: the mixed
    variants reproduce the FP64 mirror solver, which equals the dense FP64
    solver on the full particle to the equivariance of the refined Green
    function (about 1e-11)."""
    B, p, op, _ = _mirror("xy", 16, 41, (10.0, 10.0, 40.0), 1.77)
    print(f"[mirror {precision}] part {p.n} faces, full {p.full().n}", flush=True)
    exc = B.planewave(POL, DIRS, op)
    b64 = B.bemsolver(p, op, precision="fp64")
    bmx = B.bemsolver(p, op, precision=precision)
    full_op = B.bemoptions(sim="ret")
    pf = B.comparticle(p.eps, [(p.full().verts, p.full().faces)], [2, 1], 1, full_op)
    bfull = B.bemsolver(pf, full_op, precision="fp64")
    excf = B.planewave(POL, DIRS, full_op)
    for enei in (550.0, 800.0):
        c64 = np.array(exc.cross_sections(b64.solve(exc(p, enei))))
        cmx = np.array(exc.cross_sections(bmx.solve(exc(p, enei))))
        it, rr = bmx.stats()
        cf = np.array(excf.cross_sections(bfull.solve(excf(pf, enei))))
        print(f"  enei {enei}: sweeps {it}, relres {rr:.2e}; ext mirror {c64[0]} full {cf[0]}", flush=True)
        assert rr < 1e-12 and it <= 6
        assert np.allclose(cmx, c64, rtol=1e-9)
        assert np.allclose(c64, cf, rtol=1e-8)


def test_mirror_argument_errors():
    import ctypes as C
    from mnpbem_amd import _lib
    B, p, op, _ = _mirror("xy", 4, 6, (10.0, 12.0, 20.0), 1.0)
    lib = _lib.lib()
    h = C.c_void_p()
    bad = np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, 1, -1], [1, 1, -1, -1]], dtype=np.int32)
    rc = lib.mnp_bemretmirror_create(p.green.handle, _lib.MNP_SOLVER_FP64, 4, bad.ctypes.data_as(C.POINTER(C.c_int)),
                                     C.byref(h))
    assert rc == _lib.MNP_ERR_ARG and b"orthogonal" in lib.mnp_last_error()
    tab = np.ascontiguousarray(p.symtable, dtype=np.int32)
    rc = lib.mnp_bemretmirror_create(p.green.handle, _lib.MNP_SOLVER_FP64, 3, tab.ctypes.data_as(C.POINTER(C.c_int)),
                                     C.byref(h))
    assert rc == _lib.MNP_ERR_ARG
    bem = B.bemsolver(p, op)
    bem(600.0)
    rc = lib.mnp_bemretmirror_solve(bem.handle, 0, 4, 0, 1, *([None] * 12))
    assert rc == _lib.MNP_ERR_ARG and b"symmetry index" in lib.mnp_last_error()
