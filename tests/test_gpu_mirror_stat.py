"""GPU parity of the quasistatic solver with mirror symmetry (bemstatmirror)
and its plane wave (planewavestatmirror) through the C ABI: surface charges
of each excitation against the oracle's restatement, the expanded charge and
cross sections against the oracle's bemstat on the full particle, the mixed
and split-bf16 variants against FP64 and the GPU bemstat on the full
particle."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import mirror_oracle as M
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)
SEGMENT = {"xy": (0.0, np.pi / 2), "y": (0.0, np.pi), "x": (-np.pi / 2, np.pi / 2)}
POL = np.array([[1.0, 0, 0], [0, 1.0, 0], [0.6, 0.8, 0]])


def _mirror(sym, mesh, eps_out):
    from mnpbem_amd import bem as B
    op = B.bemoptions(sim="stat", sym=sym)
    p = B.comparticlemirror([B.epsconst(eps_out), B.epsdrude(*DRUDE)], [mesh], [2, 1], 1, op)
    return B, p, op


@pytest.mark.parametrize("sym", ["x", "y", "xy"])
def test_stat_mirror_fp64_matches_oracle(sym):
    from mnpbem_amd import device as dev
    lo, hi = SEGMENT[sym]
    mesh = shapes.ellipsoid_segment(4 if sym == "xy" else 8, 10, (10.0, 12.0, 20.0), lo, hi)
    B, p, op = _mirror(sym, mesh, 1.33)
    print(f"[stat mirror] sym {sym}: part {p.n} faces, full {p.full().n}", flush=True)
    bem = B.bemsolver(p, op, precision="fp64")
    exc = B.planewave(POL, None, op)
    assert isinstance(bem, B.bemstatmirror) and isinstance(exc, B.planewavestatmirror)
    V, F = M.full_mesh(*mesh, sym)
    g = O.GreenRet(V, F)
    mb = M.BemStatMirror(g, p.n, sym)
    symvals = M.planewavestat_symvals(sym)
    for enei in (540.0, 700.0):
        pot = exc(p, enei)
        sig = bem.solve(pot)
        e1 = O.eps_drude(enei, *DRUDE)
        mb.initmat(enei, e1, 1.33)
        sols = []
        for i, val in enumerate(pot.val):
            p1 = np.zeros((1, 3))
            p1[0, i] = 1.0
            phip = O.planewavestat_potential(g.nvec[:p.n], p1)
            assert np.allclose(dev.to_host(val.phip), phip, rtol=1e-14, atol=1e-15)
            s = M.symindex(sym, M.PLANEWAVE_KEYS[sym][i][-1])
            assert p.symindex(val.symval[-1]) == s
            want = mb.solve(phip, s)
            sols.append(want)
            err = np.linalg.norm(dev.to_host(sig.val[i].sig) - want) / np.linalg.norm(want)
            print(f"  enei {enei} excitation {i}: rel err {err:.2e}", flush=True)
            assert err < 1e-10
        full = exc.full(sig)
        dense = O.BemStat(g).initmat(enei, e1, 1.33).solve(O.planewavestat_potential(g.nvec, POL))
        got = dev.to_host(full.sig)
        assert got.shape == (len(F), len(POL))
        assert np.allclose(got, M.full_stat(sols, symvals, POL), rtol=1e-12, atol=1e-13 * np.abs(got).max())
        assert np.linalg.norm(got - dense) <= 1e-9 * np.linalg.norm(dense)
        ext, sca, ab = exc.cross_sections(sig)
        ext_o, sca_o, ab_o = O.stat_cross_sections(g.pos, g.area, dense, POL, enei, 1.33)
        print(f"  ext {ext} vs full oracle {ext_o}", flush=True)
        assert np.allclose(ext, ext_o, rtol=1e-9) and np.allclose(sca, sca_o, rtol=1e-9)
        assert np.allclose(ab, ab_o, rtol=1e-9)


@pytest.mark.parametrize("precision", ["mixed", "mixed_bf16"])
def test_stat_mirror_mixed_matches_fp64_and_full_solver(precision):
    """xy symmetry, octant mesh: 1,152-face part of a 4,608-face ellipsoid."""
    mesh = shapes.ellipsoid_octants(24, (10.0, 10.0, 40.0), "xy")
    B, p, op = _mirror("xy", mesh, 1.77)
    print(f"[stat mirror {precision}] part {p.n} faces, full {p.full().n}", flush=True)
    exc = B.planewave(POL, None, op)
    b64 = B.bemsolver(p, op, precision="fp64")
    bmx = B.bemsolver(p, op, precision=precision)
    full_op = B.bemoptions(sim="stat")
    pf = B.comparticle(p.eps, [(p.full().verts, p.full().faces)], [2, 1], 1, full_op)
    bfull = B.bemsolver(pf, full_op, precision="fp64")
    excf = B.planewave(POL, None, full_op)
    for enei in (560.0, 800.0):
        c64 = np.array(exc.cross_sections(b64.solve(exc(p, enei))))
        cmx = np.array(exc.cross_sections(bmx.solve(exc(p, enei))))
        sweeps = bmx.val_stats
        cf = np.array(excf.cross_sections(bfull.solve(excf(pf, enei))))
        print(f"  enei {enei}: sweeps {sweeps}; ext mirror {c64[0]} full {cf[0]}", flush=True)
        assert all(rr < 1e-12 and it <= 6 for it, rr in sweeps)
        assert np.allclose(cmx, c64, rtol=1e-9)
        assert np.allclose(c64, cf, rtol=1e-9)


def test_stat_mirror_argument_errors():
    import ctypes as C
    from mnpbem_amd import _lib
    mesh = shapes.ellipsoid_segment(4, 6, (10.0, 12.0, 20.0), *SEGMENT["xy"])
    B, p, op = _mirror("xy", mesh, 1.0)
    lib = _lib.lib()
    h = C.c_void_p()
    bad = np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, 1, -1], [1, 1, -1, -1]], dtype=np.int32)
    rc = lib.mnp_bemstatmirror_create(p.green.handle, _lib.MNP_SOLVER_FP64, 4, bad.ctypes.data_as(C.POINTER(C.c_int)),
                                      C.byref(h))
    assert rc == _lib.MNP_ERR_ARG and b"orthogonal" in lib.mnp_last_error()
    bem = B.bemsolver(p, op)
    bem(600.0)
    rc = lib.mnp_bemstatmirror_solve(bem.handle, 4, 1, None, None)
    assert rc == _lib.MNP_ERR_ARG
    x = np.zeros(1)
    ptr = x.ctypes.data_as(C.c_void_p)
    rc = lib.mnp_bemstatmirror_solve(bem.handle, 4, 1, ptr, ptr)
    assert rc == _lib.MNP_ERR_ARG and b"symmetry index" in lib.mnp_last_error()
    with pytest.raises(ValueError):
        B.planewave(np.array([[0.0, 0.0, 1.0]]), None, op)
