"""GPU parity of the iterative quasistatic solver (GMRES on -(Lambda + F) with
F an H-matrix at k = 0) through the C ABI.

Checks: the GMRES solution against the direct solution of the same
H-matrix system (the oracle's H-matrix of F: refined full blocks plus the
oracle's ACA of the static collocation kernel), the dense oracle and dense
GPU solvers to the ACA accuracy, FP32 H-matrix storage, the default options
and the preconditioner's effect on the iteration count.
"""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import hmat_oracle as HO
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)


def _particle(mesh, eps_out, **op):
    from mnpbem_amd import bem as B
    opt = B.bemoptions(sim="stat", **op)
    return B, B.comparticle([B.epsconst(eps_out), B.epsdrude(*DRUDE)], [mesh], [2, 1], 1, opt), opt


def test_bemstatiter_equals_direct_solve_of_hmatrix_system():
    """GMRES at tol 1e-11 on the FP64 H-matrix of F reproduces the direct
    solution -(Lambda + F_H)^{-1} phi' (oracle), and the dense quasistatic
    solution to the ACA accuracy."""
    from mnpbem_amd import device as dev
    from mnpbem_amd.hmatrix import admissibility, clustertree
    v, f = shapes.ellipsoid(3, (10.0, 10.0, 25.0))
    print(f"[bemstatiter] ellipsoid level 3 n={len(f)}", flush=True)
    enei, eps_out = 620.0, 1.77
    pol = np.array([[1.0, 0, 0], [0, 0, 1.0]])
    B, p, op = _particle((v, f), eps_out, iter=dict(tol=1e-11, htol=1e-6, cleaf=32))
    it = B.bemsolver(p, op)
    assert isinstance(it, B.bemstatiter)
    exc = B.planewave(pol, op=op)
    sig = it.solve(exc(p, enei))
    st = it.stats()
    print(f"  bemstatiter: {st}", flush=True)
    assert st["flag"] == 0 and st["iter"] < 100 and st["true_relres"] < 1e-10
    assert 0 < st["compression"] < 1
    got = dev.to_host(sig.sig)
    g = O.GreenRet(v, f)
    tree = clustertree(g.pos, 32)
    _, low = admissibility(tree)
    FH = HO.hmat_dense(g, tree, low, 0.0, "F", 1e-6, 100)
    lam = 2 * np.pi * (O.eps_drude(enei, *DRUDE) + eps_out) / (O.eps_drude(enei, *DRUDE) - eps_out)
    phip = O.planewavestat_potential(g.nvec, pol)
    want = -np.linalg.solve(lam * np.eye(g.n) + FH, phip)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"  rel diff to the direct H-matrix solve {err:.2e}", flush=True)
    assert err < 1e-8
    dense = O.BemStat(g).initmat(enei, O.eps_drude(enei, *DRUDE), eps_out).solve(phip)
    cs = exc.cross_sections(sig)
    cs_h = O.stat_cross_sections(g.pos, g.area, want, pol, enei, eps_out)
    cs_d = O.stat_cross_sections(g.pos, g.area, dense, pol, enei, eps_out)
    print(f"  ext {cs[0]}; direct H-matrix {cs_h[0]}; dense {cs_d[0]}", flush=True)
    for a, b in zip(cs, cs_h):
        assert np.allclose(a, b, rtol=1e-8)
    # the H-matrix system differs from the dense one by the ACA error
    for a, b in zip(cs, cs_d):
        assert np.all(np.abs(a - b) <= 1e-4 * np.abs(b).max())


def test_bemstatiter_level4_fp32_storage_and_dense():
    """5,120 faces: GMRES reaches its tolerance with FP64 and FP32 H-matrix
    storage (whose products round x to FP32, unit roundoff 6e-8, which sets
    the floor of the true residual); both agree with the dense GPU solver to
    the H-matrix accuracy, and several wavelengths reuse the one fill of F."""
    from mnpbem_amd import device as dev
    mesh = shapes.ellipsoid(4, (10.0, 10.0, 40.0))
    print(f"[bemstatiter] ellipsoid level 4 n={len(mesh[1])}", flush=True)
    eps_out = 1.77
    B, p, op = _particle(mesh, eps_out)
    exc = B.planewave([[1.0, 0, 0], [0, 0, 1.0]], op=op)
    bd = B.bemsolver(p, op, precision="fp64")
    its = {prec: B.bemstatiter(p, dict(op, iter=dict(tol=tol, htol=1e-6, cleaf=32, hprecision=prec)))
           for prec, tol in (("fp64", 1e-9), ("fp32", 1e-7))}
    for enei in (550.0, 800.0):
        pot = exc(p, enei)
        sd = bd.solve(pot)
        cs_d = np.array(exc.cross_sections(sd))
        sol, cs = {}, {}
        for prec, it in its.items():
            sol[prec] = it.solve(pot)
            st = it.stats()
            print(f"  enei {enei} {prec}: {st}", flush=True)
            assert st["flag"] == 0 and st["true_relres"] < 2 * it.iter["tol"]
            cs[prec] = np.array(exc.cross_sections(sol[prec]))
        a, b, d = (dev.to_host(x.sig) for x in (sol["fp32"], sol["fp64"], sd))
        e32 = np.linalg.norm(a - b) / np.linalg.norm(b)
        ed = np.linalg.norm(b - d) / np.linalg.norm(d)
        print(f"  sig: fp32 vs fp64 storage {e32:.2e}; H-matrix vs dense {ed:.2e}; "
              f"ext {cs['fp64'][0]} dense {cs_d[0]}", flush=True)
        assert e32 < 1e-5
        assert np.allclose(cs["fp32"], cs["fp64"], rtol=2e-5)
        scale = np.abs(cs_d).max(axis=1, keepdims=True)
        assert np.all(np.abs(cs["fp64"] - cs_d) <= 2e-3 * scale)


def _block_jacobi(A, tree, spb=32):
    """Inverses of A on the preconditioner's blocks: each leaf cluster cut into
    near-equal runs of at most spb consecutive cluster indices."""
    blocks = []
    for c in range(len(tree.son)):
        if tree.son[c][0] >= 0:
            continue
        a, m = tree.cbeg[c], tree.cend[c] - tree.cbeg[c]
        nch = -(-m // spb)
        for q in range(nch):
            idx = tree.perm[a + q * m // nch:a + (q + 1) * m // nch]
            blocks.append((idx, np.linalg.inv(A[np.ix_(idx, idx)])))

    def apply(v):
        z = np.empty_like(v)
        for idx, Mi in blocks:
            z[idx] = Mi @ v[idx]
        return z
    return apply


def test_bemstatiter_default_options_and_preconditioner():
    """bemiter.options defaults (cleaf 200, htol 1e-6, tol 1e-6), with and
    without the block preconditioner: the GPU's GMRES stops at the first
    iteration whose residual reaches tol in the oracle's GMRES history on the
    same H-matrix system and preconditioner, with the same residual.  (For
    this second-kind operator the block preconditioner does not reduce the
    iteration count; the history pins its build and apply.)"""
    from mnpbem_amd.hmatrix import admissibility
    mesh = shapes.ellipsoid(3, (10.0, 10.0, 25.0))
    enei, eps_out = 540.0, 1.0
    B, p, op = _particle(mesh, eps_out)
    exc = B.planewave([[0, 0, 1.0]], op=op)
    pot = exc(p, enei)
    g = O.GreenRet(*mesh)
    e1 = O.eps_drude(enei, *DRUDE)
    lam = 2 * np.pi * (e1 + eps_out) / (e1 - eps_out)
    phip = O.planewavestat_potential(g.nvec, np.array([[0, 0, 1.0]]))[:, 0]
    for pre in ("hmat", None):
        it = B.bemstatiter(p, dict(op, iter=dict(maxit=300, precond=pre)))
        it.solve(pot)
        st = it.stats()
        _, low = admissibility(it.tree)
        A = -(lam * np.eye(g.n) + HO.hmat_dense(g, it.tree, low, 0.0, "F", 1e-6, 100))
        res = HO.gmres_relres(A, phip, _block_jacobi(A, it.tree) if pre else None, maxit=st["iter"])
        print(f"  precond {pre}: {st}; oracle history {np.array2string(res, precision=3)}", flush=True)
        assert st["flag"] == 0 and st["true_relres"] < 1e-5
        k = st["iter"]
        assert np.isclose(st["relres"], res[k], rtol=1e-3)
        assert res[k] <= 1e-6 * (1 + 1e-3) and res[k - 1] > 1e-6 * (1 - 1e-3)


def test_bemstatiter_argument_errors():
    from mnpbem_amd import _lib
    from mnpbem_amd import device as dev
    B, p, op = _particle(shapes.icosphere(2, 10.0), 1.0, iter=dict(cleaf=32))
    it = B.bemstatiter(p, op)
    lib = _lib.lib()
    buf = dev.colmajor((p.n, 9))
    x = dev.ptr(dev.base(buf))
    # solve before initmat, then too many right-hand sides
    rc = lib.mnp_bemstatiter_solve(it.handle, 1, x, x)
    assert rc == _lib.MNP_ERR_STATE and b"initmat" in lib.mnp_last_error()
    it(600.0)
    rc = lib.mnp_bemstatiter_solve(it.handle, 9, x, x)
    assert rc == _lib.MNP_ERR_ARG and b"npol" in lib.mnp_last_error()
    rc = lib.mnp_bemstatiter_set_options(it.handle, 0.0, 10, 1)
    assert rc == _lib.MNP_ERR_ARG
