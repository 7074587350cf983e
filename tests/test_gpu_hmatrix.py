"""GPU parity of the H-matrices (full blocks with refinement + ACA blocks) and
of the iterative solver (GMRES on the H-matrix operator) through the C ABI.

Checks: ACA ranks against the oracle's ACA, H x against the oracle's dense
matrix whose admissible blocks hold the collocation kernel (the reference
fills those blocks without refinement), full-block-only H-matrices against
the dense refined matrix, and bemretiter solutions against the oracle's
exact operator and the dense GPU solver.
"""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import hmat_oracle as HO
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)
KEYS = ("G", "F", "H1", "H2")


@pytest.fixture(scope="module")
def mesh():
    v, f = shapes.ellipsoid(3, (10.0, 10.0, 25.0))
    print(f"[hmatrix] ellipsoid level 3 n={len(f)}", flush=True)
    return v, f


@pytest.fixture(scope="module")
def setup(mesh):
    from mnpbem_amd.greenfun import CompGreenRet
    from mnpbem_amd.hmatrix import clustertree
    v, f = mesh
    ref = O.GreenRet(v, f)
    gpu = CompGreenRet(v, f)
    tree = clustertree(ref.pos, 32)
    return ref, gpu, tree


def _target(ref, tree, hm, k, key):
    """Oracle dense matrix, admissible blocks replaced by the collocation kernel."""
    dense = ref.eval(k, (key,))[key].copy()
    perm = tree.perm
    kk = "G" if key == "G" else "F"
    for r0, r1, c0, c1 in hm.low:
        rows, cols = perm[r0:r1], perm[c0:c1]
        dense[np.ix_(rows, cols)] = HO.kernel_block(ref.pos, ref.nvec, ref.area, rows, cols, k, kk)
    return dense


@pytest.mark.parametrize("k", [0.08 + 0.0j, 0.11 + 0.02j])
def test_aca_ranks_match_oracle(setup, k):
    from mnpbem_amd.hmatrix import hmatrix
    ref, gpu, tree = setup
    hm = hmatrix(gpu, tree, htol=1e-6, kmax=100)
    perm = tree.perm
    for key in ("G", "F"):
        hm.fill(k, key)
        got = hm.ranks()
        want = np.empty_like(got)
        for b, (r0, r1, c0, c1) in enumerate(hm.low):
            rows, cols = perm[r0:r1], perm[c0:c1]
            M = HO.kernel_block(ref.pos, ref.nvec, ref.area, rows, cols, k, key)
            L, _ = HO.aca(lambda r: M[r], lambda c: M[:, c], len(rows), len(cols), tol=1e-6, kmax=100)
            want[b] = L.shape[1]
        diff = np.abs(got.astype(int) - want)
        print(f"  {key}: {len(got)} blocks, ranks {got.min()}..{got.max()}, "
              f"{np.count_nonzero(diff)} differ (max {diff.max()})", flush=True)
        # FP64 on both sides; the pivot search may tie-break differently
        # only on entries equal to the last bit.
        assert np.count_nonzero(diff) <= max(1, len(got) // 50) and diff.max() <= 1


@pytest.mark.parametrize("key", KEYS)
def test_matvec_matches_dense(setup, key):
    import torch
    from mnpbem_amd import device as dev
    from mnpbem_amd.hmatrix import hmatrix
    ref, gpu, tree = setup
    k = 0.11 + 0.02j
    hm = hmatrix(gpu, tree, htol=1e-6, kmax=100).fill(k, key)
    A = _target(ref, tree, hm, k, key)
    rng = np.random.default_rng(7)
    x = rng.normal(size=(gpu.n, 3)) + 1j * rng.normal(size=(gpu.n, 3))
    y = dev.to_host(hm.matvec(dev.to_device_colmajor(x)))
    want = A @ x
    err = np.linalg.norm(y - want, axis=0) / np.linalg.norm(want, axis=0)
    st = hm.stats()
    print(f"  {key}: rel err {err}, compression {st['compression']:.3f}, max rank {st['max_rank']}",
          flush=True)
    # ACA at htol 1e-6 and complex FP32 storage
    assert err.max() < 2e-5
    assert st["compression"] < 1.0


def test_full_blocks_only_equal_dense(setup):
    """With one cluster (cleaf >= n) the H-matrix is the refined dense matrix in FP32."""
    from mnpbem_amd import device as dev
    from mnpbem_amd.hmatrix import clustertree, hmatrix
    ref, gpu, _ = setup
    tree = clustertree(ref.pos, gpu.n)
    k = 0.09 + 0.01j
    rng = np.random.default_rng(3)
    x = rng.normal(size=(gpu.n, 2)) + 1j * rng.normal(size=(gpu.n, 2))
    for key in KEYS:
        hm = hmatrix(gpu, tree).fill(k, key)
        assert len(hm.low) == 0
        want = ref.eval(k, (key,))[key] @ x
        y = dev.to_host(hm.matvec(dev.to_device_colmajor(x)))
        err = np.linalg.norm(y - want) / np.linalg.norm(want)
        print(f"  {key}: rel err {err:.2e}", flush=True)
        assert err < 2e-6


def test_hmat_argument_errors(setup):
    import ctypes as C
    from mnpbem_amd import _lib
    from mnpbem_amd.hmatrix import admissibility, block_ranges
    ref, gpu, tree = setup
    full, low = admissibility(tree)
    fr = np.ascontiguousarray(block_ranges(tree, full).T)
    lr = np.ascontiguousarray(block_ranges(tree, low).T)
    perm = np.ascontiguousarray(tree.perm, dtype=np.int64)
    h = C.c_void_p()
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    lib = _lib.lib()
    # a missing block: the blocks no longer tile the matrix
    rc = lib.mnp_hmat_create(gpu.handle, P(perm), len(fr) - 1, P(fr), len(lr), P(lr), 1e-6, 100, C.byref(h))
    assert rc == _lib.MNP_ERR_ARG and b"tile" in lib.mnp_last_error()
    rc = lib.mnp_hmat_create(gpu.handle, P(perm), len(fr), P(fr), len(lr), P(lr), 1e-6, 500, C.byref(h))
    assert rc == _lib.MNP_ERR_ARG
    bad = perm.copy()
    bad[0] = bad[1]
    rc = lib.mnp_hmat_create(gpu.handle, P(bad), len(fr), P(fr), len(lr), P(lr), 1e-6, 100, C.byref(h))
    assert rc == _lib.MNP_ERR_ARG and b"permutation" in lib.mnp_last_error()


# ------------------------------------------------------------ bemretiter
def _particle(v, f, eps_out, **op):
    from mnpbem_amd import bem as B
    opt = B.bemoptions(sim="ret", **op)
    return B, B.comparticle([B.epsconst(eps_out), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, opt), opt


def test_bemretiter_solves_exact_operator(mesh):
    from mnpbem_amd import device as dev
    v, f = mesh
    enei, eps_out = 700.0, 1.77
    pol = np.array([[1.0, 0, 0], [0, 0, 1.0]])
    dirs = np.array([[0, 0, 1.0], [1.0, 0, 0]])
    B, p, op = _particle(v, f, eps_out, iter=dict(tol=1e-8, htol=1e-7, cleaf=32))
    it = B.bemsolver(p, op)
    assert isinstance(it, B.bemretiter)
    exc = B.planewave(pol, dirs)
    sig = it.solve(exc(p, enei))
    st = it.stats()
    print(f"  bemretiter: {st}", flush=True)
    assert st["flag"] == 0 and st["iter"] < 100
    # residual with the oracle's dense (exact, refined) operator
    g = O.GreenRet(v, f)
    ob = O.BemRet(g).initmat(enei, O.eps_drude(enei, *DRUDE), eps_out)
    a2, a2p = O.planewave_potential(g.pos, g.nvec, pol, dirs, enei, eps_out)
    phi, a, alpha, De = ob.excitation({"a2": a2, "a2p": a2p})
    s = {k: dev.to_host(sig[k]) for k in ("sig1", "sig2", "h1", "h2")}
    rphi, ra, rDe, ralpha = ob.afun(s)
    num = sum(np.linalg.norm(x - y) ** 2 for x, y in ((rphi, phi), (ra, a), (rDe, De), (ralpha, alpha)))
    den = sum(np.linalg.norm(y) ** 2 for y in (phi, a, De, alpha))
    res = np.sqrt(num / den)
    print(f"  residual against the dense operator {res:.2e}", flush=True)
    # the H-matrix operator differs from the dense one only in admissible
    # blocks (ACA at htol, FP32 storage, no refinement there)
    assert res < 1e-4
    # cross sections against the oracle's dense solve
    want = ob.solve({"a2": a2, "a2p": a2p})
    ext = O.extinction(g.pos, g.area, s, pol, dirs, enei, eps_out)
    ext_ref = O.extinction(g.pos, g.area, want, pol, dirs, enei, eps_out)
    print(f"  ext {ext} vs dense {ext_ref}", flush=True)
    assert np.allclose(ext, ext_ref, rtol=1e-4)


def test_bemretiter_matches_dense_solver_level4():
    from mnpbem_amd import device as dev
    v, f = shapes.ellipsoid(4, (10.0, 10.0, 40.0))
    print(f"[bemretiter] ellipsoid level 4 n={len(f)}", flush=True)
    enei, eps_out = 650.0, 1.77
    pol, dirs = np.array([[0, 0, 1.0]]), np.array([[1.0, 0, 0]])
    B, p, op = _particle(v, f, eps_out)
    exc = B.planewave(pol, dirs)
    pot = exc(p, enei)
    dense = B.bemsolver(p, op, precision="fp64")
    it = B.bemretiter(p, dict(op, iter=dict(tol=1e-8, htol=1e-7, cleaf=32)))
    sd = dense.solve(pot)
    si = it.solve(pot)
    st = it.stats()
    print(f"  bemretiter: {st}", flush=True)
    assert st["flag"] == 0 and st["true_relres"] < 1e-6
    for key in ("sig2", "h2"):
        a, b = dev.to_host(si[key]), dev.to_host(sd[key])
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"  {key}: rel diff {err:.2e}", flush=True)
        assert err < 1e-3
    ext_i, ext_d = exc.ext(si), exc.ext(sd)
    sca_i, sca_d = exc.sca(si), exc.sca(sd)
    print(f"  ext {ext_i} vs {ext_d}; sca {sca_i} vs {sca_d}", flush=True)
    assert np.allclose(ext_i, ext_d, rtol=1e-4) and np.allclose(sca_i, sca_d, rtol=1e-4)


def test_bemretiter_preconditioner_reduces_iterations(mesh):
    v, f = mesh
    B, p, op = _particle(v, f, 1.77)
    exc = B.planewave([[1.0, 0, 0]], [[0, 0, 1.0]])
    pot = exc(p, 600.0)
    its = {}
    for pre in ("hmat", None):
        it = B.bemretiter(p, dict(op, iter=dict(tol=1e-6, maxit=300, precond=pre)))
        it.solve(pot)
        its[pre] = it.stats()
    print(f"  iterations {its}", flush=True)
    assert its["hmat"]["flag"] == 0
    assert its["hmat"]["iter"] < its[None]["iter"]


def test_bemretiter_rejects_large_leaves(mesh):
    from mnpbem_amd import _lib
    v, f = mesh
    B, p, op = _particle(v, f, 1.0)
    with pytest.raises(RuntimeError, match="cleaf"):
        B.bemretiter(p, dict(op, iter=dict(cleaf=64)))
