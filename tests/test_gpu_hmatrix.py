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
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_matvec_matches_dense(setup, key, precision):
    from mnpbem_amd import device as dev
    from mnpbem_amd.hmatrix import hmatrix
    ref, gpu, tree = setup
    k = 0.11 + 0.02j
    hm = hmatrix(gpu, tree, htol=1e-6, kmax=100, precision=precision).fill(k, key)
    A = _target(ref, tree, hm, k, key)
    rng = np.random.default_rng(7)
    x = rng.normal(size=(gpu.n, 3)) + 1j * rng.normal(size=(gpu.n, 3))
    y = dev.to_host(hm.matvec(dev.to_device_colmajor(x)))
    want = A @ x
    err = np.linalg.norm(y - want, axis=0) / np.linalg.norm(want, axis=0)
    st = hm.stats()
    print(f"  {key} {precision}: rel err {err}, compression {st['compression']:.3f}, "
          f"max rank {st['max_rank']}", flush=True)
    # ACA at htol 1e-6 (and complex FP32 storage)
    assert err.max() < 2e-5
    assert st["compression"] < 1.0


@pytest.mark.parametrize("key", KEYS)
def test_fp64_matvec_equals_oracle_hmatrix(setup, key):
    """FP64 storage: H x equals the oracle's H-matrix (refined full blocks +
    the oracle's ACA factors) times x, since both run the same ACA in FP64."""
    from mnpbem_amd import device as dev
    from mnpbem_amd.hmatrix import admissibility, hmatrix
    ref, gpu, tree = setup
    k = 0.07 + 0.015j
    hm = hmatrix(gpu, tree, htol=1e-6, kmax=100, precision="fp64").fill(k, key)
    _, low = admissibility(tree)
    A = HO.hmat_dense(ref, tree, low, k, key, htol=1e-6, kmax=100)
    rng = np.random.default_rng(11)
    x = rng.normal(size=(gpu.n, 2)) + 1j * rng.normal(size=(gpu.n, 2))
    y = dev.to_host(hm.matvec(dev.to_device_colmajor(x)))
    want = A @ x
    err = np.linalg.norm(y - want) / np.linalg.norm(want)
    print(f"  {key}: rel err vs oracle H-matrix {err:.2e}", flush=True)
    assert err < 1e-11


@pytest.mark.parametrize("key", ["G", "H1"])
def test_coarse_operator_equals_aggregated_oracle_hmatrix(setup, key):
    """FP64 storage: the coarse operator P^T H P summed from the stored blocks
    and factors equals the oracle's H-matrix (same ACA) summed over the same
    tiles, to 1e-11; the tiles partition the cluster indices in runs of at
    most 32."""
    from mnpbem_amd import device as dev
    from mnpbem_amd.hmatrix import admissibility, hmatrix
    ref, gpu, tree = setup
    k = 0.07 + 0.015j
    hm = hmatrix(gpu, tree, htol=1e-6, kmax=100, precision="fp64").fill(k, key)
    tb, Cm = hm.coarsen()
    Cg = dev.to_host(Cm)
    nt = len(tb) - 1
    assert tb[0] == 0 and tb[-1] == gpu.n and np.all(np.diff(tb) >= 1) and np.all(np.diff(tb) <= 32)
    _, low = admissibility(tree)
    A = HO.hmat_dense(ref, tree, low, k, key, htol=1e-6, kmax=100)
    P = np.zeros((gpu.n, nt))
    for t in range(nt):
        P[tree.perm[tb[t]:tb[t + 1]], t] = 1.0
    want = P.T @ A @ P
    err = np.abs(Cg - want).max() / np.abs(want).max()
    print(f"  {key}: {nt} tiles, coarse operator rel err {err:.2e}", flush=True)
    assert err < 1e-11


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
def test_full_blocks_only_equal_dense(setup, precision, tol):
    """With one cluster (cleaf >= n) the H-matrix is the refined dense matrix."""
    from mnpbem_amd import device as dev
    from mnpbem_amd.hmatrix import clustertree, hmatrix
    ref, gpu, _ = setup
    tree = clustertree(ref.pos, gpu.n)
    k = 0.09 + 0.01j
    rng = np.random.default_rng(3)
    x = rng.normal(size=(gpu.n, 2)) + 1j * rng.normal(size=(gpu.n, 2))
    for key in KEYS:
        hm = hmatrix(gpu, tree, precision=precision).fill(k, key)
        assert len(hm.low) == 0
        want = ref.eval(k, (key,))[key] @ x
        y = dev.to_host(hm.matvec(dev.to_device_colmajor(x)))
        err = np.linalg.norm(y - want) / np.linalg.norm(want)
        print(f"  {key} {precision}: rel err {err:.2e}", flush=True)
        assert err < tol


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
    F64 = _lib.MNP_FP64
    # a missing block: the blocks no longer tile the matrix
    rc = lib.mnp_hmat_create(gpu.handle, F64, P(perm), len(fr) - 1, P(fr), len(lr), P(lr), 1e-6, 100,
                             C.byref(h))
    assert rc == _lib.MNP_ERR_ARG and b"tile" in lib.mnp_last_error()
    rc = lib.mnp_hmat_create(gpu.handle, F64, P(perm), len(fr), P(fr), len(lr), P(lr), 1e-6, 500, C.byref(h))
    assert rc == _lib.MNP_ERR_ARG
    rc = lib.mnp_hmat_create(gpu.handle, 7, P(perm), len(fr), P(fr), len(lr), P(lr), 1e-6, 100, C.byref(h))
    assert rc == _lib.MNP_ERR_ARG and b"precision" in lib.mnp_last_error()
    bad = perm.copy()
    bad[0] = bad[1]
    rc = lib.mnp_hmat_create(gpu.handle, F64, P(bad), len(fr), P(fr), len(lr), P(lr), 1e-6, 100, C.byref(h))
    assert rc == _lib.MNP_ERR_ARG and b"permutation" in lib.mnp_last_error()


# ------------------------------------------------------------ bemretiter
def _particle(v, f, eps_out, **op):
    from mnpbem_amd import bem as B
    opt = B.bemoptions(sim="ret", **op)
    return B, B.comparticle([B.epsconst(eps_out), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, opt), opt


def test_bemretiter_equals_direct_solve_of_hmatrix_system(mesh):
    """GMRES at tol 1e-10 on FP64 H-matrices reproduces the direct solution of
    the same H-matrix system (oracle), and solves the dense refined system to
    the ACA accuracy."""
    from mnpbem_amd import device as dev
    from mnpbem_amd.hmatrix import admissibility, clustertree
    v, f = mesh
    enei, eps_out = 700.0, 1.77
    pol = np.array([[1.0, 0, 0], [0, 0, 1.0]])
    dirs = np.array([[0, 0, 1.0], [1.0, 0, 0]])
    B, p, op = _particle(v, f, eps_out, iter=dict(tol=1e-10, htol=1e-6, cleaf=32))
    it = B.bemsolver(p, op)
    assert isinstance(it, B.bemretiter)
    exc = B.planewave(pol, dirs)
    sig = it.solve(exc(p, enei))
    st = it.stats()
    print(f"  bemretiter: {st}", flush=True)
    assert st["flag"] == 0 and st["iter"] < 100 and st["true_relres"] < 1e-9
    s = {k: dev.to_host(sig[k]) for k in ("sig1", "sig2", "h1", "h2")}
    g = O.GreenRet(v, f)
    tree = clustertree(g.pos, 32)
    _, low = admissibility(tree)
    hm_eval = lambda k, keys: {key: HO.hmat_dense(g, tree, low, k, key, 1e-6, 100) for key in keys}
    e1 = O.eps_drude(enei, *DRUDE)
    a2, a2p = O.planewave_potential(g.pos, g.nvec, pol, dirs, enei, eps_out)
    hsol = O.BemRet(g, evalfun=hm_eval).initmat(enei, e1, eps_out).solve({"a2": a2, "a2p": a2p})
    for key in ("sig1", "sig2", "h1", "h2"):
        err = np.linalg.norm(s[key] - hsol[key]) / np.linalg.norm(hsol[key])
        print(f"  {key}: rel diff to the direct H-matrix solve {err:.2e}", flush=True)
        assert err < 1e-7
    ob = O.BemRet(g).initmat(enei, e1, eps_out)
    dsol = ob.solve({"a2": a2, "a2p": a2p})
    ext = O.extinction(g.pos, g.area, s, pol, dirs, enei, eps_out)
    ext_h = O.extinction(g.pos, g.area, hsol, pol, dirs, enei, eps_out)
    ext_d = O.extinction(g.pos, g.area, dsol, pol, dirs, enei, eps_out)
    print(f"  ext {ext}; direct H-matrix {ext_h}; dense {ext_d}", flush=True)
    assert np.all(np.abs(ext - ext_h) <= 1e-7 * np.abs(ext_h).max())
    # the H-matrix system differs from the dense one by the ACA error
    assert np.all(np.abs(ext - ext_d) <= 1e-4 * np.abs(ext_d).max())


def test_bemretiter_level4_fp32_storage_and_dense():
    """5,120 faces: GMRES reaches its tolerance with FP64 and FP32 H-matrix
    storage; the two agree to the FP32 storage rounding, and both agree
    with the dense solver to the accuracy of the H-matrix scheme (the
    reference fills admissible blocks without near-field refinement; on this
    elongated mesh 33,734 refined pairs fall into admissible blocks)."""
    from mnpbem_amd import device as dev
    v, f = shapes.ellipsoid(4, (10.0, 10.0, 40.0))
    print(f"[bemretiter] ellipsoid level 4 n={len(f)}", flush=True)
    enei, eps_out = 650.0, 1.77
    pol, dirs = np.array([[0, 0, 1.0], [1.0, 0, 0]]), np.array([[1.0, 0, 0], [0, 0, 1.0]])
    B, p, op = _particle(v, f, eps_out)
    exc = B.planewave(pol, dirs)
    pot = exc(p, enei)
    sd = B.bemsolver(p, op, precision="fp64").solve(pot)
    sol, cs = {}, {}
    # FP32 storage rounds x to FP32 inside each H-matvec, which bounds the
    # attainable residual near 1e-8
    for prec, tol in (("fp64", 1e-9), ("fp32", 1e-8)):
        it = B.bemretiter(p, dict(op, iter=dict(tol=tol, htol=1e-6, cleaf=32, hprecision=prec)))
        sol[prec] = it.solve(pot)
        st = it.stats()
        print(f"  {prec}: {st}", flush=True)
        assert st["flag"] == 0 and st["true_relres"] < 2 * tol
        cs[prec] = np.array([exc.ext(sol[prec]), exc.sca(sol[prec])])
    cs_d = np.array([exc.ext(sd), exc.sca(sd)])
    print(f"  ext/sca fp64 {cs['fp64'].ravel()} fp32 {cs['fp32'].ravel()} dense {cs_d.ravel()}", flush=True)
    for key in ("sig2", "h2"):
        a, b, d = (dev.to_host(x[key]) for x in (sol["fp32"], sol["fp64"], sd))
        e32 = np.linalg.norm(a - b) / np.linalg.norm(b)
        ed = np.linalg.norm(b - d) / np.linalg.norm(d)
        print(f"  {key}: fp32 vs fp64 storage {e32:.2e}; H-matrix vs dense {ed:.2e}", flush=True)
        assert e32 < 1e-5
    # FP32 storage (inputs rounded to FP32 in every product): residual ~1e-8
    # times the conditioning of this elongated mesh
    assert np.allclose(cs["fp32"], cs["fp64"], rtol=2e-5)
    # per quantity, relative to its largest value over the polarisations
    scale = np.abs(cs_d).max(axis=1, keepdims=True)
    assert np.all(np.abs(cs["fp64"] - cs_d) <= 2e-3 * scale)


@pytest.mark.parametrize("cleaf", [200, 32])
def test_bemretiter_preconditioners(mesh, cleaf):
    """bemiter.options defaults (cleaf 200, htol 1e-6, tol 1e-6) and cleaf 32:
    the block chain reduces the iteration count, the two-level preconditioner
    (precond 'hmat') more; all three reach the same solution."""
    from mnpbem_amd import device as dev
    v, f = mesh
    B, p, op = _particle(v, f, 1.77)
    exc = B.planewave([[1.0, 0, 0]], [[0, 0, 1.0]])
    pot = exc(p, 600.0)
    its, sol = {}, {}
    for pre in ("hmat", "block", None):
        it = B.bemretiter(p, dict(op, iter=dict(maxit=300, precond=pre, cleaf=cleaf)))
        sol[pre] = dev.to_host(it.solve(pot)["sig2"])
        its[pre] = it.stats()
    print(f"  cleaf {cleaf}: iterations { {k: v['iter'] for k, v in its.items()} }", flush=True)
    for pre in its:
        assert its[pre]["flag"] == 0 and its[pre]["true_relres"] < 1e-5
        assert np.linalg.norm(sol[pre] - sol["hmat"]) <= 1e-4 * np.linalg.norm(sol["hmat"])
    assert its["hmat"]["iter"] < its["block"]["iter"] < its[None]["iter"]
