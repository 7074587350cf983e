"""Pins for the H-matrix oracle: tree invariants, block partition, ACA accuracy."""
import numpy as np

from mnpbem_amd import shapes
from oracle import hmat_oracle as H
from oracle import mnpbem_oracle as O


def _geo(level=3, axes=(10.0, 10.0, 10.0)):
    v, f = shapes.ellipsoid(level, axes)
    pos, _, _, nvec, area = O.norm_flat(v, f)
    return pos, nvec, area


def test_cluster_tree_invariants():
    pos, _, _ = _geo(3)
    t = H.cluster_tree(pos, cleaf=32)
    assert np.array_equal(np.sort(t["perm"]), np.arange(len(pos)))
    son = t["son"]
    for i in range(len(son)):
        if son[i, 0] < 0:
            assert t["cend"][i] - t["cbeg"][i] <= 32
        else:
            a, b = son[i]
            assert t["cbeg"][a] == t["cbeg"][i] and t["cend"][a] == t["cbeg"][b]
            assert t["cend"][b] == t["cend"][i]
        p = pos[t["perm"][t["cbeg"][i]:t["cend"][i]]]
        lo, hi = p.min(0), p.max(0)
        assert np.allclose(t["mid"][i], 0.5 * (lo + hi))
        assert np.isclose(t["rad"][i], 0.5 * np.linalg.norm(hi - lo))
        assert np.all(np.linalg.norm(p - t["mid"][i], axis=1) <= t["rad"][i] + 1e-12)


def test_block_tree_partitions_the_matrix():
    pos, _, _ = _geo(3, (10.0, 10.0, 25.0))
    t = H.cluster_tree(pos, cleaf=32)
    full, low = H.block_tree(t, 2.5)
    n = len(pos)
    cover = np.zeros((n, n), dtype=np.int32)
    for i1, i2 in full + low:
        cover[t["cbeg"][i1]:t["cend"][i1], t["cbeg"][i2]:t["cend"][i2]] += 1
    assert np.all(cover == 1)
    son = t["son"]
    for i1, i2 in full:
        assert son[i1, 0] < 0 and son[i2, 0] < 0
    for i1, i2 in low:
        d = np.linalg.norm(t["mid"][i1] - t["mid"][i2])
        assert 2.5 * min(t["rad"][i1], t["rad"][i2]) < d
    assert len(low) > 0


def test_aca_exact_low_rank_and_zero_blocks():
    rng = np.random.default_rng(0)
    u = rng.normal(size=(40, 1)) + 1j * rng.normal(size=(40, 1))
    v = rng.normal(size=(30, 1)) + 1j * rng.normal(size=(30, 1))
    M = u @ v.T
    L, R = H.aca(lambda r: M[r], lambda c: M[:, c], 40, 30, tol=1e-6)
    assert L.shape[1] <= 2 and np.allclose(L @ R.T, M, atol=1e-12 * np.abs(M).max())
    Z = np.zeros((20, 25), complex)
    L, R = H.aca(lambda r: Z[r], lambda c: Z[:, c], 20, 25)
    assert L.shape[1] == 1 and np.all(L @ R.T == 0)


def test_aca_approximates_admissible_green_blocks():
    pos, nvec, area = _geo(3, (10.0, 10.0, 25.0))
    t = H.cluster_tree(pos, cleaf=32)
    _, low = H.block_tree(t, 2.5)
    pp, nn, aa = pos[t["perm"]], nvec[t["perm"]], area[t["perm"]]
    k = 0.05 + 0.01j
    for i1, i2 in low[:: max(1, len(low) // 12)]:
        rows = np.arange(t["cbeg"][i1], t["cend"][i1])
        cols = np.arange(t["cbeg"][i2], t["cend"][i2])
        for key in ("G", "F"):
            M = H.kernel_block(pp, nn, aa, rows, cols, k, key)
            L, R = H.aca(lambda r: M[r], lambda c: M[:, c], len(rows), len(cols), tol=1e-6)
            err = np.linalg.norm(L @ R.T - M) / np.linalg.norm(M)
            assert err < 1e-5, (key, err, L.shape[1])
            assert L.shape[1] < min(len(rows), len(cols))


def test_host_tree_equals_oracle():
    """The product's iterative cluster/block tree reproduces the oracle's recursion."""
    from mnpbem_amd.hmatrix import admissibility, clustertree
    for level, axes, cleaf in ((3, (10.0, 10.0, 25.0), 32), (4, (10.0, 10.0, 10.0), 16)):
        pos, _, _ = _geo(level, axes)
        ref = H.cluster_tree(pos, cleaf)
        t = clustertree(pos, cleaf)
        for key in ("son", "cbeg", "cend", "perm"):
            assert np.array_equal(getattr(t, key), ref[key]), key
        assert np.allclose(t.mid, ref["mid"]) and np.allclose(t.rad, ref["rad"])
        full, low = admissibility(t)
        rf, rl = H.block_tree(ref)
        assert sorted(full) == sorted(rf) and sorted(low) == sorted(rl)


def test_hmat_dense_blocks():
    """hmat_dense: refined values in full blocks, ACA of the collocation kernel
    (no refinement, no 2 pi) in admissible blocks."""
    v, f = shapes.ellipsoid(2, (10.0, 10.0, 25.0))
    g = O.GreenRet(v, f)
    t = H.cluster_tree(g.pos, cleaf=16)
    full, low = H.block_tree(t, 2.5)
    k = 0.1 + 0.02j
    A = H.hmat_dense(g, t, low, k, "H1", htol=1e-8)
    D = g.eval(k, ("H1",))["H1"]
    perm, cb, ce = t["perm"], t["cbeg"], t["cend"]
    for i1, i2 in full:
        r, c = perm[cb[i1]:ce[i1]], perm[cb[i2]:ce[i2]]
        assert np.array_equal(A[np.ix_(r, c)], D[np.ix_(r, c)])
    assert low
    for i1, i2 in low:
        r, c = perm[cb[i1]:ce[i1]], perm[cb[i2]:ce[i2]]
        M = H.kernel_block(g.pos, g.nvec, g.area, r, c, k, "F")
        assert np.linalg.norm(A[np.ix_(r, c)] - M) <= 1e-6 * np.linalg.norm(M)


def test_gmres_relres_matches_brute_force_krylov_least_squares():
    """The GMRES residual history equals min |b - A M y| over the Krylov space
    spanned explicitly by [b, (AM) b, ..., (AM)^(k-1) b] (QR of that basis,
    small k), is non-increasing, and reaches ~0 at k = n."""
    rng = np.random.default_rng(5)
    n = 12
    A = np.eye(n) * 3 + rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    Mi = np.diag(1 + rng.random(n)) + 0.1 * rng.normal(size=(n, n))
    for mfun, AM in ((None, A), (lambda v: Mi @ v, A @ Mi)):
        res = H.gmres_relres(A, b, mfun, maxit=n)
        K = np.empty((n, 0), dtype=complex)
        v = b.astype(complex)
        for k in range(1, 7):
            K = np.column_stack([K, v])
            v = AM @ v
            Q, _ = np.linalg.qr(AM @ K)
            r = b - Q @ (Q.conj().T @ b)
            assert np.isclose(res[k], np.linalg.norm(r) / np.linalg.norm(b), rtol=1e-8)
        assert np.all(np.diff(res) <= 1e-12) and res[n] < 1e-10
