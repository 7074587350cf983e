"""CPU restatement of MNPBEM's H-matrix construction (TEST INFRASTRUCTURE ONLY).

- cluster tree by bisection: Greenfun/hmatrices/@clustertree/private/init.m:14-80,
  private/bisection.m:1-96 (longest box axis, split at the box midpoint,
  first son = positions strictly below it, sons appended depth-first);
- block tree: @clustertree/admissibility.m:14-74 (leaf x leaf -> full block,
  admissible if 2.5 min(rad1, rad2) < |mid1 - mid2|, else recurse into sons);
- partially pivoted ACA of a block of the retarded kernel: mex/hlib/aca.cpp:158-218
  with the collocation kernel of mex/acagreen/acagreen.cpp:79-127.

Indices are 0-based; cluster ranges are [begin, end) in cluster order.
"""
import numpy as np


def cluster_tree(pos, cleaf=32):
    """Returns dict(son (nnode, 2) with -1 for leaves, cbeg, cend (cluster
    ranges), mid, rad, perm (cluster index -> particle index))."""
    nodes = []

    def sphere(ind):
        lo, hi = pos[ind].min(axis=0), pos[ind].max(axis=0)
        return 0.5 * (lo + hi), 0.5 * np.linalg.norm(hi - lo)

    mid, rad = sphere(np.arange(len(pos)))
    nodes.append(dict(cbeg=0, cend=len(pos), mid=mid, rad=rad, son=[-1, -1], ind=None))

    def bisect(ic, ind):
        lo, hi = pos[ind].min(axis=0), pos[ind].max(axis=0)
        k = int(np.argmax(hi - lo))
        m = lo[k] + 0.5 * (hi[k] - lo[k])
        ind1 = ind[pos[ind, k] < m]
        ind2 = np.setdiff1d(ind, ind1)
        c0 = nodes[ic]["cbeg"]
        s = len(nodes)
        nodes[ic]["son"] = [s, s + 1]
        for sub, cb in ((ind1, c0), (ind2, c0 + len(ind1))):
            mm, rr = sphere(sub)
            nodes.append(dict(cbeg=cb, cend=cb + len(sub), mid=mm, rad=rr, son=[-1, -1], ind=None))
        for son, sub in ((s, ind1), (s + 1, ind2)):
            if len(sub) > cleaf:
                bisect(son, sub)
            else:
                nodes[son]["ind"] = sub

    bisect(0, np.arange(len(pos)))
    perm = np.zeros(len(pos), dtype=np.int64)
    for nd in nodes:
        if nd["ind"] is not None:
            perm[nd["cbeg"]:nd["cend"]] = nd["ind"]
    return dict(son=np.array([nd["son"] for nd in nodes]),
                cbeg=np.array([nd["cbeg"] for nd in nodes]),
                cend=np.array([nd["cend"] for nd in nodes]),
                mid=np.array([nd["mid"] for nd in nodes]),
                rad=np.array([nd["rad"] for nd in nodes]), perm=perm)


def block_tree(tree, eta=2.5):
    """(full, low): lists of (row node, col node), full blocks are leaf pairs."""
    son = tree["son"]
    full, low = [], []

    def rec(i1, i2):
        leaf1, leaf2 = son[i1, 0] < 0, son[i2, 0] < 0
        if leaf1 and leaf2:
            full.append((i1, i2))
            return
        d = np.linalg.norm(tree["mid"][i1] - tree["mid"][i2])
        if eta * min(tree["rad"][i1], tree["rad"][i2]) < d:
            low.append((i1, i2))
            return
        for a in ([i1] if leaf1 else son[i1]):
            for b in ([i2] if leaf2 else son[i2]):
                rec(a, b)

    rec(0, 0)
    return full, low


def kernel_block(pos, nvec, area, rows, cols, k, key):
    """Collocation kernel of acagreen.cpp:79-127 (no refinement): G = e^{ikd}/d area_c,
    F = n_r.(p_r - p_c) (ik - 1/d) e^{ikd}/d^2 area_c."""
    x = pos[rows][:, None, :] - pos[cols][None, :, :]
    d = np.sqrt(np.sum(x * x, axis=2))
    fac = np.exp(1j * k * d)
    if key == "G":
        return fac / d * area[cols][None, :]
    inn = np.sum(x * nvec[rows][:, None, :], axis=2)
    return inn * (1j * k - 1.0 / d) * fac / d ** 2 * area[cols][None, :]


def aca(getrow, getcol, m, n, tol=1e-6, kmax=100, acatol=1e-10):
    """Partially pivoted ACA (aca.cpp:158-218).  Returns L (m, r), R (n, r) with block ~ L R^T."""
    kmax = min(kmax, m, n)
    A = np.zeros((m, kmax), complex)
    B = np.zeros((n, kmax), complex)
    rows = list(range(m))
    r, nsum = 0, 0.0
    izamax = lambda v: int(np.argmax(np.abs(v.real) + np.abs(v.imag)))
    k = 0
    while k < kmax:
        B[:, k] = getrow(r)
        if np.linalg.norm(B[:, k]) < acatol:
            break
        if k:
            B[:, k] -= B[:, :k] @ A[r, :k]
        c = izamax(B[:, k])
        B[:, k] *= 1.0 / B[c, k]
        A[:, k] = getcol(c)
        if k:
            A[:, k] -= A[:, :k] @ B[c, :k]
        rows.remove(r)
        work = np.zeros(m, complex)
        work[rows] = A[rows, k]
        r = izamax(work)
        nk = np.linalg.norm(A[:, k]) * np.linalg.norm(B[:, k])
        if nk < tol * nsum or not rows:
            break
        nsum = np.sqrt(nsum ** 2 + nk ** 2)
        k += 1
    rank = min(k + 1, kmax)
    return A[:, :rank], B[:, :rank]


def hmat_dense(green, tree, blocks_low, k, key, htol=1e-6, kmax=100):
    """The H-matrix of operator ``key`` as a dense (n, n) array in face order.

    Full blocks hold the refined dense values (green.eval, as @hmatrix/fillval.m
    fills leaf blocks through greenret eval2); each admissible block
    (row node, col node) in ``blocks_low`` holds the ACA approximant L R^T of
    the collocation kernel (aca.cpp:158-218, no refinement, no 2 pi term).
    """
    dense = green.eval(k, (key,))[key].copy()
    perm = tree["perm"] if isinstance(tree, dict) else tree.perm
    cbeg = tree["cbeg"] if isinstance(tree, dict) else tree.cbeg
    cend = tree["cend"] if isinstance(tree, dict) else tree.cend
    kk = "G" if key == "G" else "F"
    for i1, i2 in blocks_low:
        rows, cols = perm[cbeg[i1]:cend[i1]], perm[cbeg[i2]:cend[i2]]
        M = kernel_block(green.pos, green.nvec, green.area, rows, cols, k, kk)
        L, R = aca(lambda r: M[r], lambda c: M[:, c], len(rows), len(cols), tol=htol, kmax=kmax)
        dense[np.ix_(rows, cols)] = L @ R.T
    return dense


def gmres_relres(A, b, mfun=None, maxit=100):
    """Residual history of GMRES without restarts, right-preconditioned:
    res[k] = min over y of |b - A M y| / |b| with y in the k-dimensional
    Krylov space of A M and b (Arnoldi with twice-iterated classical
    Gram-Schmidt, least squares on the Hessenberg matrix).  ``mfun`` applies
    M (identity when None); b is one column.  The iterative solvers of
    BEM/@bemiter/solve.m:33-36 stop at the first k with res[k] <= tol; with
    a fixed M the flexible form used on the GPU spans the same space."""
    b = np.asarray(b, dtype=complex).ravel()
    m = (lambda v: v) if mfun is None else mfun
    nb = np.linalg.norm(b)
    V = [b / nb]
    H = np.zeros((maxit + 1, maxit), dtype=complex)
    e1 = np.zeros(maxit + 1, dtype=complex)
    e1[0] = nb
    res = [1.0]
    for k in range(maxit):
        w = A @ m(V[k])
        for _ in range(2):
            Q = np.array(V).T
            h = Q.conj().T @ w
            w = w - Q @ h
            H[:k + 1, k] += h
        H[k + 1, k] = np.linalg.norm(w)
        y = np.linalg.lstsq(H[:k + 2, :k + 1], e1[:k + 2], rcond=None)[0]
        res.append(np.linalg.norm(e1[:k + 2] - H[:k + 2, :k + 1] @ y) / nb)
        if H[k + 1, k].real < 1e-300:
            break
        V.append(w / H[k + 1, k].real)
    return np.array(res)
