"""Cluster tree and H-matrix block structure (host side of MNPBEM's hmatrices).

Mirrors Greenfun/hmatrices/@clustertree (bisection tree, private/init.m:14-80,
private/bisection.m) and @clustertree/admissibility.m + @hmatrix/private/init.m
(block tree).  The geometry-only structure is built here, as the reference
builds it in Matlab before calling its MEX kernels; the blocks are then filled
(full blocks with refinement, admissible blocks by ACA) and multiplied on the
GPU (mnp_hmat_* in include/mnpbem_hip.h).
"""
import numpy as np


class clustertree:
    """Bisection cluster tree over face centroids ``pos`` (n, 3).

    Attributes (0-based): ``son`` (nnode, 2), -1 for leaves; ``cbeg``/``cend``
    cluster-index ranges; ``mid``/``rad`` bounding spheres; ``perm`` cluster
    index -> face index; ``iperm`` face index -> cluster index.
    """

    def __init__(self, pos, cleaf=32):
        pos = np.asarray(pos, dtype=np.float64)
        n = len(pos)
        self.cleaf = int(cleaf)
        son, cbeg, cend, mid, rad = [[-1, -1]], [0], [n], [], []
        lo, hi = pos.min(axis=0), pos.max(axis=0)
        mid.append(0.5 * (lo + hi))
        rad.append(0.5 * np.linalg.norm(hi - lo))
        perm = np.empty(n, dtype=np.int64)
        # depth-first, first son before second son, sons numbered when their
        # parent is split (the node order of the reference's recursion)
        stack = [(0, np.arange(n))]
        while stack:
            ic, ind = stack.pop()
            p = pos[ind]
            lo, hi = p.min(axis=0), p.max(axis=0)
            k = int(np.argmax(hi - lo))
            m = lo[k] + 0.5 * (hi[k] - lo[k])
            below = p[:, k] < m
            parts = (ind[below], ind[~below])  # ind stays sorted: setdiff order
            s = len(son)
            son[ic] = [s, s + 1]
            c0 = cbeg[ic]
            for sub, cb in ((parts[0], c0), (parts[1], c0 + len(parts[0]))):
                q = pos[sub]
                a, b = q.min(axis=0), q.max(axis=0)
                son.append([-1, -1])
                cbeg.append(cb)
                cend.append(cb + len(sub))
                mid.append(0.5 * (a + b))
                rad.append(0.5 * np.linalg.norm(b - a))
            todo = []
            for node, sub in ((s, parts[0]), (s + 1, parts[1])):
                if len(sub) > self.cleaf:
                    todo.append((node, sub))
                else:
                    perm[cbeg[node]:cend[node]] = sub
            stack.extend(reversed(todo))
        self.son = np.array(son, dtype=np.int64)
        self.cbeg = np.array(cbeg, dtype=np.int64)
        self.cend = np.array(cend, dtype=np.int64)
        self.mid = np.array(mid)
        self.rad = np.array(rad)
        self.perm = perm
        self.iperm = np.empty(n, dtype=np.int64)
        self.iperm[perm] = np.arange(n)

    @property
    def leaves(self):
        return np.nonzero(self.son[:, 0] < 0)[0]


def admissibility(tree, fadmiss=None):
    """Block tree of (tree x tree): returns (full, low) lists of (row node, col node).

    Leaf x leaf pairs are full blocks; other pairs are low-rank when
    ``fadmiss(rad1, rad2, dist)`` holds (default 2.5 min(rad1, rad2) < dist),
    else they are split into sons (a leaf stands for itself).
    """
    son, mid, rad = tree.son, tree.mid, tree.rad
    leaf = son[:, 0] < 0
    full, low = [], []
    stack = [(0, 0)]
    while stack:
        i1, i2 = stack.pop()
        if leaf[i1] and leaf[i2]:
            full.append((i1, i2))
            continue
        d = float(np.linalg.norm(mid[i1] - mid[i2]))
        ok = fadmiss(rad[i1], rad[i2], d) if fadmiss else 2.5 * min(rad[i1], rad[i2]) < d
        if ok:
            low.append((i1, i2))
            continue
        s1 = (i1,) if leaf[i1] else tuple(son[i1])
        s2 = (i2,) if leaf[i2] else tuple(son[i2])
        stack.extend((a, b) for a in reversed(s1) for b in reversed(s2))
    # column-major order of the (node x node) admissibility matrix, as find()
    full.sort(key=lambda t: (t[1], t[0]))
    low.sort(key=lambda t: (t[1], t[0]))
    return full, low


def block_ranges(tree, blocks):
    """4 x nb int64 array [row_begin; row_end; col_begin; col_end] (cluster indices)."""
    b = np.array(blocks, dtype=np.int64).reshape(-1, 2)
    return np.stack([tree.cbeg[b[:, 0]], tree.cend[b[:, 0]], tree.cbeg[b[:, 1]], tree.cend[b[:, 1]]])


class hmatrix:
    """H-matrix of one Green-function operator on the GPU (G, F, H1 or H2).

    ``hmatrix(green, tree, htol=1e-6, kmax=100)`` builds the block tree;
    ``fill`` evaluates full blocks with refinement and compresses admissible
    blocks by ACA (Greenfun/+aca/@compgreenret/eval.m); ``matvec``
    multiplies.  ``precision`` ('fp64' or 'fp32') is the storage of blocks
    and factors.
    """

    def __init__(self, green, tree, htol=1e-6, kmax=100, fadmiss=None, precision="fp64"):
        import ctypes as C
        from . import _lib
        self.green, self.tree = green, tree
        full, low = admissibility(tree, fadmiss)
        self.full = np.ascontiguousarray(block_ranges(tree, full).T)
        self.low = np.ascontiguousarray(block_ranges(tree, low).T)
        perm = np.ascontiguousarray(tree.perm, dtype=np.int64)
        h = C.c_void_p()
        prec = {"fp64": _lib.MNP_FP64, "fp32": _lib.MNP_FP32}[precision]
        _lib.call("mnp_hmat_create", green.handle, prec, perm.ctypes.data_as(C.c_void_p), len(self.full),
                  self.full.ctypes.data_as(C.c_void_p), len(self.low),
                  self.low.ctypes.data_as(C.c_void_p), float(htol), int(kmax), C.byref(h))
        self.handle = h

    def __del__(self):
        try:
            from . import _lib
        except Exception:  # interpreter shutdown: the module may be gone
            return
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib.can_release():
            _lib.lib().mnp_hmat_destroy(h)
            self.handle = None

    def fill(self, k, key="G"):
        from . import _lib
        _lib.call("mnp_hmat_fill", self.handle, _lib.MnpComplex.of(k), _lib.MNP_KEY[key])
        return self

    def fill_layer(self, k, key, comp, direct, gl, enei):
        """Component comp ('p', 'ss', 'hs', 'sh', 'hh') of the outside Green
        function above the substrate of gl (a layer.greenretlayer): key 'G' or
        'H2'; direct adds the direct part at wavenumber k
        (mnp_hmat_fill_layer)."""
        from . import _lib
        c = ("p", "ss", "hs", "sh", "hh").index(comp) if isinstance(comp, str) else int(comp)
        eu, ed = gl.tab.layer.epsilon(enei)
        _lib.call("mnp_hmat_fill_layer", self.handle, _lib.MnpComplex.of(k), _lib.MNP_KEY[key], c, int(bool(direct)),
                  gl.handle, float(enei), _lib.MnpComplex.of(eu), _lib.MnpComplex.of(ed))
        return self

    def matvec(self, x):
        """y = H x for a device (n,) or (n, nrhs) complex128 column-major tensor."""
        from . import _lib, device as dev
        n = self.green.n
        x2 = x.reshape(n, -1) if x.dim() == 1 else x
        nrhs = x2.shape[1]
        xb = dev.colmajor((n, nrhs))
        xb.copy_(x2)
        y = dev.colmajor((n, nrhs))
        _lib.call("mnp_hmat_matvec", self.handle, nrhs, dev.ptr(dev.base(xb)), dev.ptr(dev.base(y)))
        return y.reshape(x.shape)

    def ranks(self):
        import ctypes as C
        from . import _lib
        r = np.empty(len(self.low), dtype=np.int32)
        _lib.call("mnp_hmat_ranks", self.handle, r.ctypes.data_as(C.c_void_p))
        return r

    def coarsen(self):
        """(tilebeg, C): the row tiles of the product plan (first cluster
        indices, ntile + 1) and the coarse operator P^T H P summed over them
        (device ntile x ntile complex128, column-major; mnp_hmat_coarsen)."""
        import ctypes as C
        from . import _lib, device as dev
        nt = C.c_int64()
        _lib.call("mnp_hmat_coarsen", self.handle, C.byref(nt), None, None)
        tb = np.zeros(nt.value + 1, dtype=np.int64)
        Cm = dev.colmajor((nt.value, nt.value))
        _lib.call("mnp_hmat_coarsen", self.handle, C.byref(nt), tb.ctypes.data_as(C.c_void_p),
                  dev.ptr(dev.base(Cm)))
        return tb, Cm

    def stats(self):
        import ctypes as C
        from . import _lib
        a, b, m, c = C.c_int64(), C.c_int64(), C.c_int(), C.c_double()
        _lib.call("mnp_hmat_stats", self.handle, C.byref(a), C.byref(b), C.byref(m), C.byref(c))
        return dict(full_entries=a.value, lowrank_entries=b.value, max_rank=m.value, compression=c.value)
