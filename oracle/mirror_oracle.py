"""CPU restatement of MNPBEM's retarded solver with mirror symmetry
(TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product).

- full particle: Particles/@comparticlemirror/private/init.m:29-54 (symmetry
  table; copies [p, flip x] for 'x', [p, flip y] for 'y', [p, flip x,
  flip y, flip y of flip x] for 'xy'), Particles/@particle/flip.m,
  flipfaces.m (vertex coordinate negated, vertex order reversed);
- symmetry values: Particles/@comparticlemirror/symvalue.m:14-31, symindex;
- Green function per symmetry value: Greenfun/@compgreenretmirror/eval.m:23-42
  (G_s = sum_j tab[s, j] G(p, copy j), the rows of p in compgreenret(p, full(p)));
- solver: BEM/@bemretmirror/private/initmat.m:25-66 (explicit inverses, as
  written), initsigmai.m:19-30, mldivide.m:25-87, excitation.m:12-27;
- quasistatic solver: BEM/@bemstatmirror/subsref.m:29-37, mldivide.m:14-22,
  Greenfun/@compgreenstatmirror/eval.m:18-40; plane wave
  Simulation/static/@planewavestatmirror/potential.m:16-28, full.m:23-39;
- plane wave: Simulation/retarded/@planewaveretmirror/potential.m:16-41
  (pol x and pol y, each for dir +z and -z, with their symmetry values),
  full.m:22-64 and Particles/@compstructmirror/expand.m:22-46.

Pinned by tests/test_mirror_oracle.py: the expanded mirror solution equals
the oracle's dense solution (mnpbem_oracle.BemRet, BemStat) on the full particle, and
the full particle's Green matrices are equivariant under the mirror group.
"""
import numpy as np

SYMTABLE = {
    "x": np.array([[1, 1], [1, -1]]),
    "y": np.array([[1, 1], [1, -1]]),
    "xy": np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, 1, -1], [1, -1, -1, 1]]),
}
SYMVALUE = {"+": [1, 1], "-": [1, -1], "++": [1, 1, 1, 1], "+-": [1, 1, -1, -1],
            "-+": [1, -1, 1, -1], "--": [1, -1, -1, 1]}
# planewaveretmirror/potential.m:27-40: symmetry keys of the x, y, z components
# for the pol-x and pol-y excitations
PLANEWAVE_KEYS = {
    "x": (("+", "-", "-"), ("-", "+", "+")),
    "y": (("+", "-", "+"), ("-", "+", "-")),
    "xy": (("++", "--", "-+"), ("--", "++", "+-")),
}


def flip(verts, faces, k):
    v = np.array(verts, dtype=float, copy=True)
    v[:, k] = -v[:, k]
    return v, np.asarray(faces)[:, ::-1].copy()


def full_mesh(verts, faces, sym):
    """Vertices and faces of the full particle, copies in table-column order."""
    parts = [(np.asarray(verts, float), np.asarray(faces))]
    for k, keys in ((0, ("x", "xy")), (1, ("y", "xy"))):
        if sym in keys:
            parts += [flip(v, f, k) for v, f in parts]
    off = np.cumsum([0] + [len(v) for v, _ in parts])
    return (np.vstack([v for v, _ in parts]),
            np.vstack([f + off[i] for i, (_, f) in enumerate(parts)]))


def symindex(sym, key):
    return int(np.flatnonzero(np.all(SYMTABLE[sym] == np.array(SYMVALUE[key]), axis=1))[0])


class BemRetMirror:
    """``green`` is mnpbem_oracle.GreenRet of the full particle; the first n
    faces are the symmetry part."""

    def __init__(self, green, n, sym):
        self.g, self.n, self.sym = green, n, sym
        self.tab = SYMTABLE[sym]
        self.nvec = green.nvec[:n]

    def _contract(self, M):
        n, J = self.n, len(self.tab)
        return [sum(self.tab[s, j] * M[:n, j * n:(j + 1) * n] for j in range(J)) for s in range(J)]

    def initmat(self, enei, eps1, eps2):
        k = 2 * np.pi / enei
        m1 = self.g.eval(k * np.sqrt(complex(eps1)), ("G", "H1"))
        m2 = self.g.eval(k * np.sqrt(complex(eps2)), ("G", "H2"))
        G1, H1 = self._contract(m1["G"]), self._contract(m1["H1"])
        G2, H2 = self._contract(m2["G"]), self._contract(m2["H2"])
        self.G1i = [np.linalg.inv(a) for a in G1]
        self.G2i = [np.linalg.inv(a) for a in G2]
        self.Sigma1 = [h @ gi for h, gi in zip(H1, self.G1i)]
        self.Sigma2 = [h @ gi for h, gi in zip(H2, self.G2i)]
        self.Deltai = [np.linalg.inv(a - b) for a, b in zip(self.Sigma1, self.Sigma2)]
        self.enei, self.k, self.eps1, self.eps2 = enei, k, eps1, eps2
        self.mats = dict(G1=G1, G2=G2, H1=H1, H2=H2)
        return self

    def sigmai(self, x, y, z):
        """initsigmai.m:19-30 with scalar L1 = eps1, L2 = eps2."""
        L1, L2, k, nv = self.eps1, self.eps2, self.k, self.nvec
        L = L1 - L2
        S = self.Sigma1[z] * L1 - self.Sigma2[z] * L2
        for c, s in enumerate((x, y, z)):
            S = S + k ** 2 * ((L * self.Deltai[s]) * np.outer(nv[:, c], nv[:, c])) * L
        return np.linalg.inv(S)

    def solve(self, exc, sx):
        """mldivide.m:25-87 for one excitation; sx = (x, y, z) symmetry indices."""
        n, k, nv = self.n, self.k, self.nvec
        e1, e2 = self.eps1, self.eps2
        z0 = lambda shape: np.zeros(shape, complex)
        npol = next(v.shape[-1] for v in exc.values() if v is not None)
        get = lambda key, shape: exc[key] if exc.get(key) is not None else z0(shape)
        phi1, phi2 = get("phi1", (n, npol)), get("phi2", (n, npol))
        phi1p, phi2p = get("phi1p", (n, npol)), get("phi2p", (n, npol))
        a1, a2 = get("a1", (n, 3, npol)), get("a2", (n, 3, npol))
        a1p, a2p = get("a1p", (n, 3, npol)), get("a2p", (n, 3, npol))
        phi, a = phi2 - phi1, a2 - a1
        alpha = a2p - a1p - 1j * k * (nv[:, :, None] * (e2 * phi2)[:, None, :] - nv[:, :, None] * (e1 * phi1)[:, None, :])
        De = e2 * phi2p - e1 * phi1p - 1j * k * (np.einsum("ic,icp->ip", nv, e2 * a2) - np.einsum("ic,icp->ip", nv, e1 * a1))
        x, y, z = sx
        S1, Di, G1i, G2i = self.Sigma1, self.Deltai, self.G1i, self.G2i
        L1, L2 = e1, e2
        al = [alpha[:, c] - S1[s] @ a[:, c] + 1j * k * nv[:, c:c + 1] * (L1 * phi) for c, s in enumerate(sx)]
        De = De - S1[z] @ (L1 * phi) + 1j * k * sum(nv[:, c:c + 1] * (L1 * a[:, c]) for c in range(3))
        sig2 = self.sigmai(x, y, z) @ (De + 1j * k * sum(nv[:, c:c + 1] * ((L1 - L2) * (Di[s] @ al[c]))
                                                         for c, s in enumerate(sx)))
        h2 = [Di[s] @ (1j * k * nv[:, c:c + 1] * ((L1 - L2) * sig2) + al[c]) for c, s in enumerate(sx)]
        return {"sig1": G1i[z] @ (sig2 + phi), "sig2": G2i[z] @ sig2,
                "h1": np.stack([G1i[s] @ (h2[c] + a[:, c]) for c, s in enumerate(sx)], axis=1),
                "h2": np.stack([G2i[s] @ h2[c] for c, s in enumerate(sx)], axis=1)}


class BemStatMirror:
    """Quasistatic solver with mirror symmetry (BEM/@bemstatmirror):
    F_s = sum_j tab[s, j] F(p, copy j) (Greenfun/@compgreenstatmirror/
    eval.m:18-40, the rows of p in compgreenstat(p, full(p))), mat_s =
    -inv(diag(Lambda) + F_s) (bemstatmirror/subsref.m:29-37), sig = mat_s
    phi' with s the index of the excitation's z-key (mldivide.m:14-22).
    ``green`` is mnpbem_oracle.GreenRet of the full particle; the first n
    faces are the symmetry part."""

    def __init__(self, green, n, sym):
        self.n, self.sym, self.tab = n, sym, SYMTABLE[sym]
        F = green.eval(0.0, ("F",))["F"]
        J = len(self.tab)
        self.F = [sum(self.tab[s, j] * F[:n, j * n:(j + 1) * n] for j in range(J)) for s in range(J)]

    def initmat(self, enei, eps1, eps2):
        lam = 2 * np.pi * (eps1 + eps2) / (eps1 - eps2)
        self.mat = [-np.linalg.inv(lam * np.eye(self.n) + F) for F in self.F]
        return self

    def solve(self, phip, s):
        return self.mat[s] @ phip


def planewavestat_symvals(sym):
    """Symmetry values (3 x J) of the quasistatic pol-x and pol-y
    excitations (Simulation/static/@planewavestatmirror/potential.m:16-28),
    with the pol-y z-key of sym 'x' taken as '+' (the reference's table has
    '-' there; the potential -y is even under x -> -x)."""
    return planewave_symvals(sym)


def full_stat(sols, symvals, pol):
    """planewavestatmirror/full.m:23-39: sig = pol_x sig_x + pol_y sig_y on
    the full particle, each expanded with its symmetry values."""
    fx, fy = (expand(s, v) for s, v in zip(sols, symvals))
    pol = np.atleast_2d(pol)
    return fx @ pol[:, 0][None, :] + fy @ pol[:, 1][None, :]


def expand(val, symval):
    """compstructmirror/expand.m:22-46: symval is 3 x J (x, y, z rows);
    scalars take the last row, vector component c row c."""
    J = symval.shape[1]
    if val.ndim == 2:
        return np.concatenate([symval[-1, j] * val for j in range(J)], axis=0)
    return np.concatenate([val * symval[:, j][None, :, None] for j in range(J)], axis=0)


def planewave_symvals(sym):
    """(symval of the pol-x excitation, of the pol-y excitation), each 3 x J."""
    return tuple(np.array([SYMVALUE[k] for k in keys]) for keys in PLANEWAVE_KEYS[sym])


def full_solution(sols, symvals, pol, dirs):
    """planewaveretmirror/full.m:22-64: sols = (solution of the pol-x
    excitation, of the pol-y one), each with columns (dir +z, dir -z);
    returns the full particle's fields for the requested pol/dir pairs."""
    out = {}
    for key in sols[0]:
        v1, v2 = expand(sols[0][key], symvals[0]), expand(sols[1][key], symvals[1])
        cols = []
        for p, d in zip(np.atleast_2d(pol), np.atleast_2d(dirs)):
            j = 0 if d[2] > 0 else 1
            cols.append(p[0] * v1[..., j] + p[1] * v2[..., j])
        out[key] = np.stack(cols, axis=-1)
    return out
