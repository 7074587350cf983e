"""CPU restatement (NumPy, complex FP64) of MNPBEM's dense retarded BEM path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``mnpbem_amd`` imports this module; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, as the checker.  Each function cites the MNPBEM file and
lines whose algorithm it follows (paths relative to the MNPBEM tree).

Conventions: faces are 0-based int (n, 3) flat triangles; matrices are
(rows = observation faces, cols = source faces); refined pairs are listed in
column-major order (sorted by column, then row), which is the order of
Matlab's ``find(ir)``.

Pins (see tests/test_oracle_*.py): quadrature exactness and closed-form
self-integrals (rule 18, polar rule); brute-force refinement selection;
Taylor tables against direct quadrature of exp(ikr)/r; the Gauss sum rule;
closed-form kernels and reciprocity; algebraic chain identities and the
exact BEM operator residual; Mie theory for spheres (BEM converges to it);
Rayleigh limit and energy conservation of the Mie restatement.
"""
import math

import numpy as np

# Misc/units.m:2
EV2NM = 1.0 / 8.0655477e-4
FACT = [float(math.factorial(n)) for n in range(16)]


# --------------------------------------------------------------------------
# quadrature (Misc/integration)
# --------------------------------------------------------------------------
def triangle_rule18():
    """28-point, precision-11 rule on the unit triangle (Lyness & Jespersen).

    Misc/integration/@quadface/private/triangle_unit_set.m:471-505.
    Returns x, y, w with sum(w) == 1.
    """
    a = 1.0 / 3.0
    b, c = 0.9480217181434233, 0.02598914092828833
    d, e = 0.8114249947041546, 0.09428750264792270
    f, g = 0.01072644996557060, 0.4946367750172147
    p, q = 0.5853132347709715, 0.2073433826145142
    r, s = 0.1221843885990187, 0.4389078057004907
    t, u = 0.6779376548825902, 0.04484167758913055
    v = 0.27722066752827925
    w_, x_, y_ = 0.8588702812826364, 0.0, 0.1411297187173636
    w1, w2, w3 = 0.08797730116222190, 0.008744311553736190, 0.03808157199393533
    w4, w5, w6 = 0.01885544805613125, 0.07215969754474100, 0.06932913870553720
    w7, w8 = 0.04105631542928860, 0.007362383783300573
    xt = [a, b, c, c, d, e, e, f, g, g, p, q, q, r, s, s, t, t, u, u, v, v,
          w_, w_, x_, x_, y_, y_]
    yt = [a, c, b, c, e, d, e, g, f, g, q, p, q, s, r, s, u, v, t, v, t, u,
          x_, y_, w_, y_, w_, x_]
    wt = [w1, w2, w2, w2, w3, w3, w3, w4, w4, w4, w5, w5, w5, w6, w6, w6,
          w7, w7, w7, w7, w7, w7, w8, w8, w8, w8, w8, w8]
    return np.array(xt), np.array(yt), np.array(wt)


def lglnodes(n):
    """Legendre-Gauss-Lobatto nodes/weights, n+1 points, Newton from Chebyshev.

    Misc/integration/lglnodes.m:13-32 (node order: +1 down to -1).
    """
    n1 = n + 1
    x = np.cos(np.pi * np.arange(n + 1) / n)
    p = np.zeros((n1, n1))
    xold = 2.0
    while np.max(np.abs(x - xold)) > np.finfo(float).eps:
        xold = x
        p[:, 0] = 1.0
        p[:, 1] = x
        for k in range(2, n + 1):
            p[:, k] = ((2 * k - 1) * x * p[:, k - 1] - (k - 1) * p[:, k - 2]) / k
        x = xold - (x * p[:, n1 - 1] - p[:, n - 1]) / (n1 * p[:, n1 - 1])
    w = 2.0 / (n * n1 * p[:, n1 - 1] ** 2)
    return x, w


def polar_rule(npol=(7, 5)):
    """Polar self-integration rule for a triangle (144 points for npol=[7,5]).

    Misc/integration/@quadface/private/init.m:25-53.  Returns x3, y3, w3 in
    unit-triangle coordinates (point = x*v1 + y*v2 + (1-x-y)*v3), sum(w3)==1.
    """
    x1, w1 = lglnodes(npol[0])
    rho = 0.5 * (x1 + 1 + 1e-6)
    x2, w2 = lglnodes(npol[1])
    phi = (270 + 60 * x2) / 180 * np.pi
    phi0 = 120 / 180 * np.pi
    # ndgrid(rho, phi) flattened column-major: rho varies fastest
    R, P = np.meshgrid(rho, phi, indexing="ij")
    rho_f, phi_f = R.ravel(order="F"), P.ravel(order="F")
    rad = 1.0 / np.abs(2 * np.sin(phi_f))
    phis = np.stack([phi_f, phi_f + phi0, phi_f + 2 * phi0], axis=1)
    x = (np.cos(phis) * (rho_f * rad)[:, None]).ravel(order="F")
    y = (np.sin(phis) * (rho_f * rad)[:, None]).ravel(order="F")
    x, y = (1 - np.sqrt(3) * x - y) / 3, (1 + np.sqrt(3) * x - y) / 3
    W = np.outer(w1, w2).ravel(order="F")
    w = np.tile(W * rho_f * rad ** 2, 3)
    w = w / np.sum(w)
    return x, y, w


# --------------------------------------------------------------------------
# particle geometry (Particles/@particle, Misc/+misc)
# --------------------------------------------------------------------------
def norm_flat(verts, faces):
    """Centroids, areas and basis vectors of flat triangles.

    Particles/@particle/private/norm_flat.m:17-50.
    Returns pos, tvec1, tvec2, nvec (n,3) and area (n,).
    """
    v1, v2, v3 = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    pos = (v1 + v2 + v3) / 3
    vec1, vec2 = v1 - v2, v3 - v2
    nvec = np.cross(vec1, vec2)
    area = 0.5 * np.sqrt(np.sum(nvec * nvec, axis=1))
    vec1 = vec1 / np.sqrt(np.sum(vec1 * vec1, axis=1))[:, None]
    nvec = nvec / np.sqrt(np.sum(nvec * nvec, axis=1))[:, None]
    vec2 = np.cross(nvec, vec1)
    return pos, vec1, vec2, nvec, area


def bradius(verts, faces, pos):
    """Max centroid-vertex distance per face (Misc/+misc/bradius.m:11-32)."""
    r = np.zeros(len(faces))
    for i in range(3):
        r = np.maximum(r, np.sqrt(np.sum((pos - verts[faces[:, i]]) ** 2, axis=1)))
    return r


def pdist2(p1, p2):
    """Distances via |a|^2+|b|^2-2a.b with d^2<1e-10 -> 0 (Misc/+misc/pdist2.m:13-17)."""
    d = np.sum(p1 * p1, axis=1)[:, None] + np.sum(p2 * p2, axis=1)[None, :] - 2 * p1 @ p2.T
    d[d < 1e-10] = 0
    return np.sqrt(d)


def refinematrix(pos, rad, relcutoff=3.0, abscutoff=0.0):
    """Refinement selector (Greenfun/+green/refinematrix.m:21-68).

    Returns (rows, cols, kind) in column-major order; kind 2 = coincident
    centroids (self term), 1 = near-field pair.  The relative distance is
    ``(d_ij - rad_j) / rad_i`` (``rad`` of the observation face, :56-60).
    """
    d = pdist2(pos, pos)
    d2 = d - rad[None, :]
    idr = d2 / rad[:, None]
    sel2 = d == 0
    sel1 = ((d2 < abscutoff) | (idr < relcutoff)) & (d != 0)
    kind = np.zeros(d.shape, dtype=np.int8)
    kind[sel1] = 1
    kind[sel2] = 2
    cols, rows = np.nonzero(kind.T)  # column-major order
    return rows.astype(np.int64), cols.astype(np.int64), kind[rows, cols]


def quad_flat(verts, faces, ind):
    """Rule-18 points and area-scaled weights for faces ``ind``.

    Particles/@particle/private/quad_flat.m:21-51 (area from
    cross(v2-v1, v3-v1)).  Returns pos (len(ind), 28, 3), w (len(ind), 28).
    """
    x, y, w = triangle_rule18()
    f = faces[ind]
    v1, v2, v3 = verts[f[:, 0]], verts[f[:, 1]], verts[f[:, 2]]
    nv = np.cross(v2 - v1, v3 - v1)
    area = 0.5 * np.sqrt(np.sum(nv * nv, axis=1))
    pts = (x[None, :, None] * v1[:, None, :] + y[None, :, None] * v2[:, None, :]
           + (1 - x - y)[None, :, None] * v3[:, None, :])
    return pts, w[None, :] * area[:, None]


def quadpol_flat(verts, faces, area, ind, npol=(7, 5)):
    """Polar-rule points and weights (Particles/@particle/private/quadpol_flat.m:33-50)."""
    x, y, w = polar_rule(npol)
    f = faces[ind]
    v1, v2, v3 = verts[f[:, 0]], verts[f[:, 1]], verts[f[:, 2]]
    pts = (x[None, :, None] * v1[:, None, :] + y[None, :, None] * v2[:, None, :]
           + (1 - x - y)[None, :, None] * v3[:, None, :])
    return pts, w[None, :] * area[ind][:, None]


# --------------------------------------------------------------------------
# Green function (Greenfun/@greenret, @compgreenret)
# --------------------------------------------------------------------------
class GreenRet:
    """Retarded Green function of a closed flat-triangle particle with itself.

    Construction follows Greenfun/@greenret/private/init.m (deriv='cart')
    and Greenfun/@compgreenret/private/initclosed.m (closed-surface sum rule);
    evaluation follows Greenfun/@greenret/private/eval1.m.
    """

    def __init__(self, verts, faces, order=5, relcutoff=3.0, abscutoff=0.0,
                 npol=(7, 5), closed=True):
        self.verts = np.asarray(verts, dtype=np.float64)
        self.faces = np.asarray(faces, dtype=np.int64)
        self.order = order
        self.pos, self.tvec1, self.tvec2, self.nvec, self.area = norm_flat(self.verts, self.faces)
        self.n = len(self.faces)
        self.rad = bradius(self.verts, self.faces, self.pos)
        self.rows, self.cols, self.kind = refinematrix(self.pos, self.rad, relcutoff, abscutoff)
        self._init_tables(npol)
        if closed:
            self._init_closed()

    # greenret/private/init.m:42-177
    def _init_tables(self, npol):
        order = self.order
        nnz = len(self.rows)
        g = np.zeros((nnz, order + 1))
        f = np.zeros((nnz, 3, order + 1))
        pos, nvec, t1, t2 = self.pos, self.nvec, self.tvec1, self.tvec2
        # diagonal (self) elements, polar integration: init.m:44-95
        sd = np.nonzero(self.kind == 2)[0]
        if len(sd):
            face = self.rows[sd]
            face2 = self.cols[sd]
            pts, w = quadpol_flat(self.verts, self.faces, self.area, face2, npol)
            vec = pos[face][:, None, :] - pts
            r = np.sqrt(np.sum(vec * vec, axis=2))
            for o in range(order + 1):
                g[sd, o] = np.sum(w * r ** (o - 1), axis=1) / FACT[o]
            inn = np.sum(vec * nvec[face][:, None, :], axis=2)
            rr = np.maximum(r, 1e-4 * np.max(r))
            in1 = np.sum(vec * t1[face][:, None, :], axis=2)
            in2 = np.sum(vec * t2[face][:, None, :], axis=2)
            for o in range(order + 1):
                fn = np.sum(w * ((o - 1) * inn * r ** (o - 3)), axis=1) / FACT[o]
                f1 = np.sum(w * ((o - 1) * in1 * r ** o / rr ** 3), axis=1) / FACT[o]
                f2 = np.sum(w * ((o - 1) * in2 * r ** o / rr ** 3), axis=1) / FACT[o]
                f[sd, :, o] = (nvec[face] * fn[:, None] + t1[face] * f1[:, None]
                               + t2[face] * f2[:, None])
        # near-field elements, rule-18 integration: init.m:98-177
        sn = np.nonzero(self.kind == 1)[0]
        if len(sn):
            i, j = self.rows[sn], self.cols[sn]
            pts, w = quad_flat(self.verts, self.faces, j)
            x = pos[i][:, None, :] - pts
            r = np.sqrt(np.sum(x * x, axis=2))
            vec0 = -(pos[i] - pos[j])
            r0 = np.sqrt(np.sum(vec0 * vec0, axis=1))
            dr = r - r0[:, None]
            for o in range(order + 1):
                g[sn, o] = np.sum((dr ** o / r) / FACT[o] * w, axis=1)
            f[sn, :, 0] = -np.sum(x * (w / r ** 3)[:, :, None], axis=1)
            for o in range(1, order + 1):
                fac = -dr ** o / (r ** 3 * FACT[o]) + dr ** (o - 1) / (r ** 2 * FACT[o - 1])
                f[sn, :, o] = np.sum(x * (fac * w)[:, :, None], axis=1)
        self.g, self.f = g, f

    def _F_static(self):
        """Static (k=0) surface derivative F with refinement (eval2.m with k=0)."""
        pos, nvec, area = self.pos, self.nvec, self.area
        dx = pos[:, None, :] - pos[None, :, :]
        d = np.maximum(np.sqrt(np.sum(dx * dx, axis=2)), np.finfo(float).eps)
        inn = np.sum(nvec[:, None, :] * dx, axis=2)
        F = inn * (-1.0 / d) / d ** 2 * area[None, :]
        F[self.rows, self.cols] = np.sum(nvec[self.rows] * self.f[:, :, 0], axis=1)
        return F

    # compgreenret/private/initclosed.m:50-89 and greenret/diag.m:10-30
    def _init_closed(self):
        F = self._F_static()
        fsum = np.sum(self.area[:, None] * (F * (1.0 / self.area)[None, :]), axis=0)
        sd = np.nonzero(self.kind == 2)[0]
        face = self.rows[sd]
        self.f[sd, :, 0] += self.nvec[face] * (-2 * np.pi - fsum[face])[:, None]

    def ref_G(self, k):
        """Refined G entries before the phase factor: sum_n g_n (ik)^n."""
        return self.g @ ((1j * k) ** np.arange(self.order + 1))

    def ref_F(self, k):
        """Refined F entries before the phase factor: n_i . sum_n f_n (ik)^n."""
        fr = (1j * k) ** np.arange(self.order + 1)
        return np.sum(self.nvec[self.rows][:, :, None] * self.f, axis=1) @ fr

    def eval(self, k, keys=("G", "F")):
        """Full matrices G, F, H1 = F + 2 pi I, H2 = F - 2 pi I at wavenumber k.

        Greenfun/@greenret/private/eval1.m:18-40 (G) and :110-124, :140-143
        (cart, F-only branch).
        """
        pos, nvec, area = self.pos, self.nvec, self.area
        x = pos[:, 0][:, None] - pos[:, 0][None, :]
        y = pos[:, 1][:, None] - pos[:, 1][None, :]
        z = pos[:, 2][:, None] - pos[:, 2][None, :]
        d = np.maximum(np.sqrt(x ** 2 + y ** 2 + z ** 2), np.finfo(float).eps)
        ph = np.exp(1j * k * d)
        out = {}
        if "G" in keys:
            G = (1.0 / d) * area[None, :] + 0j
            G[self.rows, self.cols] = self.ref_G(k)
            out["G"] = G * ph
        if any(key in keys for key in ("F", "H1", "H2")):
            fa = (1j * k - 1.0 / d) / d ** 2
            F = ((fa * x) * nvec[:, 0][:, None] + (fa * y) * nvec[:, 1][:, None]
                 + (fa * z) * nvec[:, 2][:, None]) * area[None, :]
            F[self.rows, self.cols] = self.ref_F(k)
            F = F * ph
            if "F" in keys:
                out["F"] = F
            I2 = 2 * np.pi * np.eye(self.n)
            if "H1" in keys:
                out["H1"] = F + I2
            if "H2" in keys:
                out["H2"] = F - I2
        return out


# --------------------------------------------------------------------------
# materials (Material/)
# --------------------------------------------------------------------------
def eps_drude(enei, eps_inf, wp, gamma):
    """Drude function eps_inf - wp^2/(w(w+i gamma)), w = eV2nm/enei.

    Material/@epsdrude/subsref.m:16-28 (parameters in eV).
    """
    w = EV2NM / np.asarray(enei, dtype=np.float64)
    return eps_inf - wp ** 2 / (w * (w + 1j * gamma))


def wavenumber(enei, eps):
    """k = 2 pi / enei * sqrt(eps) (principal root), Material/@epsdrude/subsref.m:24."""
    return 2 * np.pi / enei * np.sqrt(np.asarray(eps, dtype=complex))


# --------------------------------------------------------------------------
# dense solver (BEM/@bemret)
# --------------------------------------------------------------------------
class BemRet:
    """Dense retarded BEM solver for one closed particle (inside medium eps1,
    outside medium eps2).  BEM/@bemret/private/initmat.m, mldivide.m.
    """

    def __init__(self, green, evalfun=None):
        """``evalfun(k, keys)`` -> dict of dense matrices; defaults to green.eval.
        Passing the H-matrix operators here gives the direct solution of the
        system that bemretiter iterates on."""
        self.g = green
        self.evalfun = evalfun or green.eval

    def initmat(self, enei, eps1, eps2):
        """BEM/@bemret/private/initmat.m:5-72 (scalar eps1, eps2; con{1,2}==0)."""
        g = self.g
        nvec = g.nvec
        k = 2 * np.pi / enei
        k1 = 2 * np.pi / enei * np.sqrt(complex(eps1))
        k2 = 2 * np.pi / enei * np.sqrt(complex(eps2))
        m1 = self.evalfun(k1, ("G", "H1"))
        m2 = self.evalfun(k2, ("G", "H2"))
        G1, H1 = m1["G"], m1["H1"]
        G2, H2 = m2["G"], m2["H2"]
        G1i = np.linalg.inv(G1)
        G2i = np.linalg.inv(G2)
        L1, L2 = eps1, eps2
        Sigma1 = H1 @ G1i
        Sigma2 = H2 @ G2i
        Deltai = np.linalg.inv(Sigma1 - Sigma2)
        L = L1 - L2
        Sigma = Sigma1 * L1 - Sigma2 * L2 + k ** 2 * ((L * Deltai) * (nvec @ nvec.T)) * L
        self.enei, self.k, self.eps1, self.eps2 = enei, k, eps1, eps2
        self.G1, self.G2, self.H1, self.H2 = G1, G2, H1, H2
        self.G1i, self.G2i = G1i, G2i
        self.L1, self.L2 = L1, L2
        self.Sigma1, self.Deltai = Sigma1, Deltai
        self.Sigmai = np.linalg.inv(Sigma)
        return self

    def excitation(self, exc):
        """BEM/@bemret/private/excitation.m:21-29 for phi1=phi2=0 plane waves.

        ``exc`` holds a1, a1p, a2, a2p as (n, 3, npol) arrays (or None) and
        optional phi1, phi1p, phi2, phi2p (n, npol).
        """
        n, k, nvec = self.g.n, self.k, self.g.nvec
        e1, e2 = self.eps1, self.eps2
        npol = next(v.shape[-1] for v in exc.values() if v is not None)
        phi1 = exc.get("phi1") if exc.get("phi1") is not None else np.zeros((n, npol))
        phi2 = exc.get("phi2") if exc.get("phi2") is not None else np.zeros((n, npol))
        phi1p = exc.get("phi1p") if exc.get("phi1p") is not None else np.zeros((n, npol))
        phi2p = exc.get("phi2p") if exc.get("phi2p") is not None else np.zeros((n, npol))
        a1 = exc.get("a1") if exc.get("a1") is not None else np.zeros((n, 3, npol))
        a2 = exc.get("a2") if exc.get("a2") is not None else np.zeros((n, 3, npol))
        a1p = exc.get("a1p") if exc.get("a1p") is not None else np.zeros((n, 3, npol))
        a2p = exc.get("a2p") if exc.get("a2p") is not None else np.zeros((n, 3, npol))
        phi = phi2 - phi1
        a = a2 - a1
        alpha = a2p - a1p - 1j * k * (nvec[:, :, None] * (phi2 * e2)[:, None, :]
                                      - nvec[:, :, None] * (phi1 * e1)[:, None, :])
        De = e2 * phi2p - e1 * phi1p - 1j * k * (
            np.einsum("ic,icp->ip", nvec, a2 * e2) - np.einsum("ic,icp->ip", nvec, a1 * e1))
        return phi, a, alpha, De

    def solve(self, exc):
        """BEM/@bemret/mldivide.m:11-49.  Returns sig1, sig2 (n,npol), h1, h2 (n,3,npol)."""
        phi, a, alpha, De = self.excitation(exc)
        k, nvec = self.k, self.g.nvec
        L1, L2 = self.L1, self.L2
        mm = lambda A, X: np.einsum("ij,j...->i...", A, X)
        inner = lambda X: np.einsum("ic,icp->ip", nvec, X)
        outer = lambda X: nvec[:, :, None] * X[:, None, :]
        alpha = alpha - mm(self.Sigma1, a) + 1j * k * outer(L1 * phi)
        De = De - mm(self.Sigma1, L1 * phi) + 1j * k * inner(L1 * a)
        sig2 = mm(self.Sigmai, De + 1j * k * inner((L1 - L2) * mm(self.Deltai, alpha)))
        h2 = mm(self.Deltai, 1j * k * outer((L1 - L2) * sig2) + alpha)
        sig1 = mm(self.G1i, sig2 + phi)
        h1 = mm(self.G1i, h2 + a)
        sig2 = mm(self.G2i, sig2)
        h2 = mm(self.G2i, h2)
        return {"sig1": sig1, "sig2": sig2, "h1": h1, "h2": h2}

    def afun(self, sig):
        """Exact BEM operator [sig1,h1,sig2,h2] -> [phi,a,De,alpha].

        BEM/@bemretiter/private/afun.m:6-35 with dense G1, G2, H1, H2.
        """
        k, nvec, e1, e2 = self.k, self.g.nvec, self.eps1, self.eps2
        mm = lambda A, X: np.einsum("ij,j...->i...", A, X)
        inner = lambda X: np.einsum("ic,icp->ip", nvec, X)
        outer = lambda X: nvec[:, :, None] * X[:, None, :]
        Gs1, Gh1 = mm(self.G1, sig["sig1"]), mm(self.G1, sig["h1"])
        Gs2, Gh2 = mm(self.G2, sig["sig2"]), mm(self.G2, sig["h2"])
        Hs1, Hh1 = mm(self.H1, sig["sig1"]), mm(self.H1, sig["h1"])
        Hs2, Hh2 = mm(self.H2, sig["sig2"]), mm(self.H2, sig["h2"])
        phi = Gs1 - Gs2
        a = Gh1 - Gh2
        alpha = Hh1 - Hh2 - 1j * k * outer(e1 * Gs1 - e2 * Gs2)
        De = e1 * Hs1 - e2 * Hs2 - 1j * k * inner(e1 * Gh1 - e2 * Gh2)
        return phi, a, De, alpha


# --------------------------------------------------------------------------
# plane-wave excitation and cross sections (Simulation/retarded)
# --------------------------------------------------------------------------
def planewave_potential(pos, nvec, pol, dirs, enei, eps_medium):
    """a, a' on faces whose outside is the embedding medium.

    Simulation/retarded/@planewaveret/potential.m:47-88 (inout column 2 is
    the medium for a particle embedded in it, so the result is a2, a2p).
    Returns a2, a2p as (n, 3, npol).
    """
    pol = np.atleast_2d(pol).astype(float)
    dirs = np.atleast_2d(dirs).astype(float)
    assert not np.any(np.sum(pol * dirs, axis=1))
    nb = np.sqrt(complex(eps_medium))
    k0 = 2 * np.pi / enei
    k = k0 * nb
    n, npol = len(pos), len(pol)
    a = np.zeros((n, 3, npol), complex)
    ap = np.zeros((n, 3, npol), complex)
    for i in range(npol):
        phase = np.exp(1j * k * (pos @ dirs[i])) / (1j * k0)
        a[:, :, i] = phase[:, None] * pol[i][None, :]
        ap[:, :, i] = ((1j * k * (nvec @ dirs[i])) * phase)[:, None] * pol[i][None, :]
    return a, ap


def farfield(pos, area, sig2, h2, dirs, enei, eps_medium):
    """Far fields of the outside surface charges/currents.

    Simulation/retarded/@spectrumret/farfield.m:42-100 (only the faces whose
    outside is the medium contribute).  Returns e, h as (ndir, 3, npol).
    """
    k = 2 * np.pi / enei * np.sqrt(complex(eps_medium))
    k0 = 2 * np.pi / enei
    phase = np.exp(-1j * k * (dirs @ pos.T)) * area[None, :]
    ph2 = np.einsum("dj,jcp->dcp", phase, h2)
    ps2 = phase @ sig2
    e = 1j * k0 * ph2 - 1j * k * dirs[:, :, None] * ps2[:, None, :]
    h = 1j * k * np.cross(dirs[:, :, None], ph2, axisa=1, axisb=1, axisc=1)
    return e, h, k


def extinction(pos, area, sig, pol, dirs, enei, eps_medium):
    """sigma_ext = 4 pi / k Im(pol . e(dir))  (@planewaveret/extinction.m:12-16)."""
    pol = np.atleast_2d(pol).astype(float)
    dirs = np.atleast_2d(dirs).astype(float)
    e, _, k = farfield(pos, area, sig["sig2"], sig["h2"], dirs, enei, eps_medium)
    val = np.einsum("ic,icp->ip", pol, e)
    return np.real(4 * np.pi / k * np.imag(np.diag(val)))


def scattering(pos, area, sig, enei, eps_medium, pinfty_nvec, pinfty_area):
    """sigma_sca: Poynting flux over the unit sphere / (0.5 nb).

    @spectrumret/scattering.m -> Simulation/retarded/scattering.m:14-31 and
    @planewaveret/scattering.m:13-19.
    """
    e, h, _ = farfield(pos, area, sig["sig2"], sig["h2"], pinfty_nvec, enei, eps_medium)
    s = np.cross(e, np.conj(h), axisa=1, axisb=1, axisc=1)
    dsca = 0.5 * np.real(np.einsum("dc,dcp->dp", pinfty_nvec, s))
    sca = pinfty_area @ dsca
    nb = np.sqrt(complex(eps_medium))
    return np.real(sca / (0.5 * nb))


# --------------------------------------------------------------------------
# quasistatic solver (BEM/@bemstat) and plane wave (Simulation/static)
# --------------------------------------------------------------------------
class BemStat:
    """Quasistatic BEM solver for one closed particle: sig = -(diag(Lambda) + F)^{-1} phi'
    with Lambda = 2 pi (eps1 + eps2) / (eps1 - eps2) (BEM/@bemstat/subsref.m:31-40,
    mldivide.m:13-15).  F is the static surface derivative with refinement and the
    closed-surface correction (Greenfun/@compgreenstat/init.m:24-55,
    @greenstat/private/init.m), which equals the retarded F at k = 0."""

    def __init__(self, green):
        self.g = green
        self.F = green.eval(0.0, ("F",))["F"]

    def initmat(self, enei, eps1, eps2):
        lam = 2 * np.pi * (eps1 + eps2) / (eps1 - eps2)
        self.enei = enei
        self.mat = -np.linalg.inv(np.diag(np.full(self.g.n, lam)) + self.F)
        return self

    def solve(self, phip):
        return self.mat @ phip


def planewavestat_potential(nvec, pol):
    """phi' = -nvec . pol (Simulation/static/@planewavestat/potential.m:13)."""
    return -nvec @ np.atleast_2d(pol).T


def stat_cross_sections(pos, area, sig, pol, enei, eps_medium):
    """(ext, sca, abs) from the dipole moment p = sum area pos sig
    (@planewavestat/scattering.m:14-18, absorption.m:14-19, extinction.m)."""
    pol = np.atleast_2d(pol)
    dip = (area[:, None] * pos).T @ sig                     # (3, npol)
    k = 2 * np.pi / enei * np.sqrt(complex(eps_medium))
    sca = 8 * np.pi / 3 * k ** 4 * np.sum(np.abs(dip) ** 2, axis=0)
    ab = 4 * np.pi * k * np.imag(np.sum(pol.T * dip, axis=0))
    return np.real(sca + ab), np.real(sca), np.real(ab)
