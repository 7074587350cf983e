"""Compound particles for the retarded solver (TEST INFRASTRUCTURE ONLY).

CPU restatement, for the tests, of MNPBEM's handling of several particles
and nested media:
  - the compound mesh and its connectivity: Particles/@compound/connect.m
    (con{i,j}(a, b) = the medium that side i of part a and side j of part b
    both face, or 0) and Greenfun/@compgreenret/private/eval1.m:16-51 (block
    (a, b) of g{i,j} is the Green function at that medium's wavenumber, or
    zero);
  - the closed-surface sum rule per closed surface:
    Greenfun/@compgreenret/private/initclosed.m:19-89 with
    Particles/@comparticle/closedparticle.m and closed.m (the rows summed
    for a face are those of the parts listed in its part's closed argument);
  - the dense solver with per-face dielectric functions and, when some
    inside medium is some outside medium (con{1,2} ~= 0), the full matrices
    L1 = G1 eps1 G1^{-1}, L2 = G2 eps2 G2^{-1}: BEM/@bemret/private/
    initmat.m:19-72, mldivide.m:11-49, excitation.m:21-29, and the operator
    that chain inverts (BemRetCompound.afun: BEM/@bemretiter/private/
    afun.m:6-35 with eps on the potentials, or, for nested media, eps on the
    charges, which is what the full L matrices amount to);
  - plane-wave potentials on the faces that face the embedding medium from
    either side (Simulation/retarded/@planewaveret/potential.m:47-88) and the
    far field of both sides (@spectrumret/farfield.m:30-60);
  - Mie theory of a coated sphere (Bohren and Huffman, Absorption and
    Scattering of Light by Small Particles, section 8.1, the coefficients
    of the core-shell sphere), the reference's Mie/ has no coated sphere.
"""
import numpy as np
from scipy.special import spherical_jn, spherical_yn

from . import mnpbem_oracle as O


def compound_mesh(parts):
    """Concatenated vertices and faces of several (verts, faces) parts and
    the face range of each part (Particles/@comparticle/norm.m: pc =
    vertcat(p{:}))."""
    verts, faces, index, off, f0 = [], [], [], 0, 0
    for v, f in parts:
        v = np.asarray(v, float)
        f = np.asarray(f, np.int64)
        verts.append(v)
        faces.append(f + off)
        index.append(np.arange(f0, f0 + len(f)))
        off += len(v)
        f0 += len(f)
    return np.vstack(verts), np.vstack(faces), index


def closed_sets(nparts, closed_args):
    """closed{i} of Particles/@comparticle/closed.m: every part listed in an
    argument (not set before) gets that argument's list.  Only positive
    indices (closed surfaces made of the parts themselves)."""
    out = [None] * nparts
    for arg in closed_args:
        arg = [int(a) for a in np.atleast_1d(arg)]
        if any(a <= 0 for a in arg):
            raise NotImplementedError("flipped parts in closed surfaces")
        for a in arg:
            if out[a - 1] is None:
                out[a - 1] = [x - 1 for x in arg]
    return out


class CompGreenRet(O.GreenRet):
    """Green function of a compound of parts with media inout (nparts x 2,
    1-based) and closed arguments; eval_side gives G1, H1 or G2, H2."""

    def __init__(self, parts, inout, closed_args, order=5, relcutoff=3.0, abscutoff=0.0, npol=(7, 5)):
        verts, faces, self.index = compound_mesh(parts)
        self.inout = np.atleast_2d(np.asarray(inout, int))
        self.part = np.concatenate([np.full(len(ix), i) for i, ix in enumerate(self.index)])
        self.closed = closed_sets(len(parts), closed_args)
        super().__init__(verts, faces, order, relcutoff, abscutoff, npol, closed=True)
        self.fin = self.inout[self.part, 0]
        self.fout = self.inout[self.part, 1]

    # initclosed.m:50-89 per closed surface
    def _init_closed(self):
        F = self._F_static()
        w = self.area[:, None] * (F * (1.0 / self.area)[None, :])
        sd = np.nonzero(self.kind == 2)[0]
        for idx in sd:
            c = self.rows[idx]
            cs = self.closed[self.part[c]]
            if cs is None:
                continue
            rows = np.isin(self.part, cs)
            fsum = np.sum(w[rows, c])
            self.f[idx, :, 0] += self.nvec[c] * (-2 * np.pi - fsum)

    def eval_side(self, kmed, side):
        """(G, H) of side 1 (G1 = g{1,1} - g{2,1}, H1 = F + 2 pi I) or 2 (G2 =
        g{2,2} - g{1,2}, H2 = F - 2 pi I) with kmed[m - 1] the wavenumber of
        medium m (initmat.m:27-31, eval1.m:25-39): column c takes the
        wavenumber of its side's medium m_c; row r enters with + when its
        same side faces m_c, with - when its other side does."""
        colmed = self.fin if side == 1 else self.fout
        same = self.fin if side == 1 else self.fout
        other = self.fout if side == 1 else self.fin
        n = self.n
        G = np.zeros((n, n), complex)
        F = np.zeros((n, n), complex)
        for m in np.unique(colmed):
            cols = colmed == m
            ev = self.eval(kmed[m - 1], ("G", "F"))
            sgn = (same[:, None] == m).astype(float) - (other[:, None] == m).astype(float)
            G[:, cols] = (sgn * ev["G"])[:, cols]
            F[:, cols] = (sgn * ev["F"])[:, cols]
        I2 = 2 * np.pi * np.eye(n)
        return G, (F + I2 if side == 1 else F - I2)


class BemRetCompound:
    """bemret for a compound particle: per-face eps1, eps2 and, when an
    inside medium of a part is an outside medium of a part (con{1,2} ~= 0),
    L1 = G1 eps1 G1^{-1}, L2 = G2 eps2 G2^{-1} (initmat.m:34-42)."""

    def __init__(self, green):
        self.g = green
        self.full_L = bool(np.any(green.inout[:, 0][:, None] == green.inout[:, 1][None, :]))

    def initmat(self, enei, eps_med):
        g = self.g
        eps_med = np.asarray(eps_med, complex)
        k = 2 * np.pi / enei
        kmed = k * np.sqrt(eps_med)
        self.G1, self.H1 = g.eval_side(kmed, 1)
        self.G2, self.H2 = g.eval_side(kmed, 2)
        self.e1 = eps_med[g.fin - 1]
        self.e2 = eps_med[g.fout - 1]
        self.G1i = np.linalg.inv(self.G1)
        self.G2i = np.linalg.inv(self.G2)
        if self.full_L:
            self.L1 = self.G1 @ (self.e1[:, None] * self.G1i)
            self.L2 = self.G2 @ (self.e2[:, None] * self.G2i)
        else:
            self.L1, self.L2 = np.diag(self.e1), np.diag(self.e2)
        self.Sigma1 = self.H1 @ self.G1i
        Sigma2 = self.H2 @ self.G2i
        self.Deltai = np.linalg.inv(self.Sigma1 - Sigma2)
        L = self.L1 - self.L2
        nn = g.nvec @ g.nvec.T
        Sigma = self.Sigma1 @ self.L1 - Sigma2 @ self.L2 + k ** 2 * ((L @ self.Deltai) * nn) @ L
        self.Sigmai = np.linalg.inv(Sigma)
        self.enei, self.k = enei, k
        return self

    def excitation(self, exc):
        """excitation.m:21-29 with eps1, eps2 per face."""
        n, k, nvec = self.g.n, self.k, self.g.nvec
        npol = next(v.shape[-1] for v in exc.values() if v is not None)
        z1, z3 = np.zeros((n, npol), complex), np.zeros((n, 3, npol), complex)
        get = lambda key, z: exc.get(key) if exc.get(key) is not None else z
        phi1, phi2, phi1p, phi2p = get("phi1", z1), get("phi2", z1), get("phi1p", z1), get("phi2p", z1)
        a1, a2, a1p, a2p = get("a1", z3), get("a2", z3), get("a1p", z3), get("a2p", z3)
        e1, e2 = self.e1[:, None], self.e2[:, None]
        phi = phi2 - phi1
        a = a2 - a1
        alpha = a2p - a1p - 1j * k * (nvec[:, :, None] * (e2 * phi2)[:, None, :]
                                      - nvec[:, :, None] * (e1 * phi1)[:, None, :])
        De = e2 * phi2p - e1 * phi1p - 1j * k * (np.einsum("ic,icp->ip", nvec, a2) * e2
                                                 - np.einsum("ic,icp->ip", nvec, a1) * e1)
        return phi, a, alpha, De

    def solve(self, exc):
        """mldivide.m:11-49 with matrices L1, L2."""
        phi, a, alpha, De = self.excitation(exc)
        k, nvec = self.k, self.g.nvec
        mm = lambda A, X: np.einsum("ij,j...->i...", A, X)
        inner = lambda X: np.einsum("ic,icp->ip", nvec, X)
        outer = lambda X: nvec[:, :, None] * X[:, None, :]
        L1, L = self.L1, self.L1 - self.L2
        alpha = alpha - mm(self.Sigma1, a) + 1j * k * outer(mm(L1, phi))
        De = De - mm(self.Sigma1, mm(L1, phi)) + 1j * k * inner(mm(L1, a))
        sig2 = mm(self.Sigmai, De + 1j * k * inner(mm(L, mm(self.Deltai, alpha))))
        h2 = mm(self.Deltai, 1j * k * outer(mm(L, sig2)) + alpha)
        return {"sig1": mm(self.G1i, sig2 + phi), "h1": mm(self.G1i, h2 + a),
                "sig2": mm(self.G2i, sig2), "h2": mm(self.G2i, h2)}

    def afun(self, sig):
        """The operator that bemret's chain inverts: afun.m:6-35 with eps1,
        eps2 per face, multiplying the potentials at the faces (rows),
        eps1 G1 sig1, when L1, L2 are diagonal; with the full L1 = G1 eps1
        G1^{-1} of nested media the chain's algebra (mldivide.m:29-40 solved
        for the excitation) puts eps on the charges instead, G1 (eps1 sig1),
        which differs from afun.m's row form wherever G1 couples faces of
        different eps1."""
        k, nvec, e1, e2 = self.k, self.g.nvec, self.e1[:, None], self.e2[:, None]
        mm = lambda A, X: np.einsum("ij,j...->i...", A, X)
        inner = lambda X: np.einsum("ic,icp->ip", nvec, X)
        outer = lambda X: nvec[:, :, None] * X[:, None, :]
        s1, h1, s2, h2 = sig["sig1"], sig["h1"], sig["sig2"], sig["h2"]
        phi = mm(self.G1, s1) - mm(self.G2, s2)
        a = mm(self.G1, h1) - mm(self.G2, h2)
        Hh = mm(self.H1, h1) - mm(self.H2, h2)
        if not self.full_L:
            alpha = Hh - 1j * k * outer(e1 * mm(self.G1, s1) - e2 * mm(self.G2, s2))
            De = e1 * mm(self.H1, s1) - e2 * mm(self.H2, s2) - 1j * k * (e1 * inner(mm(self.G1, h1))
                                                                         - e2 * inner(mm(self.G2, h2)))
        else:
            E1, E2 = self.e1[:, None, None], self.e2[:, None, None]
            alpha = Hh - 1j * k * outer(mm(self.G1, e1 * s1) - mm(self.G2, e2 * s2))
            De = mm(self.H1, e1 * s1) - mm(self.H2, e2 * s2) - 1j * k * inner(mm(self.G1, E1 * h1)
                                                                               - mm(self.G2, E2 * h2))
        return phi, a, De, alpha


def planewave_excitation(green, pol, dirs, enei, eps_medium, medium=1):
    """exc(p, enei) of a plane wave in `medium`: a1, a1p on faces whose inside
    is the medium, a2, a2p on faces whose outside is it (potential.m:47-88)."""
    a, ap = O.planewave_potential(green.pos, green.nvec, pol, dirs, enei, eps_medium)
    m1 = (green.fin == medium)[:, None, None]
    m2 = (green.fout == medium)[:, None, None]
    return {"a1": a * m1, "a1p": ap * m1, "a2": a * m2, "a2p": ap * m2}


def side_charges(green, sig, medium=1):
    """The charges and currents that radiate into `medium`: sig1, h1 of faces
    whose inside is the medium plus sig2, h2 of faces whose outside is it
    (farfield.m:30-60 sums the two contributions)."""
    m1 = green.fin == medium
    m2 = green.fout == medium
    return {"sig2": sig["sig1"] * m1[:, None] + sig["sig2"] * m2[:, None],
            "h2": sig["h1"] * m1[:, None, None] + sig["h2"] * m2[:, None, None]}


def cross_sections(green, sig, pol, dirs, enei, eps_medium, pinfty_nvec, pinfty_area, medium=1):
    s = side_charges(green, sig, medium)
    ext = O.extinction(green.pos, green.area, s, pol, dirs, enei, eps_medium)
    sca = O.scattering(green.pos, green.area, s, enei, eps_medium, pinfty_nvec, pinfty_area)
    return ext, sca, ext - sca


# --------------------------------------------------------------------------
# Mie theory of a coated sphere (Bohren and Huffman, section 8.1)
# --------------------------------------------------------------------------
def _psi(n, z):
    return z * spherical_jn(n, z), spherical_jn(n, z) + z * spherical_jn(n, z, derivative=True)


def _chi(n, z):
    # chi_n(z) = -z y_n(z)
    return -z * spherical_yn(n, z), -(spherical_yn(n, z) + z * spherical_yn(n, z, derivative=True))


def coated_mie_cross_sections(enei, eps_core, eps_shell, eps_out, r_core, r_shell, lmax=30):
    """(ext, sca, abs) in nm^2 of a sphere of radius r_shell with a core of
    radius r_core, from the coefficients a_n, b_n of the coated sphere
    (Bohren and Huffman eqs. 8.2): x = k r_core, y = k r_shell, m1, m2 the
    relative refractive indices of core and shell; sigma_ext = 2 pi / k^2
    sum (2n + 1) Re(a_n + b_n), sigma_sca = 2 pi / k^2 sum (2n + 1)
    (|a_n|^2 + |b_n|^2)."""
    nb = np.sqrt(complex(eps_out))
    k = (2 * np.pi / enei * nb).real
    m1 = np.sqrt(complex(eps_core)) / nb
    m2 = np.sqrt(complex(eps_shell)) / nb
    x, y = k * r_core, k * r_shell
    n = np.arange(1, lmax + 1)
    p1, p1d = _psi(n, m1 * x)
    p2x, p2xd = _psi(n, m2 * x)
    c2x, c2xd = _chi(n, m2 * x)
    A = (m2 * p2x * p1d - m1 * p2xd * p1) / (m2 * c2x * p1d - m1 * c2xd * p1)
    B = (m2 * p1 * p2xd - m1 * p2x * p1d) / (m2 * c2xd * p1 - m1 * p1d * c2x)
    p2y, p2yd = _psi(n, m2 * y)
    c2y, c2yd = _chi(n, m2 * y)
    py, pyd = _psi(n, y + 0j)
    cy, cyd = _chi(n, y + 0j)
    xi, xid = py - 1j * cy, pyd - 1j * cyd
    a = (py * (p2yd - A * c2yd) - m2 * pyd * (p2y - A * c2y)) / (xi * (p2yd - A * c2yd) - m2 * xid * (p2y - A * c2y))
    b = (m2 * py * (p2yd - B * c2yd) - pyd * (p2y - B * c2y)) / (m2 * xi * (p2yd - B * c2yd) - xid * (p2y - B * c2y))
    ext = 2 * np.pi / k ** 2 * np.sum((2 * n + 1) * np.real(a + b))
    sca = 2 * np.pi / k ** 2 * np.sum((2 * n + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2))
    return ext, sca, ext - sca
