"""Oracle (test infrastructure only): CPU restatement of the retarded
substrate path, one interface at z0 with the particle in the upper medium.

  * reflection and transmission coefficients of surface charges and
    currents (Particles/@layerstructure/reflectionsubs.m:17-60);
  * the reflected Green function and its r, z derivatives as Sommerfeld
    integrals over k_parallel, on the reference's path (semi-ellipse to
    2 k1max, then the real axis or a Hankel split up and down the imaginary
    direction; Particles/@layerstructure/green.m:39-127, private/intbessel.m,
    private/inthankel.m).  The reference integrates the path as an ODE
    (ode45, AbsTol 1e-6); here each leg is composite Gauss-Legendre (fixed
    panels, see sommerfeld_nodes), with scipy's Bessel and Hankel functions;
  * the table grid for one particle above the interface and its bilinear
    interpolation of d G, d^3 Fr / r, d^3 Fz / zmin (Particles/@layerstructure/
    private/tabspace2.m, tabspace1.m, adjust.m; Greenfun/@greentablayer/
    private/init.m, norm.m, private/interp2.m; Misc/@igrid2/finterp.m), with
    linear interpolation in wavelength between precomputed tables
    (Greenfun/@greentablayer/eval.m:24-50);
  * the reflected Green matrices G, Gp, F of the particle with itself,
    with the near-pair refinement (shape-function weights of the vertex
    values, Greenfun/@greenretlayer/private/init2.m:23-99) and the polar
    quadrature of diagonal elements (initrefl2.m:39-99); the connected
    outside Green functions of compgreenretlayer (ss = hh = p = direct +
    reflected; sh, hs reflected only; Greenfun/@compgreenretlayer/private/
    eval1.m:31-50);
  * the bemretlayer chain (BEM/@bemretlayer/private/initmat.m:17-95,
    mldivide.m:13-76, private/excitation.m) and the boundary operator the
    chain inverts (afun below: continuity of phi and A with the structured
    outside Green function, and the two normal-derivative conditions);
  * potentials and fields at points above the interface from the reflected
    part (compgreenretlayer(pt, p): potential.m, field.m, private/matmul2.m,
    matmul3.m; greenretlayer with p1 a compoint);
  * the plane wave above the substrate (Simulation/retarded/
    @planewaveretlayer/potential.m:20-105, with Particles/@layerstructure/
    fresnel.m) and its cross sections through the far field with the
    interface (Simulation/retarded/@spectrumretlayer/farfield.m:20-104,
    @planewaveretlayer/extinction.m, scattering.m; efresnel.m);
  * dipoles above the interface (Simulation/retarded/@dipoleretlayer:
    private/greenderiv.m, potential.m, field.m with 'refl', farfield.m,
    decayrate.m), with the reference's forms of the mixed derivative
    F{2,3}, of field.m's h_y and of potential.m's hh term of a2p(:, 3).

Pins (tests/test_layerret_oracle.py): the Sommerfeld identity (reflection
coefficient 1 gives exp(i k D)/D and its derivatives) on both path kinds;
a brute-force real-axis integral for a lossy substrate; the perfect-conductor
and quasistatic limits of the p and ss coefficients; equal media reduce the
solver, the excitation and the cross sections to the free-space ones; the
chain inverts afun for random structured matrices; a small sphere follows the
quasistatic substrate solver; the table interpolation reproduces direct
integration between grid points; a dipole alone above glass has the
Novotny-Hecht rates, and its fields and surface derivatives are finite
differences of its potentials (tests/test_dipolelayer_oracle.py).
"""
import numpy as np
from scipy import special as sp

from . import mnpbem_oracle as O

COMPONENTS = ("p", "ss", "hs", "sh", "hh")  # fieldnames order of reflectionsubs.m


# --------------------------------------------------------------------------
# layer structure with one interface (Particles/@layerstructure)
# --------------------------------------------------------------------------
class Substrate:
    """Upper medium eps_up above z0, lower medium eps_down below; options as
    layerstructure.options (ztol, rmin, zmin, semi, ratio)."""

    def __init__(self, eps_up, eps_down, z0=0.0, ztol=2e-2, rmin=1e-2, zmin=1e-2, semi=0.1, ratio=2.0):
        self.eps_up, self.eps_down = eps_up, eps_down
        self.z0 = float(z0)
        self.ztol, self.rmin, self.zmin, self.semi, self.ratio = ztol, rmin, zmin, semi, ratio

    def eps(self, enei):
        f = lambda e: complex(e(enei)) if callable(e) else complex(e)
        return f(self.eps_up), f(self.eps_down)

    def k(self, enei):
        k0 = 2 * np.pi / enei
        e1, e2 = self.eps(enei)
        return k0, k0 * np.sqrt(e1), k0 * np.sqrt(e2)

    def mindist(self, z):
        return np.abs(np.asarray(z, dtype=float) - self.z0)

    def round(self, z, zmin=None):
        """round.m:13-24: points closer than zmin to the interface move to zmin."""
        zmin = self.zmin if zmin is None else zmin
        z = np.array(z, dtype=float, copy=True)
        d = np.abs(z - self.z0)
        s = np.sign(z - self.z0)
        sel = d <= zmin
        z[sel] = self.z0 + s[sel] * zmin
        return z

    def indlayer(self, z):
        """indlayer.m: 1 above (z > z0... histc of -z), 2 below."""
        z = np.asarray(z, dtype=float)
        return np.where(-z < -self.z0, 1, 2)


def kz_of(k, kpar, offset=False):
    """sqrt(k^2 - kpar^2) on the sheet Im >= 0 (reflectionsubs.m:24-25;
    fresnel.m:19 adds 1e-10 i before taking the sign)."""
    kz = np.sqrt(k * k - kpar * kpar + 0j)
    if offset:
        kz = kz + 1e-10j
        return kz * np.sign(np.imag(kz))
    return kz * np.sign(np.imag(kz + 1e-10j))


def coefficients(k0, e1, e2, kpar, k1z=None, k2z=None):
    """2 x 2 reflection/transmission matrices of every component at kpar
    (reflectionsubs.m:27-60): entry [i, j] for the observer in medium i+1 and
    the source in medium j+1 (1 = above).  Returns dict name -> (2, 2, ...)."""
    k1 = k0 * np.sqrt(e1 + 0j)
    k2 = k0 * np.sqrt(e2 + 0j)
    if k1z is None:
        k1z = kz_of(k1, kpar)
    if k2z is None:
        k2z = kz_of(k2, kpar)
    rr = (k1z - k2z) / (k2z + k1z)
    D = (k2z + k1z) * (e1 * k2z + e2 * k1z)

    def mat1(r1, r2):
        return np.array([[r1, k1z / k2z * (r2 + 1)], [k2z / k1z * (r1 + 1), r2]])

    def mat2(r1, r2):
        return np.array([[r1, -k1z / k2z * r2], [k2z / k1z * r1, -r2]])

    out = {"p": np.array([[rr, 1 + rr], [1 - rr, -rr]])}
    out["ss"] = mat1((k1z + k2z) * (2 * e1 * k1z - e2 * k1z - e1 * k2z) / D,
                     (k2z + k1z) * (2 * e2 * k2z - e1 * k2z - e2 * k1z) / D)
    out["hs"] = mat2(-2 * k0 * (e2 - e1) * e1 * k1z / D, -2 * k0 * (e1 - e2) * e2 * k2z / D)
    out["sh"] = mat2(-2 * k0 * (e2 - e1) * k1z / D, -2 * k0 * (e1 - e2) * k2z / D)
    out["hh"] = mat1((k1z - k2z) * (2 * e1 * k1z - e2 * k1z + e1 * k2z) / D,
                     (k2z - k1z) * (2 * e2 * k2z - e1 * k2z + e2 * k1z) / D)
    return out


# --------------------------------------------------------------------------
# Sommerfeld integrals (green.m, intbessel.m, inthankel.m)
# --------------------------------------------------------------------------
PANELS_ELLIPSE, NODES = 16, 16
TAIL_DECAY = 40.0  # the real-axis and Hankel legs end where exp(-decay) is reached


def sommerfeld_nodes(n=NODES):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1), 0.5 * w  # on [0, 1]


def tail_panels(k1max, scale):
    """Panel edges of a tail leg s in [0, TAIL_DECAY * scale]: widths
    k1max / 4 doubling up to scale (the decay length 1 / Z or 1 / r), then
    constant, so that both the structure of kpar / kz near 2 k1max and the
    decay are resolved."""
    L = TAIL_DECAY * scale
    edges = [0.0]
    h = 0.25 * k1max
    while edges[-1] < L:
        edges.append(min(edges[-1] + min(h, scale), L))
        h *= 2
    return np.array(edges)


def _legs(k1max, semi, r, Z, ratio):
    """Nodes kpar and weights dkpar of the legs for one (r, Z): the
    semi-ellipse kpar = k1max (1 - cos t - i semi sin t), t in [0, pi], and
    either the real axis kpar = 2 k1max + s when Z >= r / ratio
    (green.m:54-57), or the Hankel leg kpar = 2 k1max + i s (its mirror
    conj(kpar) taken by the caller), s on tail_panels."""
    x, w = sommerfeld_nodes()
    t = (np.arange(PANELS_ELLIPSE)[:, None] + x[None, :]).ravel() * np.pi / PANELS_ELLIPSE
    wt = np.tile(w, PANELS_ELLIPSE) * np.pi / PANELS_ELLIPSE
    ka = k1max * (1 - np.cos(t) - 1j * semi * np.sin(t))
    wa = k1max * (np.sin(t) - 1j * semi * np.cos(t)) * wt
    real_leg = Z >= r / ratio
    e = tail_panels(k1max, 1.0 / (Z if real_leg else r))
    h = np.diff(e)
    s = (e[:-1, None] + h[:, None] * x[None, :]).ravel()
    ws = (h[:, None] * w[None, :]).ravel()
    if real_leg:
        return ka, wa, 2 * k1max + s + 0j, ws + 0j, True
    return ka, wa, 2 * k1max + 1j * s, 1j * ws, False


def _integrand(k0, e1, e2, kpar, Z, bes0, bes1):
    """Integrands of G, Fr, Fz for every component at nodes kpar, both points
    above the interface (intbessel.m:33-58 with reflectionsubs' phases;
    the z derivative is taken at the observer, sign(z1 - z0) = +1)."""
    k1 = k0 * np.sqrt(e1 + 0j)
    k1z = kz_of(k1, kpar)
    c = coefficients(k0, e1, e2, kpar)
    ph = np.exp(1j * k1z * Z)
    out = np.empty((5, 3) + np.shape(kpar), dtype=complex)
    for i, nm in enumerate(COMPONENTS):
        R = c[nm][0, 0] * ph
        out[i, 0] = 1j * bes0 * R * kpar / k1z
        out[i, 1] = 1j * bes1 * R * (-kpar * kpar / k1z)
        out[i, 2] = 1j * bes0 * 1j * R * kpar
    return out


def reflected_green(sub, enei, r, Z, coeff=None):
    """Reflected G, Fr, Fz (each (5,)) for one radius r and height sum
    Z = |z1 - z0| + |z2 - z0|, both points above (green.m:22-127).
    coeff: optional override of the coefficients, a function kpar -> (5, ...)
    array (used to pin the quadrature with the Sommerfeld identity)."""
    k0, k1, k2 = sub.k(enei)
    e1, e2 = sub.eps(enei)
    k1max = max(np.real(k1), np.real(k2)) + k0
    ka, wa, kt, wt, real_leg = _legs(k1max, sub.semi, r, Z, sub.ratio)

    def integ(kp, b0, b1):
        if coeff is None:
            return _integrand(k0, e1, e2, kp, Z, b0, b1)
        k1z = kz_of(k1, kp)
        R = coeff(kp) * np.exp(1j * k1z * Z)
        out = np.empty((R.shape[0], 3) + np.shape(kp), dtype=complex)
        out[:, 0] = 1j * b0 * R * kp / k1z
        out[:, 1] = 1j * b1 * R * (-kp * kp / k1z)
        out[:, 2] = -b0 * R * kp
        return out

    res = np.sum(integ(ka, sp.jv(0, ka * r), sp.jv(1, ka * r)) * wa, axis=-1)
    if real_leg:
        res = res + np.sum(integ(kt, sp.jv(0, kt * r), sp.jv(1, kt * r)) * wt, axis=-1)
    else:
        # inthankel.m:37-57: 0.5 H1(kpar r) f(kpar) dkpar + 0.5 H2(kpar* r) f(kpar*) dkpar*
        up = integ(kt, sp.hankel1(0, kt * r), sp.hankel1(1, kt * r)) * wt
        kc = np.conj(kt)
        dn = integ(kc, sp.hankel2(0, kc * r), sp.hankel2(1, kc * r)) * np.conj(wt)
        res = res + 0.5 * np.sum(up + dn, axis=-1)
    return res[:, 0], res[:, 1], res[:, 2]


# --------------------------------------------------------------------------
# table grid and interpolation (tabspace2.m, tabspace1.m, greentablayer)
# --------------------------------------------------------------------------
def tab_grid(sub, verts, nr=30, nz=30, scale=1.05, pts=None):
    """Radii and observer heights of the 2D table for one particle above the
    interface, and for points above it when given (tabspace2.m:16-75 with
    limits, adjust.m:18-33, tabspace1.m:24-48 'log' spacing,
    greentablayer/private/init.m:8-14).  Returns r (nr,), z1 (nz,) and the
    source height z2 of the table."""
    v = np.asarray(verts, dtype=float)
    obs = v if pts is None else np.vstack([v, np.atleast_2d(np.asarray(pts, dtype=float))])
    d = np.sqrt(np.sum((obs[:, None, :2] - v[None, :, :2]) ** 2, axis=-1))
    rlo, rhi = d.min(), d.max()
    m = 0.5 * (rlo + rhi)
    r = [max(m + 0.5 * scale * (rlo - rhi), 0.0), max(m - 0.5 * scale * (rlo - rhi), 0.0)]

    def enlarge(lo, hi):
        if abs(lo - hi) < 1e-10:
            z = [lo - 1e-2, lo + 1e-2]
        else:
            mz = 0.5 * (lo + hi)
            z = [mz + 0.5 * scale * (lo - hi), mz - 0.5 * scale * (lo - hi)]
        if z[0] <= sub.z0:  # adjust.m:45 has <, which leaves a bound exactly at z0 (log10(0) later)
            z[0] = sub.z0 + 1e-10
        return z

    z1 = enlarge(obs[:, 2].min(), obs[:, 2].max())
    z2 = enlarge(v[:, 2].min(), v[:, 2].max())
    z1 = [z1[0] + z2[0] - sub.z0, z1[1] + z2[1] - sub.z0]
    z2s = sub.z0 + 1e-10
    rg = np.logspace(np.log10(max(r[0], sub.rmin)), np.log10(r[1]), nr)
    zz = np.sort(sub.round(np.array(z1), zmin=0.999 * sub.zmin))
    if abs(zz[0] - zz[1]) < 1e-3:
        raise NotImplementedError("table: flat particle at the interface (tabspace1 expand) is not supported")
    zg = sub.z0 + np.logspace(np.log10(zz[0] - sub.z0), np.log10(zz[1] - sub.z0), nz)
    rg = np.unique(np.maximum(sub.rmin, rg))
    zg = np.unique(sub.round(zg))
    z2t = float(sub.round(np.array([z2s]))[0])
    return rg, zg, z2t


class GreenTable:
    """greentablayer for one particle above the interface: tables of the
    normalised d G, d^3 Fr / r, d^3 Fz / zmin on (r, z1) at the source height
    z2, at one or more wavelengths (set.m, eval.m)."""

    def __init__(self, sub, r, z1, z2):
        self.sub, self.r, self.z1, self.z2 = sub, np.asarray(r), np.asarray(z1), float(z2)
        self.enei = None

    def compute(self, enei):
        """Raw G, Fr, Fz on the grid, (3, 5, nr, nz), and the normalised
        tables (norm.m:12-27)."""
        sub = self.sub
        Zt = sub.mindist(self.z1) + sub.mindist(self.z2)
        raw = np.empty((3, 5, len(self.r), len(self.z1)), dtype=complex)
        for i, r in enumerate(self.r):
            for j, Z in enumerate(Zt):
                raw[:, :, i, j] = reflected_green(sub, enei, r, Z)
        R, ZZ = np.meshgrid(self.r, Zt, indexing="ij")
        d = np.sqrt(R ** 2 + ZZ ** 2)
        norm = np.stack([raw[0] * d, raw[1] * d ** 3 / R, raw[2] * d ** 3 / ZZ])
        return raw, norm

    def set(self, eneis):
        self.enei = np.atleast_1d(np.asarray(eneis, dtype=float))
        self.saved = np.stack([self.compute(e)[1] for e in self.enei])
        return self

    def normalised(self, enei):
        """eval.m:17-50: the table at enei (new when nothing is saved,
        linear interpolation in wavelength otherwise)."""
        if self.enei is None:
            return self.compute(enei)[1]
        e = self.enei
        if not (e.min() <= enei <= e.max()):
            raise ValueError("greentablayer: wavelength outside the tabulated range")
        if len(e) == 1:
            return self.saved[0]
        i = min(np.searchsorted(e, enei, side="right") - 1, len(e) - 2)
        t = (enei - e[i]) / (e[i + 1] - e[i])
        return (1 - t) * self.saved[i] + t * self.saved[i + 1]

    def interp(self, tabn, r, z1, z2):
        """interp2.m:17-44 with finterp.m:17-33: G, Fr, Fz (each (5,) + shape)
        and the clamped r, zmin."""
        sub = self.sub
        r = np.maximum(sub.rmin, np.asarray(r, dtype=float))
        z1 = sub.round(z1)
        z2 = sub.round(z2)
        z = z1 + sub.mindist(sub.round(z2))
        zmin = sub.mindist(z)
        d = np.sqrt(r ** 2 + zmin ** 2)
        x, y = self.r, self.z1
        ix = np.searchsorted(x, r.ravel(), side="right")
        iy = np.searchsorted(y, z.ravel(), side="right")
        if np.any(ix == 0) or np.any(iy == 0) or np.any(r.ravel() > x[-1]) or np.any(z.ravel() > y[-1]):
            raise ValueError("greentablayer: point outside the table")
        ix = np.minimum(ix, len(x) - 1) - 1
        iy = np.minimum(iy, len(y) - 1) - 1
        xx = (r.ravel() - x[ix]) / (x[ix + 1] - x[ix])
        yy = (z.ravel() - y[iy]) / (y[iy + 1] - y[iy])
        w = [(1 - xx) * (1 - yy), xx * (1 - yy), (1 - xx) * yy, xx * yy]

        def f(v):
            out = (w[0] * v[..., ix, iy] + w[1] * v[..., ix + 1, iy] + w[2] * v[..., ix, iy + 1]
                   + w[3] * v[..., ix + 1, iy + 1])
            return out.reshape(v.shape[:-2] + r.shape)

        G = f(tabn[0]) / d
        Fr = f(tabn[1]) * r / d ** 3
        Fz = f(tabn[2]) * zmin / d ** 3
        return G, Fr, Fz, r, zmin


# --------------------------------------------------------------------------
# reflected Green matrices of a particle with itself (greenretlayer)
# --------------------------------------------------------------------------
class GreenRetLayer:
    """Reflected G, Gp, F (each (5, n, n) / (5, n, 3, n)) of a flat-triangle
    particle above the interface (Greenfun/@greenretlayer, deriv 'cart')."""

    def __init__(self, sub, verts, faces, table, relcutoff=3.0, abscutoff=0.0, npol=(7, 5)):
        self.sub, self.tab = sub, table
        self.verts = np.asarray(verts, dtype=float)
        self.faces = np.asarray(faces, dtype=np.int64)
        self.pos, _, _, self.nvec, self.area = O.norm_flat(self.verts, self.faces)
        self.n = len(self.faces)
        self.npol = npol
        self.ir = self.refinematrix(relcutoff, abscutoff)
        self._init_refine()

    def refinematrix(self, relcutoff, abscutoff):
        """refinematrixlayer.m:24-56: 1 for near pairs measured to the image,
        2 on refined diagonal elements."""
        sub, pos = self.sub, self.pos
        rad = O.bradius(self.verts, self.faces, pos)
        r = O.pdist2(pos[:, :2], pos[:, :2])
        z = sub.mindist(pos[:, 2])[:, None] + sub.mindist(pos[:, 2])[None, :]
        d = np.sqrt(r ** 2 + z ** 2)
        d2 = d - rad[None, :]
        ir = ((d2 < abscutoff) | (d2 / rad[:, None] < relcutoff)).astype(np.int8)
        ir[np.diag_indices(self.n)] *= 2
        return ir

    def _dist(self, p1, p2):
        """init2.m:103-119."""
        sub = self.sub
        x = p1[:, None, 0] - p2[None, :, 0]
        y = p1[:, None, 1] - p2[None, :, 1]
        z = sub.round(sub.mindist(p1[:, 2]))[:, None] + sub.round(sub.mindist(p2[:, 2]))[None, :]
        r = np.maximum(sub.rmin, np.sqrt(x * x + y * y))
        d = np.sqrt(r * r + z * z)
        return r, z, d, x, y

    def _init_refine(self):
        """init2.m:23-99: per refined pair and vertex, the weights that turn
        the table's values at the vertex into the face integral."""
        rows, cols = np.nonzero(self.ir.T == 1)
        rows, cols = cols, rows  # column-major order of find
        self.ref = []
        x, y, _ = O.triangle_rule18()
        s = np.stack([x, y, 1 - x - y], axis=1)  # shapefunction.m for triangles
        for face in np.unique(cols):
            nb = rows[cols == face]
            qp, qw = O.quad_flat(self.verts, self.faces, [face])
            qp, qw = qp[0], qw[0]
            vts = self.verts[self.faces[face]]
            r0, z0, d0, _, _ = self._dist(self.pos[nb], vts)
            _, z, d, xq, yq = self._dist(self.pos[nb], qp)
            z1 = self.sub.round(self.pos[nb, 2])
            for i in range(3):
                g = (d0[:, i:i + 1] * s[None, :, i] / d) @ qw
                fr0 = (d0[:, i] ** 3 / r0[:, i])[:, None] * s[None, :, i]
                fz0 = (d0[:, i] ** 3 / z0[:, i])[:, None] * s[None, :, i]
                f1 = (fr0 * xq / d ** 3) @ qw
                f2 = (fr0 * yq / d ** 3) @ qw
                fz = (fz0 * z / d ** 3) @ qw
                self.ref.append((nb, face, r0[:, i], z1, np.full(len(nb), self.sub.round(vts[i, 2])),
                                 g, f1, f2, fz))

    def eval(self, enei):
        """initrefl2.m:17-152 for one wavelength."""
        sub, tab, n = self.sub, self.tab, self.n
        tabn = tab.normalised(enei)
        pos = self.pos
        x = pos[:, None, 0] - pos[None, :, 0]
        y = pos[:, None, 1] - pos[None, :, 1]
        r = np.sqrt(x * x + y * y)
        z1 = np.repeat(pos[:, 2:3], n, axis=1)
        z2 = np.repeat(pos[None, :, 2], n, axis=0)
        G, Fr, Fz, _, _ = tab.interp(tabn, r, z1, z2)
        G = G * self.area[None, None, :]
        rr = np.maximum(r, 1e-10)
        Gp = np.stack([Fr * x / rr, Fr * y / rr, Fz], axis=2) * self.area[None, None, None, :]
        # diagonal elements with the polar rule
        idg = np.nonzero(np.diag(self.ir) == 2)[0]
        if len(idg):
            qp, qw = O.quadpol_flat(self.verts, self.faces, self.area, idg, self.npol)
            for a, f in enumerate(idg):
                p1 = qp[a]
                xq = p1[:, 0] - pos[f, 0]
                yq = p1[:, 1] - pos[f, 1]
                rq = np.sqrt(xq * xq + yq * yq)
                g, fr, fz, rc, _ = tab.interp(tabn, rq, p1[:, 2], np.full(len(rq), pos[f, 2]))
                G[:, f, f] = g @ qw[a]
                rcc = np.maximum(rc, 1e-10)
                Gp[:, f, 0, f] = (fr * xq / rcc) @ qw[a]
                Gp[:, f, 1, f] = (fr * yq / rcc) @ qw[a]
                Gp[:, f, 2, f] = fz @ qw[a]
        # refined off-diagonal pairs: sum over the vertices of the face
        acc = {}
        for nb, face, r0, z1v, z2v, g, f1, f2, fz in self.ref:
            gv, frv, fzv, _, _ = tab.interp(tabn, r0, z1v, z2v)
            key = face
            if key not in acc:
                acc[key] = [nb, np.zeros((5, len(nb)), complex), np.zeros((5, 3, len(nb)), complex)]
            acc[key][1] += gv * g
            acc[key][2][:, 0] += frv * f1
            acc[key][2][:, 1] += frv * f2
            acc[key][2][:, 2] += fzv * fz
        for face, (nb, gs, gps) in acc.items():
            G[:, nb, face] = gs
            Gp[:, nb, :, face] = np.transpose(gps, (2, 0, 1))  # advanced indices first
        F = np.einsum("ic,kicj->kij", self.nvec, Gp)
        return {"G": G, "Gp": Gp, "F": F}


# --------------------------------------------------------------------------
# bemretlayer (BEM/@bemretlayer)
# --------------------------------------------------------------------------
def _mm(A, X):
    return np.einsum("ij,j...->i...", A, X)


class BemRetLayer:
    """Dense retarded solver above the substrate for one closed particle in
    the upper medium (inside eps1, outside the upper medium)."""

    def __init__(self, green, greflayer):
        self.g, self.gr = green, greflayer

    def matrices(self, enei, eps1):
        """G1, H1 and the structured G2, H2, with eps applied (initmat.m:29-55)."""
        sub = self.gr.sub
        e2, _ = sub.eps(enei)
        k0 = 2 * np.pi / enei
        m1 = self.g.eval(k0 * np.sqrt(complex(eps1)), ("G", "H1"))
        m2 = self.g.eval(k0 * np.sqrt(complex(e2)), ("G", "H2"))
        rf = self.gr.eval(enei)
        G2, H2 = {}, {}
        for i, nm in enumerate(COMPONENTS):
            base = nm in ("ss", "hh", "p")
            G2[nm] = (m2["G"] if base else 0) + rf["G"][i]
            H2[nm] = (m2["H2"] if base else 0) + rf["F"][i]
        return m1["G"], m1["H1"], G2, H2, e2

    def initmat(self, enei, eps1, mats=None):
        G1, H1, G2, H2, e2 = self.matrices(enei, eps1) if mats is None else mats
        return self.initmat_from(enei, eps1, e2, G1, H1, G2, H2)

    def initmat_from(self, enei, eps1, eps2, G1, H1, G2, H2):
        """initmat.m:29-95 for unique eps1 (inside) and eps2 (outside)."""
        k = 2 * np.pi / enei
        nvec = self.g.nvec
        G1e, H1e = eps1 * G1, eps1 * H1
        G2e = {nm: eps2 * G2[nm] for nm in COMPONENTS}
        H2e = {nm: eps2 * H2[nm] for nm in COMPONENTS}
        G1i = np.linalg.inv(G1)
        G2pi = np.linalg.inv(G2["p"])
        Sigma1 = H1 @ G1i
        Sigma1e = H1e @ G1i
        Sigma2p = H2["p"] @ G2pi
        L1 = G1e @ G1i
        L2p = G2e["p"] @ G2pi
        nperp = nvec[:, 2]
        npar = nvec.copy()
        npar[:, 2] = 0
        Gamma = np.linalg.inv(Sigma1 - Sigma2p)
        Gammapar = 1j * k * ((L1 - L2p) @ Gamma) * (npar @ npar.T)
        ss = L1 @ G2["ss"] - G2e["ss"]
        sh = L1 @ G2["sh"] - G2e["sh"]
        hh = L1 @ G2["hh"] - G2e["hh"]
        m11 = Sigma1e @ G2["ss"] - H2e["ss"] - 1j * k * (Gammapar @ ss + nperp[:, None] * sh)
        m12 = Sigma1e @ G2["sh"] - H2e["sh"] - 1j * k * (Gammapar @ sh + nperp[:, None] * hh)
        m21 = Sigma1 @ G2["hs"] - H2["hs"] - 1j * k * nperp[:, None] * ss
        m22 = Sigma1 @ G2["hh"] - H2["hh"] - 1j * k * nperp[:, None] * sh
        self.k, self.eps1, self.eps2, self.enei = k, eps1, eps2, enei
        self.nvec, self.npar = nvec, npar
        self.G1, self.H1, self.G2, self.H2, self.G2e, self.H2e = G1, H1, G2, H2, G2e, H2e
        self.G1i, self.G2pi, self.L1, self.L2p = G1i, G2pi, L1, L2p
        self.Sigma1, self.Sigma1e, self.Gamma = Sigma1, Sigma1e, Gamma
        self.M = np.block([[m11, m12], [m21, m22]])
        return self

    def excitation(self, exc):
        """private/excitation.m:17-29 (scalar eps1, eps2)."""
        n, k, nvec = self.g.n, self.k, self.nvec
        e1, e2 = self.eps1, self.eps2
        npol = next(v.shape[-1] for v in exc.values() if v is not None)
        z = lambda key, shp: exc.get(key) if exc.get(key) is not None else np.zeros(shp, complex)
        phi1, phi1p, phi2, phi2p = (z(kk, (n, npol)) for kk in ("phi1", "phi1p", "phi2", "phi2p"))
        a1, a1p, a2, a2p = (z(kk, (n, 3, npol)) for kk in ("a1", "a1p", "a2", "a2p"))
        phi = phi2 - phi1
        a = a2 - a1
        alpha = a2p - a1p - 1j * k * (nvec[:, :, None] * (e2 * phi2)[:, None, :]
                                      - nvec[:, :, None] * (e1 * phi1)[:, None, :])
        De = e2 * phi2p - e1 * phi1p - 1j * k * (np.einsum("ic,icp->ip", nvec, e2 * a2)
                                                 - np.einsum("ic,icp->ip", nvec, e1 * a1))
        return phi, a, alpha, De

    def solve(self, exc):
        """mldivide.m:31-76."""
        n, k, nvec, npar = self.g.n, self.k, self.nvec, self.npar
        phi, a, alpha, De = self.excitation(exc)
        inner = lambda v, X: np.einsum("ic,icp->ip", v, X)
        outer = lambda v, X: v[:, :, None] * X[:, None, :]
        aperp = a[:, 2]
        apar = a.copy()
        apar[:, 2] = 0
        L1, G2, G2e = self.L1, self.G2, self.G2e
        alpha = alpha - _mm(self.Sigma1, a) + 1j * k * outer(nvec, _mm(L1, phi))
        De = (De - _mm(self.Sigma1e, phi) + 1j * k * inner(nvec, _mm(L1, a))
              + 1j * k * inner(npar, _mm((L1 - self.L2p) @ self.Gamma, alpha)))
        alphaperp = alpha[:, 2]
        alphapar = alpha.copy()
        alphapar[:, 2] = 0
        xi = np.linalg.solve(self.M, np.vstack([De, alphaperp]))
        sig2, h2perp = xi[:n], xi[n:]
        h2par = _mm(self.G2pi @ self.Gamma, alphapar + 1j * k * outer(
            npar, _mm(L1 @ G2["ss"] - G2e["ss"], sig2) + _mm(L1 @ G2["sh"] - G2e["sh"], h2perp)))
        h2 = h2par.copy()
        h2[:, 2] += h2perp
        sig1 = _mm(self.G1i, _mm(G2["ss"], sig2) + _mm(G2["sh"], h2perp) + phi)
        h1perp = _mm(self.G1i, _mm(G2["hs"], sig2) + _mm(G2["hh"], h2perp) + aperp)
        h1 = _mm(self.G1i, _mm(G2["p"], h2par) + apar)
        h1[:, 2] += h1perp
        return {"sig1": sig1, "sig2": sig2, "h1": h1, "h2": h2}

    def afun(self, sig):
        """The boundary conditions the chain solves, [sig1, h1, sig2, h2] ->
        [phi, a, De, alpha]: continuity of phi and A with the outside
        potentials phi2 = Gss sig2 + Gsh h2z, A2par = Gp h2par, A2z = Ghs sig2
        + Ghh h2z, and the normal-derivative conditions with eps (the scalar
        equations of afun.m with the structured G2, H2).  initmat.m:78-79
        (m{1,1}) takes, in the n.A term of the De condition, L1 G2.sh - G2e.sh
        for the charge where the continuity of A has G2.hs; the operator here
        is the one the chain inverts, so De carries the difference of the two
        (zero when the upper medium is vacuum, where the tables give
        hs = eps_up sh = sh)."""
        k, nvec, e1 = self.k, self.nvec, self.eps1
        G1, H1, G2, H2, G2e, H2e = self.G1, self.H1, self.G2, self.H2, self.G2e, self.H2e
        s1, h1, s2, h2 = sig["sig1"], sig["h1"], sig["sig2"], sig["h2"]
        h2z = h2[:, 2]
        h2p = h2.copy()
        h2p[:, 2] = 0
        inner = lambda X: np.einsum("ic,icp->ip", nvec, X)
        outer = lambda X: nvec[:, :, None] * X[:, None, :]

        def vec(Mp, Mhs, Mhh):
            out = _mm(Mp, h2p)
            out[:, 2] += _mm(Mhs, s2) + _mm(Mhh, h2z)
            return out

        phi2 = _mm(G2["ss"], s2) + _mm(G2["sh"], h2z)
        phi2e = _mm(G2e["ss"], s2) + _mm(G2e["sh"], h2z)
        A2 = vec(G2["p"], G2["hs"], G2["hh"])
        A2e = vec(G2e["p"], G2e["hs"], G2e["hh"])
        dphi2e = _mm(H2e["ss"], s2) + _mm(H2e["sh"], h2z)
        dA2 = vec(H2["p"], H2["hs"], H2["hh"])
        phi = _mm(G1, s1) - phi2
        a = _mm(G1, h1) - A2
        alpha = _mm(H1, h1) - dA2 - 1j * k * outer(e1 * _mm(G1, s1) - phi2e)
        De = e1 * _mm(H1, s1) - dphi2e - 1j * k * inner(e1 * _mm(G1, h1) - A2e)
        De = De + 1j * k * nvec[:, 2:3] * _mm(e1 * (G2["hs"] - G2["sh"]) - (G2e["hs"] - G2e["sh"]), s2)
        return phi, a, De, alpha


# --------------------------------------------------------------------------
# plane wave above the substrate and cross sections
# --------------------------------------------------------------------------
def _phase_coeff(sub, enei, kpar, z1, ind1, z2, ind2, fresnel=False):
    """reflectionsubs' coefficient times g1 g2 for points z1 (medium ind1)
    and z2 (medium ind2), elementwise; fresnel=True multiplies by k2z / k1z
    of fresnel.m (with its 1e-10 i offset)."""
    k0, k1, k2 = sub.k(enei)
    e1, e2 = sub.eps(enei)
    ks = np.array([k1, k2])
    kz = np.array([kz_of(k1, kpar), kz_of(k2, kpar)])
    c = coefficients(k0, e1, e2, kpar, kz[0], kz[1])
    g1 = np.exp(1j * kz[ind1 - 1] * np.abs(np.asarray(z1) - sub.z0))
    g2 = np.exp(1j * kz[ind2 - 1] * np.abs(np.asarray(z2) - sub.z0))
    out = {nm: c[nm][ind1 - 1, ind2 - 1] * g1 * g2 for nm in COMPONENTS}
    if fresnel:
        kzo = np.array([kz_of(k1, kpar, offset=True), kz_of(k2, kpar, offset=True)])
        ratio = kzo[ind2 - 1] / kzo[ind1 - 1]
        out = {nm: v * ratio for nm, v in out.items()}
    return out, kz, ks


def planewave_potential(sub, pos, nvec, pol, dirs, enei, eps_particle_out=None, derivative="analytic"):
    """Outside potentials of plane waves above the substrate at the faces
    (potential.m:20-105; every face lies above the interface and its outside
    is the upper medium).  derivative 'analytic' differentiates the phase
    factors; 'fd' restates the reference's finite difference with eta 1e-8.
    Returns phi2, phi2p (n, npol), a2, a2p (n, 3, npol)."""
    pol = np.atleast_2d(np.asarray(pol, float))
    dirs = np.atleast_2d(np.asarray(dirs, float))
    n = len(pos)
    k0 = 2 * np.pi / enei
    e1, e2 = sub.eps(enei)
    phi = np.zeros((n, len(dirs)), complex)
    phip = np.zeros_like(phi)
    a = np.zeros((n, 3, len(dirs)), complex)
    ap = np.zeros_like(a)
    eta = 1e-8
    for i, (pl, d) in enumerate(zip(pol, dirs)):
        if d[2] < 0:
            medium, z2, ind2 = 1, sub.z0 + 1e-10, 1
        else:
            medium, z2, ind2 = 2, sub.z0 - 1e-10, 2
        nb = np.sqrt(e1 if medium == 1 else e2)
        kpar = nb * k0 * np.linalg.norm(d[:2])

        def pot(p):
            r, kz, _ = _phase_coeff(sub, enei, kpar, p[:, 2], 1, np.full(len(p), z2), ind2, fresnel=True)
            inn = p[:, :2] @ d[:2] + z2 * d[2]
            fac = np.exp(1j * k0 * nb * inn)[:, None] / (1j * k0) * pl[None, :]
            av = np.empty((len(p), 3), complex)
            av[:, :2] = fac[:, :2] * r["p"][:, None]
            av[:, 2] = fac[:, 2] * r["hh"]
            return av, fac[:, 2] * r["sh"], kz

        av, ph, kz = pot(pos)
        if derivative == "fd":
            av2, ph2, _ = pot(pos + eta * nvec)
            avp, php = (av2 - av) / eta, (ph2 - ph) / eta
        else:
            # d/dn of exp(i k0 nb xy.dir_xy) exp(i k1z (z - z0)) (z above z0)
            dph = 1j * (k0 * nb * (nvec[:, :2] @ d[:2]) + kz[0] * nvec[:, 2])
            avp, php = av * dph[:, None], ph * dph
        a[:, :, i], phi[:, i] = av, ph
        ap[:, :, i], phip[:, i] = avp, php
        if medium == 1:
            a0 = np.exp(1j * k0 * nb * (pos @ d))[:, None] / (1j * k0) * pl[None, :]
            a[:, :, i] += a0
            ap[:, :, i] += a0 * (1j * k0 * nb * (nvec @ d))[:, None]
    return phi, phip, a, ap


def farfield(sub, pos, area, sig, dirs, enei):
    """spectrumretlayer/farfield.m:20-104 for a particle above the interface:
    e, h (ndir, 3, npol) of the outside charges and currents, direct and
    reflected in the upper half space, transmitted below (only through a
    lossless medium, else zero)."""
    k0, k1, k2 = sub.k(enei)
    dirs = np.atleast_2d(np.asarray(dirs, float))
    s2, h2 = sig["sig2"], sig["h2"]
    npol = s2.shape[1]
    a = np.zeros((len(dirs), 3, npol), complex)
    e = np.zeros_like(a)
    for i, d in enumerate(dirs):
        up = d[2] >= 0
        km = k1 if up else k2
        if np.imag(km) != 0:
            continue
        kpar = km * np.sqrt(1 - d[2] ** 2)
        r, _, _ = _phase_coeff(sub, enei, kpar, np.full(len(pos), sub.z0), 1 if up else 2, pos[:, 2], 1)
        dist = pos[:, :2] @ d[:2] + sub.z0 * d[2]
        phase = np.exp(-1j * km * dist) * area
        a[i, 0] = (r["p"] * phase) @ h2[:, 0]
        a[i, 1] = (r["p"] * phase) @ h2[:, 1]
        a[i, 2] = (r["hh"] * phase) @ h2[:, 2] + (r["hs"] * phase) @ s2
    e = 1j * k0 * a
    upm = dirs[:, 2] >= 0
    if np.any(upm):
        ed, _, _ = O.farfield(pos, area, s2, h2, dirs[upm], enei, sub.eps(enei)[0])
        e[upm] += ed
    e = e - dirs[:, :, None] * np.einsum("dc,dcp->dp", dirs, e)[:, None, :]
    kk = np.where(upm, k1, k2)
    h = (kk / k0)[:, None, None] * np.cross(dirs[:, :, None], e, axisa=1, axisb=1, axisc=1)
    return e, h


def efresnel(sub, pol, dirs, enei):
    """efresnel.m:19-66: reflected and transmitted fields and wavevectors."""
    k0, k1, k2 = sub.k(enei)
    ks = [k1, k2]
    z1, z2 = sub.z0 + 1e-10, sub.z0 - 1e-10
    out = {key: np.zeros((len(dirs), 3), complex) for key in ("er", "et", "kr", "kt")}
    for i, (pl, d) in enumerate(zip(pol, dirs)):
        if d[2] < 0:
            posr = (z1, 1, z1, 1)
            post = (z2, 2, z1, 1)
        else:
            posr = (z2, 2, z2, 2)
            post = (z1, 1, z2, 2)
        kparv = ks[post[3] - 1] * d[:2]
        kp2 = np.dot(kparv, kparv)
        kzr = np.sqrt(ks[posr[1] - 1] ** 2 - kp2 + 0j)
        kzr = kzr * np.sign(np.imag(kzr + 1e-10j))
        kzt = np.sqrt(ks[post[1] - 1] ** 2 - kp2 + 0j)
        kzt = kzt * np.sign(np.imag(kzt + 1e-10j))
        kr = np.array([kparv[0], kparv[1], -np.sign(d[2]) * kzr])
        kt = np.array([kparv[0], kparv[1], np.sign(d[2]) * kzt])
        kpn = np.sqrt(kp2 + 0j)
        r, _, _ = _phase_coeff(sub, enei, kpn, posr[0], posr[1], posr[2], posr[3], fresnel=True)
        t, _, _ = _phase_coeff(sub, enei, kpn, post[0], post[1], post[2], post[3], fresnel=True)
        er = np.array([r["p"] * pl[0], r["p"] * pl[1], r["hh"] * pl[2]]) - kr / k0 * r["sh"] * pl[2]
        et = np.array([t["p"] * pl[0], t["p"] * pl[1], t["hh"] * pl[2]]) - kt / k0 * t["sh"] * pl[2]
        er = np.exp(-1j * kr[2] * posr[2] - 1j * kr[2] * posr[0]) * er
        et = np.exp(-1j * kr[2] * post[2] - 1j * kt[2] * post[0]) * et
        out["er"][i], out["et"][i], out["kr"][i], out["kt"][i] = er, et, kr, kt
    return out


def extinction(sub, pos, area, sig, pol, dirs, enei):
    """planewaveretlayer/extinction.m:17-48 (Matlab dot conjugates its first
    argument)."""
    pol = np.atleast_2d(np.asarray(pol, float))
    dirs = np.atleast_2d(np.asarray(dirs, float))
    ef = efresnel(sub, pol, dirs, enei)
    ext = np.zeros(len(dirs))
    for i in range(len(dirs)):
        kr, kt = ef["kr"][i], ef["kt"][i]
        knr = np.linalg.norm(kr)
        esr = farfield(sub, pos, area, sig, (kr / knr).real[None, :], enei)[0][0, :, i]
        if abs(np.imag(kt[2])) > 1e-10:
            est = np.zeros(3, complex)
        else:
            est = farfield(sub, pos, area, sig, (kt / np.linalg.norm(kt)).real[None, :], enei)[0][0, :, i]
        extr = 4 * np.pi / knr * np.imag(np.vdot(ef["er"][i], esr))
        extt = 4 * np.pi / knr * np.imag(np.vdot(ef["et"][i], est))
        ext[i] = extr + extt
    return ext


def scattering(sub, pos, area, sig, dirs, enei, pinf_nvec, pinf_area):
    """planewaveretlayer/scattering.m:13-24 over the sphere pinfty."""
    e, h = farfield(sub, pos, area, sig, pinf_nvec, enei)
    s = np.cross(e, np.conj(h), axisa=1, axisb=1, axisc=1)
    dsca = 0.5 * np.real(np.einsum("dc,dcp->dp", pinf_nvec, s))
    sca = pinf_area @ dsca
    e1, e2 = sub.eps(enei)
    dz = np.atleast_2d(dirs)[:, 2]
    fac = np.where(dz < 0, 0.5 * np.sqrt(e1 + 0j), np.where(dz > 0, 0.5 * np.sqrt(e2 + 0j), 1.0))
    return np.real(sca / fac)


# --------------------------------------------------------------------------
# points above the interface: compgreenretlayer(pt, p) (potential.m, field.m)
# --------------------------------------------------------------------------
class PointGreenLayer(GreenRetLayer):
    """Reflected G (5, npt, n) and Gp (5, npt, 3, n) from the faces of the
    particle to points above the interface (greenretlayer with p1 a
    compoint: refinematrixlayer.m divides by the face radius, as compoint has
    none; init2.m with p1 the points; no diagonal elements)."""

    def __init__(self, sub, pts, verts, faces, table, relcutoff=3.0, abscutoff=0.0):
        self.pts = np.atleast_2d(np.asarray(pts, dtype=float))
        self.sub, self.tab = sub, table
        self.verts = np.asarray(verts, dtype=float)
        self.faces = np.asarray(faces, dtype=np.int64)
        _, _, _, self.nvec, self.area = O.norm_flat(self.verts, self.faces)
        self.fpos = O.norm_flat(self.verts, self.faces)[0]
        self.n = len(self.faces)
        rad = O.bradius(self.verts, self.faces, self.fpos)
        r = O.pdist2(self.pts[:, :2], self.fpos[:, :2])
        z = sub.mindist(self.pts[:, 2])[:, None] + sub.mindist(self.fpos[:, 2])[None, :]
        d2 = np.sqrt(r ** 2 + z ** 2) - rad[None, :]
        self.ir = ((d2 < abscutoff) | (d2 / rad[None, :] < relcutoff)).astype(np.int8)
        self.pos = self.pts  # rows of the refinement weights
        self._init_refine()

    def eval(self, enei):
        tab = self.tab
        tabn = tab.normalised(enei)
        p1, p2 = self.pts, self.fpos
        x = p1[:, None, 0] - p2[None, :, 0]
        y = p1[:, None, 1] - p2[None, :, 1]
        r = np.sqrt(x * x + y * y)
        z1 = np.repeat(p1[:, 2:3], self.n, axis=1)
        z2 = np.repeat(p2[None, :, 2], len(p1), axis=0)
        G, Fr, Fz, _, _ = tab.interp(tabn, r, z1, z2)
        G = G * self.area[None, None, :]
        rr = np.maximum(r, 1e-10)
        Gp = np.stack([Fr * x / rr, Fr * y / rr, Fz], axis=2) * self.area[None, None, None, :]
        acc = {}
        for nb, face, r0, z1v, z2v, g, f1, f2, fz in self.ref:
            gv, frv, fzv, _, _ = tab.interp(tabn, r0, z1v, z2v)
            if face not in acc:
                acc[face] = [nb, np.zeros((5, len(nb)), complex), np.zeros((5, 3, len(nb)), complex)]
            acc[face][1] += gv * g
            acc[face][2][:, 0] += frv * f1
            acc[face][2][:, 1] += frv * f2
            acc[face][2][:, 2] += fzv * fz
        for face, (nb, gs, gps) in acc.items():
            G[:, nb, face] = gs
            Gp[:, nb, :, face] = np.transpose(gps, (2, 0, 1))
        return {"G": G, "Gp": Gp}


def points_reflected(pg, enei, sig):
    """Reflected potentials and fields at the points from the outside charges
    and currents (compgreenretlayer/potential.m, field.m with matmul2.m,
    matmul3.m): phi = Gss s2 + Gsh h2z, a = [Gp h2x, Gp h2y, Ghh h2z + Ghs s2],
    e = i k0 a - (Gp_ss s2 + Gp_sh h2z), h = curl of the structured product."""
    k0 = 2 * np.pi / enei
    m = pg.eval(enei)
    G, Gp = m["G"], m["Gp"]
    ic = {nm: i for i, nm in enumerate(COMPONENTS)}
    s2, h2 = sig["sig2"], sig["h2"]
    mm = lambda A, X: np.einsum("ij,jq->iq", A, X)
    phi = mm(G[ic["ss"]], s2) + mm(G[ic["sh"]], h2[:, 2])
    a = np.stack([mm(G[ic["p"]], h2[:, 0]), mm(G[ic["p"]], h2[:, 1]),
                  mm(G[ic["hh"]], h2[:, 2]) + mm(G[ic["hs"]], s2)], axis=1)
    e = 1j * k0 * a - np.stack([mm(Gp[ic["ss"]][:, c], s2) + mm(Gp[ic["sh"]][:, c], h2[:, 2]) for c in range(3)],
                               axis=1)

    def M(i1, i2):
        if i2 < 2:
            return mm(Gp[ic["p"]][:, i1], h2[:, i2])
        return mm(Gp[ic["hh"]][:, i1], h2[:, 2]) + mm(Gp[ic["hs"]][:, i1], s2)

    h = np.stack([M(1, 2) - M(2, 1), M(2, 0) - M(0, 2), M(0, 1) - M(1, 0)], axis=1)
    return {"phi": phi, "a": a, "e": e, "h": h}


# --------------------------------------------------------------------------
# dipoles above the interface (Simulation/retarded/@dipoleretlayer)
# --------------------------------------------------------------------------
def greenderiv(table, tabn, pos1, pos2):
    """@dipoleretlayer/private/greenderiv.m:15-99: reflected Green functions
    G (5, n1, n2) between observers pos1 and dipoles pos2 and their
    derivatives F[(a, b)] (5, n1, n2): a = 2, 3, 4 differentiates in x, y, z
    at the observer, b = 2, 3, 4 in x', y', z' at the dipole (1: none).
    Second radial derivatives and the z' derivatives are finite differences
    with eta = 1e-6 of the interpolated table; equal point sets move the
    dipoles by rmin along x (the self-interaction)."""
    sub = table.sub
    pos1, pos2 = np.atleast_2d(pos1), np.atleast_2d(pos2)
    if pos1.shape == pos2.shape and np.all(pos1 == pos2):
        pos2 = pos2 + np.array([sub.rmin, 0.0, 0.0])
    x = pos1[:, 0, None] - pos2[None, :, 0]
    y = pos1[:, 1, None] - pos2[None, :, 1]
    r = np.sqrt(x ** 2 + y ** 2)
    z1 = np.broadcast_to(pos1[:, 2, None], r.shape)
    z2 = np.broadcast_to(pos2[None, :, 2], r.shape)
    eta = 1e-6
    r = np.where(r < sub.rmin, sub.rmin, r)
    ux, uy = x / r, y / r
    G, Fr, F1 = table.interp(tabn, r, z1, z2)[:3]
    _, fr, fz = table.interp(tabn, r + eta, z1, z2)[:3]
    Frr, Fr1 = (fr - Fr) / eta, (fz - F1) / eta
    g, fr, fz = table.interp(tabn, r, z1, z2 + eta)[:3]
    F2, Fr2, F12 = (g - G) / eta, (fr - Fr) / eta, (fz - F1) / eta
    F = {(1, 2): -Fr * ux, (1, 3): -Fr * uy, (2, 1): Fr * ux, (3, 1): Fr * uy, (4, 1): F1, (1, 4): F2,
         (2, 2): -Fr * uy ** 2 / r - Frr * ux ** 2,
         (2, 3): -Fr * ux * uy / r - Frr * ux * uy,
         (3, 3): -Fr * ux ** 2 / r - Frr * uy ** 2,
         (2, 4): Fr2 * ux, (3, 4): Fr2 * uy, (4, 2): -Fr1 * ux, (4, 3): -Fr1 * uy, (4, 4): F12}
    F[(3, 2)] = F[(2, 3)]
    return G, F


def _dip_moments(pts, dip, eps_pt):
    """dip (npt, 3, ndip) and dip / eps of the point's medium."""
    pts = np.atleast_2d(pts)
    dip = np.asarray(dip, dtype=float).reshape(len(pts), 3, -1)
    return dip, dip / np.asarray(eps_pt, dtype=complex).reshape(-1, 1, 1)


def _cols(v):
    """(..., npt, ndip) -> (..., npt ndip), column ipt + npt idip."""
    return v.reshape(v.shape[:-2] + (-1,), order="F")


def dipole_layer_potential(table, tabn, enei, pos, nvec, pts, dip, eps_pt):
    """Reflected part of the dipole potentials at the faces pos (normals
    nvec), @dipoleretlayer/potential.m:29-96: phi2, phi2p (n, npt ndip),
    a2, a2p (n, 3, npt ndip); the caller adds dipoleret's direct part.  The
    hh term of a2p(:, 3) takes the y component of the moment, as
    potential.m:82 does."""
    k0 = 2 * np.pi / enei
    ik = -1j * k0
    dp, d2 = _dip_moments(pts, dip, eps_pt)
    G, F = greenderiv(table, tabn, pos, pts)
    c = {nm: i for i, nm in enumerate(COMPONENTS)}
    f = lambda g, d: g[:, :, None] * d[None, :, :]
    dd = lambda nm, b: (F[(2, b)][c[nm]] * nvec[:, 0, None] + F[(3, b)][c[nm]] * nvec[:, 1, None]
                        + F[(4, b)][c[nm]] * nvec[:, 2, None])
    hs = lambda M, b: M[(1, b)][c["hs"]]
    a = np.stack([ik * f(G[c["p"]], dp[:, 0]), ik * f(G[c["p"]], dp[:, 1]),
                  ik * f(G[c["hh"]], dp[:, 2]) + f(hs(F, 2), d2[:, 0]) + f(hs(F, 3), d2[:, 1])
                  + f(hs(F, 4), d2[:, 2])], axis=1)
    phi = (f(F[(1, 2)][c["ss"]], d2[:, 0]) + f(F[(1, 3)][c["ss"]], d2[:, 1]) + f(F[(1, 4)][c["ss"]], d2[:, 2])
           + ik * f(G[c["sh"]], dp[:, 2]))
    ap = np.stack([ik * f(dd("p", 1), dp[:, 0]), ik * f(dd("p", 1), dp[:, 1]),
                   ik * f(dd("hh", 1), dp[:, 1]) + f(dd("hs", 2), d2[:, 0]) + f(dd("hs", 3), d2[:, 1])
                   + f(dd("hs", 4), d2[:, 2])], axis=1)
    phip = (f(dd("ss", 2), d2[:, 0]) + f(dd("ss", 3), d2[:, 1]) + f(dd("ss", 4), d2[:, 2])
            + ik * f(dd("sh", 1), dp[:, 2]))
    return _cols(phi), _cols(phip), _cols(a), _cols(ap)


def dipole_layer_field(table, tabn, enei, obs, pts, dip, eps_pt):
    """Reflected field of the dipoles at the points obs,
    @dipoleretlayer/field.m:45-128 with key 'refl': e, h (nobs, 3, npt ndip),
    e = i k0 a - grad phi and h = -(the curl of a) as field.m:122-128 writes
    them (phi3 takes the z' derivative of sh, as field.m:104 does)."""
    k0 = 2 * np.pi / enei
    ik = -1j * k0
    dp, d2 = _dip_moments(pts, dip, eps_pt)
    G, F = greenderiv(table, tabn, obs, pts)
    c = {nm: i for i, nm in enumerate(COMPONENTS)}
    f = lambda g, d: g[:, :, None] * d[None, :, :]
    Fc = lambda a, b, nm: F[(a, b)][c[nm]]

    def hsum(a, nm):  # F{a,2}.nm d2x + F{a,3}.nm d2y + F{a,4}.nm d2z
        return f(Fc(a, 2, nm), d2[:, 0]) + f(Fc(a, 3, nm), d2[:, 1]) + f(Fc(a, 4, nm), d2[:, 2])

    a1 = ik * f(G[c["p"]], dp[:, 0])
    a2 = ik * f(G[c["p"]], dp[:, 1])
    a3 = ik * f(G[c["hh"]], dp[:, 2]) + hsum(1, "hs")
    a23 = ik * f(Fc(3, 1, "hh"), dp[:, 2]) + hsum(3, "hs")
    a32 = ik * f(Fc(4, 1, "p"), dp[:, 1])
    a31 = ik * f(Fc(4, 1, "p"), dp[:, 0])
    a13 = ik * f(Fc(4, 1, "hh"), dp[:, 2]) + hsum(4, "hs")
    a12 = ik * f(Fc(2, 1, "p"), dp[:, 1])
    a21 = ik * f(Fc(3, 1, "p"), dp[:, 0])
    phi1 = hsum(2, "ss") + ik * f(Fc(2, 1, "sh"), dp[:, 2])
    phi2 = hsum(3, "ss") + ik * f(Fc(3, 1, "sh"), dp[:, 2])
    phi3 = hsum(4, "ss") + ik * f(Fc(1, 4, "sh"), dp[:, 2])
    e = 1j * k0 * np.stack([a1, a2, a3], axis=1) - np.stack([phi1, phi2, phi3], axis=1)
    h = -np.stack([a23 - a32, a31 - a13, a12 - a21], axis=1)
    return _cols(e), _cols(h)


def dipole_layer_farfield(sub, enei, pts, dip, eps_pt, dirs):
    """Far field of dipoles above the interface, @dipoleretlayer/
    farfield.m:21-177: e, h (ndir, 3, npt ndip).  Upward directions: direct
    plus reflected; downward: transmitted (each only through a lossless
    medium); the reflected and transmitted parts take the gradient at the
    dipole by forward differences with eta = 1e-6 (farfield.m:41-50)."""
    k0, k1, k2 = sub.k(enei)
    dirs = np.atleast_2d(np.asarray(dirs, float))
    dp, d2 = _dip_moments(pts, dip, eps_pt)
    pts = np.atleast_2d(pts)
    npt, _, ndip = dp.shape
    eta = 1e-6
    p4 = np.concatenate([pts[:, None, :], pts[:, None, :] + eta * np.eye(3)[None]], axis=1)  # (npt, 4, 3)
    a = np.zeros((len(dirs), 3, npt, ndip), complex)
    for i, d in enumerate(dirs):
        up = d[2] >= 0
        km = k1 if up else k2
        if np.imag(km) != 0:
            continue
        kpar = km * np.sqrt(1 - d[2] ** 2)
        z2 = p4[:, :, 2].ravel()
        r, _, _ = _phase_coeff(sub, enei, kpar, np.full(z2.shape, sub.z0), 1 if up else 2, z2, 1)
        dist = p4[:, :, :2].reshape(-1, 2) @ d[:2] + sub.z0 * d[2]
        phase = np.exp(-1j * km * dist)
        rr = {}
        for nm in ("p", "hh", "hs"):
            v = (r[nm] * phase).reshape(npt, 4)
            v[:, 1:] = (v[:, 1:] - v[:, :1]) / eta
            rr[nm] = v
        a[i, 0] = -1j * k0 * rr["p"][:, 0, None] * dp[:, 0]
        a[i, 1] = -1j * k0 * rr["p"][:, 0, None] * dp[:, 1]
        a[i, 2] = (-1j * k0 * rr["hh"][:, 0, None] * dp[:, 2] + rr["hs"][:, 1, None] * d2[:, 0]
                   + rr["hs"][:, 2, None] * d2[:, 1] + rr["hs"][:, 3, None] * d2[:, 2])
    e = 1j * k0 * a
    upm = dirs[:, 2] >= 0
    if np.any(upm):
        from . import dipole_oracle as DO
        ed, _ = DO.dipole_farfield(dirs[upm], pts, dp, enei, sub.eps(enei)[0])
        e[upm] += ed.reshape(int(upm.sum()), 3, npt, ndip, order="F")
    e = e - dirs[:, :, None, None] * np.einsum("dc,dcpq->dpq", dirs, e)[:, None]
    kk = np.where(upm, k1, k2)
    h = (kk / k0)[:, None, None, None] * np.cross(dirs[:, :, None, None], e, axisa=1, axisb=1, axisc=1)
    return _cols(e), _cols(h)


def dipole_layer_decayrate(table, tabn, sub, enei, pts, dip, e_sig, e_sig_far, h_sig_far, pinfty_nvec,
                           pinfty_area):
    """@dipoleretlayer/decayrate.m:14-60: (tot, rad, rad0) (npt, ndip) for
    dipoles above the interface in the upper medium.  e_sig (npt, 3,
    npt ndip) is the field of the surface charges at the dipoles (direct and
    reflected, compgreenretlayer field.m), e_sig_far/h_sig_far their far
    fields on the pinfty normals (spectrumretlayer farfield.m); zero for no
    particle."""
    pts = np.atleast_2d(pts)
    eu = sub.eps(enei)[0]
    dp, _ = _dip_moments(pts, dip, np.full(len(pts), eu))
    npt, _, ndip = dp.shape
    k0 = 2 * np.pi / enei
    nb = np.sqrt(eu)
    gamma = 4 / 3 * k0 ** 3
    er, _ = dipole_layer_field(table, tabn, enei, pts, pts, dp, np.full(npt, eu))
    e = e_sig + er
    ed, hd = dipole_layer_farfield(sub, enei, pts, dp, np.full(npt, eu), pinfty_nvec)
    from . import dipole_oracle as DO
    sca = DO.radiated_power(e_sig_far + ed, h_sig_far + hd, pinfty_nvec, pinfty_area)
    tot = np.zeros((npt, ndip))
    rad = np.zeros((npt, ndip))
    for p in range(npt):
        for i in range(ndip):
            q = p + npt * i
            tot[p, i] = np.real(1 + np.imag(e[p, :, q] @ dp[p, :, i]) / (0.5 * nb * gamma))
            rad[p, i] = np.real(sca[q] / (2 * np.pi * k0) / (0.5 * nb * gamma))
    return tot, rad, np.real(nb * gamma) * np.ones((npt, ndip))
