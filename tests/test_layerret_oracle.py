"""CPU pins of the oracle's retarded substrate restatement
(oracle/layerret_oracle.py): the Sommerfeld integrals against the closed form
of the Sommerfeld identity and against brute-force real-axis integration, the
perfect-conductor and quasistatic limits of the coefficients, the table
interpolation against direct integration, the chain against the boundary
operator it inverts, the reduction to the free-space solver and observables
for equal media, and a small sphere against the quasistatic substrate."""
import numpy as np
import pytest
from scipy import integrate, special as sp

from mnpbem_amd import shapes
from oracle import layer_oracle as LS
from oracle import layerret_oracle as L
from oracle import mnpbem_oracle as O

DRUDE = (10.0, 9.065, 0.0708)
POL = np.array([[1.0, 0, 0], [0.6, 0, 0.8]])
DIR = np.array([[0, 0, -1.0], [0.8, 0, -0.6]])


def _one(kp):
    return np.ones((1,) + np.shape(kp), complex)


@pytest.mark.parametrize("eps_up", [1.0, 1.77])
def test_sommerfeld_identity(eps_up):
    """Coefficient 1: G = exp(i k D) / D, Fr, Fz its r and z derivatives, on
    both the real-axis leg (Z >= r / 2) and the Hankel leg."""
    sub = L.Substrate(eps_up, 2.25, 0.0)
    for enei in (400.0, 1000.0):
        k1 = sub.k(enei)[1]
        for r, Z in ((0.01, 0.03), (0.3, 1.0), (4.0, 3.0), (10.0, 2.0), (40.0, 1.0), (150.0, 80.0)):
            G, Fr, Fz = L.reflected_green(sub, enei, r, Z, coeff=_one)
            D = np.hypot(r, Z)
            g = np.exp(1j * k1 * D) / D
            dg = (1j * k1 - 1 / D) * g
            assert abs(G[0] / g - 1) < 1e-11
            assert abs(Fr[0] - dg * r / D) < 1e-11 * abs(dg)
            assert abs(Fz[0] - dg * Z / D) < 1e-11 * abs(dg)


def _brute(sub, enei, r, Z, comp, which):
    k0, k1, k2 = sub.k(enei)
    e1, e2 = sub.eps(enei)

    def f(x):
        kp = complex(x)
        k1z = L.kz_of(k1, kp)
        R = L.coefficients(k0, e1, e2, kp)[comp][0, 0] * np.exp(1j * k1z * Z)
        if which == 0:
            return 1j * sp.jv(0, x * r) * R * kp / k1z
        if which == 1:
            return 1j * sp.jv(1, x * r) * R * (-kp * kp / k1z)
        return -sp.jv(0, x * r) * R * kp

    pts = sorted({abs(k1), abs(k2.real)})
    kw = dict(points=pts, limit=4000, epsabs=1e-13, epsrel=1e-11)
    re = integrate.quad(lambda x: f(x).real, 0, 45.0 / Z, **kw)[0]
    im = integrate.quad(lambda x: f(x).imag, 0, 45.0 / Z, **kw)[0]
    return re + 1j * im


@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
@pytest.mark.parametrize("eps_down", [2.25, -10 + 1.2j])
def test_complex_path_equals_real_axis(eps_down):
    """The deformed path (semi-ellipse, real axis or Hankel leg) gives the
    integral along the real axis, for a dielectric and a lossy metal substrate."""
    sub = L.Substrate(1.0, eps_down, 0.0)
    for r, Z in ((1.0, 3.0), (10.0, 3.0)):
        out = L.reflected_green(sub, 600.0, r, Z)
        for comp in ("p", "ss", "hs", "hh"):
            ic = L.COMPONENTS.index(comp)
            for w in range(3):
                b = _brute(sub, 600.0, r, Z, comp, w)
                assert abs(out[w][ic] - b) < 1e-8 * abs(b) + 1e-14


def test_coefficient_limits():
    """A perfect conductor below reflects p with -1 (image of opposite sign);
    at long wavelengths ss is the image charge (eps1 - eps2) / (eps1 + eps2) / D
    and p, hh, sh, hs vanish; hs = eps_up sh above the interface."""
    r, Z = 3.0, 4.0
    D = np.hypot(r, Z)
    pec = L.Substrate(1.0, -1e12, 0.0)
    G, _, _ = L.reflected_green(pec, 600.0, r, Z)
    k1 = pec.k(600.0)[1]
    assert abs(G[0] / (-np.exp(1j * k1 * D) / D) - 1) < 1e-3
    qs = L.Substrate(1.77, 2.25, 0.0)
    G, _, Fz = L.reflected_green(qs, 1e8, r, Z)  # k D ~ 4e-7
    f = (1.77 - 2.25) / (1.77 + 2.25)
    assert abs(G[1] * D / f - 1) < 1e-6
    assert abs(Fz[1] / (-f * Z / D ** 3) - 1) < 1e-6
    for c in (0, 2, 3, 4):
        assert abs(G[c]) < 1e-5 * abs(G[1])
    G, Fr, Fz = L.reflected_green(qs, 600.0, r, Z)
    for a in (G, Fr, Fz):
        assert abs(a[2] - 1.77 * a[3]) < 1e-12 * abs(a[2])


def test_table_interpolation_between_grid_points():
    """Bilinear interpolation of the normalised table at cell midpoints
    reproduces direct integration there to the grid's discretisation error."""
    sub = L.Substrate(1.0, 2.25, 0.0)
    v, _ = shapes.icosphere(1, 5.0)
    v = v + [0, 0, 6.0]
    rg, zg, z2 = L.tab_grid(sub, v, nr=16, nz=16)
    tab = L.GreenTable(sub, rg, zg, z2)
    tabn = tab.normalised(550.0)
    for i, j in ((2, 3), (7, 9), (12, 5)):
        r = 0.5 * (rg[i] + rg[i + 1])
        z1 = 0.5 * (zg[j] + zg[j + 1]) - (z2 - sub.z0)
        G, Fr, Fz, _, _ = tab.interp(tabn, np.array([r]), np.array([z1]), np.array([z2]))
        Gd, Frd, Fzd = L.reflected_green(sub, 550.0, r, (z1 - sub.z0) + (z2 - sub.z0))
        for a, b in ((G[:, 0], Gd), (Fr[:, 0], Frd), (Fz[:, 0], Fzd)):
            assert np.max(np.abs(a - b)) < 2e-2 * np.max(np.abs(b))
    with pytest.raises(ValueError):
        tab.interp(tabn, np.array([2 * rg[-1]]), np.array([zg[0]]), np.array([z2]))


def test_chain_inverts_the_boundary_operator():
    """mldivide's solution satisfies afun for random structured matrices:
    continuity of phi and A, and the two normal-derivative conditions in the
    form of initmat.m (the charge term of n.A with G2.sh)."""
    rng = np.random.default_rng(3)
    n = 12

    class G:
        pass

    g = G()
    g.n = n
    nv = rng.normal(size=(n, 3))
    g.nvec = nv / np.linalg.norm(nv, axis=1)[:, None]
    R = lambda: rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    G2 = {nm: R() + (5 * np.eye(n) if nm in ("p", "ss", "hh") else 0) for nm in L.COMPONENTS}
    H2 = {nm: R() for nm in L.COMPONENTS}
    b = L.BemRetLayer(g, None).initmat_from(500.0, 2.0 + 0.3j, 1.5, R() + 5 * np.eye(n), R(), G2, H2)
    exc = {kk: rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2)) for kk in ("phi1", "phi1p", "phi2", "phi2p")}
    exc.update({kk: rng.normal(size=(n, 3, 2)) + 1j * rng.normal(size=(n, 3, 2))
                for kk in ("a1", "a1p", "a2", "a2p")})
    sig = b.solve(exc)
    phi, a, alpha, De = b.excitation(exc)
    for x, want in zip(b.afun(sig), (phi, a, De, alpha)):
        assert np.abs(x - want).max() < 1e-12 * np.abs(want).max()


@pytest.fixture(scope="module")
def sphere():
    v, f = shapes.icosphere(2, 5.0)
    v = v + [0, 0, 6.0]  # 1 nm above the interface z = 0
    return v, f, O.GreenRet(v, f)


def _solve(sub, v, f, g, lam, nr=20):
    rg, zg, z2 = L.tab_grid(sub, v, nr=nr, nz=nr)
    gr = L.GreenRetLayer(sub, v, f, L.GreenTable(sub, rg, zg, z2))
    bem = L.BemRetLayer(g, gr).initmat(lam, O.eps_drude(lam, *DRUDE))
    phi, phip, a, ap = L.planewave_potential(sub, g.pos, g.nvec, POL, DIR, lam)
    return bem.solve({"phi2": phi, "phi2p": phip, "a2": a, "a2p": ap})


def test_equal_media_is_the_free_particle(sphere, pinfty):
    """eps_down = eps_up: no reflection; the solver, the plane wave and the
    cross sections are those of the particle in the homogeneous medium."""
    v, f, g = sphere
    sub = L.Substrate(1.33 ** 2, 1.33 ** 2, 0.0)
    lam = 560.0
    sig = _solve(sub, v, f, g, lam, nr=6)
    b0 = O.BemRet(g).initmat(lam, O.eps_drude(lam, *DRUDE), 1.33 ** 2)
    a2, a2p = O.planewave_potential(g.pos, g.nvec, POL, DIR, lam, 1.33 ** 2)
    ref = b0.solve({"a2": a2, "a2p": a2p})
    for k in ref:
        assert np.abs(sig[k] - ref[k]).max() < 1e-11 * np.abs(ref[k]).max()
    phi, phip, a, ap = L.planewave_potential(sub, g.pos, g.nvec, POL, DIR, lam)
    assert np.abs(phi).max() == 0 and np.abs(a - a2).max() < 1e-14 and np.abs(ap - a2p).max() < 1e-13
    ext = L.extinction(sub, g.pos, g.area, sig, POL, DIR, lam)
    sca = L.scattering(sub, g.pos, g.area, sig, DIR, lam, *pinfty)
    e0 = O.extinction(g.pos, g.area, ref, POL, DIR, lam, 1.33 ** 2)
    s0 = O.scattering(g.pos, g.area, ref, lam, 1.33 ** 2, *pinfty)
    assert np.allclose(ext, e0, rtol=1e-10) and np.allclose(sca, s0, rtol=1e-10)


def test_potential_derivative_matches_finite_difference(sphere):
    """The exact normal derivative of the plane wave's potentials equals the
    reference's finite difference with eta = 1e-8 to its rounding error."""
    v, f, g = sphere
    sub = L.Substrate(1.0, -8 + 1.5j, 0.0)
    dirs = np.vstack([DIR, [[0.6, 0, 0.8]]])
    pols = np.vstack([POL, [[0.8, 0, -0.6]]])
    a = L.planewave_potential(sub, g.pos, g.nvec, pols, dirs, 600.0)
    b = L.planewave_potential(sub, g.pos, g.nvec, pols, dirs, 600.0, derivative="fd")
    for x, y in ((a[1], b[1]), (a[3], b[3])):
        assert np.abs(x - y).max() < 1e-5 * np.abs(y).max()


def test_small_sphere_follows_the_quasistatic_substrate(sphere):
    """Gold sphere (r 5 nm) 1 nm above glass, near its plasmon resonance: the
    substrate's effect on the extinction (retarded layer / free retarded)
    matches the quasistatic one (bemstatlayer / bemstat) within 2%.  (Off
    resonance, at 650 nm, the two differ by 1% and 6% for the two waves: the
    weak extinction there is sensitive to the discretisation.)"""
    v, f, g = sphere
    eu, ed = 1.0, 2.25
    sub = L.Substrate(eu, ed, 0.0)
    Fr = LS.reflected_F(v, f, 0.0)
    for lam in (450.0, 520.0):
        eg = O.eps_drude(lam, *DRUDE)
        ext = L.extinction(sub, g.pos, g.area, _solve(sub, v, f, g, lam), POL, DIR, lam)
        a2, a2p = O.planewave_potential(g.pos, g.nvec, POL, DIR, lam, eu)
        e0 = O.extinction(g.pos, g.area, O.BemRet(g).initmat(lam, eg, eu).solve({"a2": a2, "a2p": a2p}),
                          POL, DIR, lam, eu)
        up, _ = LS.field(POL, DIR, lam, eu, ed)
        sig_s = LS.BemStatLayer(g, Fr).initmat(eg, eu, eu, ed).solve(LS.potential(g.nvec, up))
        es = LS.extinction(LS.dipole_moment(g.pos, g.area, sig_s), POL, DIR, lam, eu, ed)
        s0 = O.stat_cross_sections(g.pos, g.area, O.BemStat(g).initmat(lam, eg, eu).solve(
            O.planewavestat_potential(g.nvec, POL)), POL, lam, eu)[0]
        assert np.all(np.abs((ext / e0) / (es / s0) - 1) < 0.02)


class _ExactTable:
    """The reflected Green function by direct quadrature at every pair, in
    place of the interpolated table (same clamping and rounding)."""

    def __init__(self, sub):
        self.sub = sub

    def normalised(self, enei):
        return enei

    def interp(self, enei, r, z1, z2):
        sub = self.sub
        r = np.maximum(sub.rmin, np.asarray(r, dtype=float))
        Z = sub.mindist(sub.round(z1)) + sub.mindist(sub.round(z2))
        out = [np.empty((5,) + r.shape, complex) for _ in range(3)]
        for idx in np.ndindex(r.shape):
            for o, v in zip(out, L.reflected_green(sub, enei, r[idx], Z[idx])):
                o[(slice(None),) + idx] = v
        return out[0], out[1], out[2], r, Z


def _charges(n, ncol, seed=3):
    rng = np.random.default_rng(seed)
    c = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    return {"sig2": c(n, ncol), "h2": c(n, 3, ncol)}


def test_point_fields_are_the_potentials_derivatives():
    """Reflected fields at points above the interface: e = i k0 a - grad phi
    and h = curl a, the derivatives by central differences (step 1e-3 nm) of
    the reflected potentials, with the Green function integrated exactly at
    every point-face pair.  Pins the direction of Gp, the structured products
    (ss, sh into phi; p, hh, hs into a) and the curl's index order."""
    v, f = shapes.icosphere(1, 4.0)
    v = v + [0, 0, 5.0]
    sub = L.Substrate(1.0, -8 + 1.5j, 0.0)
    pts = np.array([[12.0, 3.0, 9.0], [-4.0, 10.0, 14.0], [2.0, -1.0, 16.0]])
    lam, hstep = 600.0, 1e-3
    sig = _charges(len(f), 2)
    tab = _ExactTable(sub)
    pg = L.PointGreenLayer(sub, pts, v, f, tab)
    assert not pg.ir.any()  # far points: centroid values throughout
    ref = L.points_reflected(pg, lam, sig)
    dphi = np.zeros_like(ref["a"])
    da = np.zeros(ref["a"].shape[:2] + (3,) + ref["a"].shape[2:], complex)  # (npt, d/dx_c, comp, ncol)
    for c in range(3):
        dx = np.zeros(3)
        dx[c] = hstep
        up = L.points_reflected(L.PointGreenLayer(sub, pts + dx, v, f, tab), lam, sig)
        dn = L.points_reflected(L.PointGreenLayer(sub, pts - dx, v, f, tab), lam, sig)
        dphi[:, c] = (up["phi"] - dn["phi"]) / (2 * hstep)
        da[:, c] = (up["a"] - dn["a"]) / (2 * hstep)
    k0 = 2 * np.pi / lam
    e = 1j * k0 * ref["a"] - dphi
    h = np.stack([da[:, 1, 2] - da[:, 2, 1], da[:, 2, 0] - da[:, 0, 2], da[:, 0, 1] - da[:, 1, 0]], axis=1)
    ee = np.abs(ref["e"] - e).max() / np.abs(e).max()
    eh = np.abs(ref["h"] - h).max() / np.abs(h).max()
    assert ee < 1e-6 and eh < 1e-6, (ee, eh)


def test_point_green_matches_the_particle_green():
    """Points at the particle's own centroids: the point Green function equals
    the particle's reflected G and Gp off the diagonal wherever both refine
    (or both do not), and the refined entries differ from the centroid ones
    (the near-pair weights are in use)."""
    v, f = shapes.icosphere(1, 4.0)
    v = v + [0, 0, 5.0]
    sub = L.Substrate(1.0, 2.25, 0.0)
    pos = O.norm_flat(v, f)[0]
    rg, zg, z2 = L.tab_grid(sub, v, nr=12, nz=12, pts=pos)
    tab = L.GreenTable(sub, rg, zg, z2)
    gp = L.GreenRetLayer(sub, v, f, tab)
    pp = L.PointGreenLayer(sub, pos, v, f, tab)
    a, b = gp.eval(550.0), pp.eval(550.0)
    same = (gp.ir == 1) == (pp.ir == 1)
    np.fill_diagonal(same, False)
    assert same.sum() > 0.9 * (len(f) ** 2 - len(f)) and (pp.ir == 1).sum() > len(f)
    for c in range(5):
        assert np.abs(a["G"][c][same] - b["G"][c][same]).max() < 1e-13 * np.abs(a["G"][c]).max()
        for d in range(3):
            x, y = a["Gp"][c][:, d][same], b["Gp"][c][:, d][same]
            assert np.abs(x - y).max() < 1e-13 * np.abs(a["Gp"][c][:, d]).max()
    # the refined pairs are not the centroid values
    cen = L.PointGreenLayer(sub, pos, v, f, tab, relcutoff=0.0).eval(550.0)
    ir = pp.ir == 1
    np.fill_diagonal(ir, False)
    assert np.abs(cen["G"][1][ir] - b["G"][1][ir]).max() > 1e-3 * np.abs(b["G"][1][ir]).max()


def test_point_fields_vanish_for_equal_media():
    v, f = shapes.icosphere(1, 4.0)
    v = v + [0, 0, 5.0]
    sub = L.Substrate(1.77, 1.77, 0.0)
    pts = np.array([[0.0, 0.0, 10.0], [6.0, 0.0, 5.0]])
    rg, zg, z2 = L.tab_grid(sub, v, nr=8, nz=8, pts=pts)
    pg = L.PointGreenLayer(sub, pts, v, f, L.GreenTable(sub, rg, zg, z2))
    out = L.points_reflected(pg, 600.0, _charges(len(f), 1))
    for k in out:
        assert np.abs(out[k]).max() < 1e-12, k
