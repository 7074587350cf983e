"""Pins for the substrate oracle (quasistatic solver and plane wave above one
interface)."""
import numpy as np

from mnpbem_amd import shapes
from oracle import layer_oracle as L
from oracle import mnpbem_oracle as O

DRUDE = (10.0, 9.065, 0.0708)


def test_fresnel_energy_and_normal_incidence():
    enei, e1, e2 = 600.0, 1.0, 2.25
    d = np.array([[0.0, 0.0, -1.0], [np.sin(0.5), 0, -np.cos(0.5)], [0, np.sin(1.1), np.cos(1.1)]])
    r = L.fresnel(d, enei, e1, e2)
    n1, n2 = 1.0, 1.5
    assert np.isclose(r["re"][0], (n1 - n2) / (n1 + n2)) and np.isclose(r["rm"][0], -(n1 - n2) / (n1 + n2))
    for i in range(3):
        k1z, k2z = r["kr"][i][2], -r["kt"][i][2]
        if d[i, 2] > 0:  # incidence from below: the media swap roles
            k1z, k2z = -r["kr"][i][2], r["kt"][i][2]
            ea, eb = e2, e1
        else:
            ea, eb = e1, e2
        te_flux = abs(r["re"][i]) ** 2 + (k2z / k1z).real * abs(r["te"][i]) ** 2
        tm_flux = abs(r["rm"][i]) ** 2 + (ea * k2z / (eb * k1z)).real * abs(r["tm"][i]) ** 2
        assert np.isclose(te_flux, 1.0) and np.isclose(tm_flux, 1.0)


def test_equal_media_reduce_to_free_space(pinfty):
    v, f = shapes.icosphere(2, 8.0)
    v = v + np.array([0, 0, 20.0])
    g = O.GreenRet(v, f)
    enei, em = 560.0, 1.33
    ein = O.eps_drude(enei, *DRUDE)
    pol = np.array([[1.0, 0, 0], [0, 1.0, 0]])
    dirs = np.array([[0, 0, -1.0], [0, 0, -1.0]])
    up, down = L.field(pol, dirs, enei, em, em)
    assert np.allclose(up, pol) and np.allclose(down, pol)
    Fr = L.reflected_F(v, f, 0.0)
    phip = L.potential(g.nvec, up)
    sig = L.BemStatLayer(g, Fr).initmat(ein, em, em, em).solve(phip)
    free = O.BemStat(g).initmat(enei, ein, em).solve(O.planewavestat_potential(g.nvec, pol))
    assert np.allclose(sig, free, rtol=1e-12, atol=1e-14 * np.abs(free).max())
    dip = L.dipole_moment(g.pos, g.area, sig)
    ext = L.extinction(dip, pol, dirs, enei, em, em)
    sca = L.scattering(dip, dirs, enei, em, em, *pinfty)
    ext0, sca0, abs0 = O.stat_cross_sections(g.pos, g.area, free, pol, enei, em)
    # extinction by the optical theorem is what planewavestat calls absorption;
    # efarfield's transmitted branch regularises k1z by 1e-10i (about 1e-8 here)
    assert np.allclose(ext, abs0, rtol=1e-6)
    # the flat 256-point sphere integrates |p_perp|^2 to its faceting (sum area / 4 pi)
    assert np.allclose(sca / sca0, pinfty[1].sum() / (4 * np.pi), rtol=2e-3)


def test_reflected_F_is_the_image_kernel_away_from_the_interface():
    v, f = shapes.icosphere(2, 6.0)
    v = v + np.array([0, 0, 9.0])
    z0 = 1.0
    Fr = L.reflected_F(v, f, z0)
    pos, _, _, nvec, area = O.norm_flat(v, f)
    img = pos * np.array([1, 1, -1.0]) + np.array([0, 0, 2 * z0])
    d = pos[:, None, :] - img[None, :, :]
    r = np.linalg.norm(d, axis=2)
    want = -np.einsum("ic,ijc->ij", nvec, d) / r ** 3 * area[None, :]
    # pairs with (d - rad_source) / rad_observer < RelCutoff (3) are refined
    # (refinematrix.m:56-60); the others are the collocation kernel at the
    # image centroids
    rad = O.bradius(v, f, pos)
    far = r > rad[None, :] + 3.2 * rad[:, None]
    assert far.mean() > 0.5 and (~far).sum() > 0
    assert np.allclose(Fr[far], want[far], rtol=1e-10, atol=0)
    # the refined pairs differ from it, and not by a sign
    assert np.abs(Fr[~far] - want[~far]).max() > 1e-6 * np.abs(want).max()
    assert np.all(np.sign(Fr[~far].real) == np.sign(want[~far]))


def test_small_sphere_above_substrate_follows_image_dipole_model():
    """Uniform field on a sphere of radius a at height h above a substrate:
    p = alpha0 E0 / (1 + c alpha0 f2 / (2h)^3), c = 1 parallel, 2 normal, with
    the image p' = f2 (px, py, -pz) and alpha0 from the free sphere."""
    a, h = 5.0, 15.0
    v, f = shapes.icosphere(3, a)
    v = v + np.array([0, 0, h])
    g = O.GreenRet(v, f)
    eps_in, em, es = -2.0 + 0.4j, 1.0, 2.25
    f2 = -(es - em) / (es + em)
    E0 = np.array([[1.0, 0, 0], [0, 0, 1.0]])
    phip = -g.nvec @ E0.T
    Fr = L.reflected_F(v, f, 0.0)
    sub = L.dipole_moment(g.pos, g.area, L.BemStatLayer(g, Fr).initmat(eps_in, em, em, es).solve(phip))
    free = L.dipole_moment(g.pos, g.area, L.BemStatLayer(g, Fr).initmat(eps_in, em, em, em).solve(phip))
    D3 = (2 * h) ** 3
    for i, (c, comp) in enumerate(((1, 0), (2, 2))):
        alpha0 = free[i, comp]
        model = 1 / (1 + c * alpha0 * f2 / D3) - 1
        bem = sub[i, comp] / free[i, comp] - 1
        assert abs(model) > 5e-3
        assert abs(bem - model) < 0.05 * abs(model), (c, bem, model)
    # a sign slip in f2 would move the shift the other way
    assert abs(bem - (1 / (1 - 2 * alpha0 * f2 / D3) - 1)) > 0.5 * abs(model)


def _fine_sphere():
    """5,120 directions and weights of the unit sphere, tilted by 0.1 rad so
    that no direction is horizontal (where equal media give 0 / 0)."""
    from mnpbem_amd import shapes
    v, f = shapes.ellipsoid(4, (1.0, 1.0, 1.0))
    c, s = np.cos(0.1), np.sin(0.1)
    v = v @ np.array([[1, 0, 0], [0, c, s], [0, -s, c]])
    g = O.GreenRet(v, f)
    return g.nvec, g.area / np.sum(g.area) * 4 * np.pi


def test_dipolestatlayer_free_limit():
    """With eps_down = eps_up the substrate disappears: the potential is the
    free dipole's, tot - rad equals dipolestat's tot and rad its |p_tot|^2
    (the radiated power of a free dipole, to the quadrature of the sphere)."""
    from mnpbem_amd import shapes
    from oracle import dipole_oracle as D
    v, f = shapes.ellipsoid(2, (8.0, 8.0, 12.0))
    v = v + np.array([0, 0, 20.0])
    g = O.GreenRet(v, f)
    V, Fc = L.image_mesh(v, f, 0.0)
    gi = O.GreenRet(V, Fc, closed=False)
    enei, eb = 560.0, 1.7
    ein = O.eps_drude(enei, 10.0, 9.065, 0.0708)
    pts = np.array([[0, 0, 36.0], [3.0, 0, 37.0]])
    dip = np.zeros((2, 3, 2))
    dip[:, 2, 0] = 1
    dip[:, 0, 1] = 1
    e_free = D.dipolestat_field(g.pos, pts, dip, eb)
    assert np.allclose(L.dipolestatlayer_field(g.pos, pts, dip, eb, eb, 0.0), e_free, rtol=1e-14)
    sig = O.BemStat(g).initmat(enei, ein, eb).solve(D.dipolestat_potential(g.nvec, e_free))
    nv, ar = _fine_sphere()
    tot, rad, _ = L.dipolestatlayer_decayrate(g, gi, pts, dip, enei, eb, eb, sig, nv, ar)
    tf, rf, _ = D.dipolestat_decayrate(g, pts, dip, enei, eb, sig)
    assert np.allclose(tot - rad, tf, rtol=1e-12)
    assert np.allclose(rad, rf, rtol=2e-3), (rad, rf)


def test_dipolestatlayer_decayrate0():
    """Without a particle: a lossless substrate gives tot = rad; the total
    rate of a lossy one exceeds the radiative rate by the image term, which
    decays as the cube of the height; equal media radiate as a free dipole."""
    nv, ar = _fine_sphere()
    pts = np.array([[0, 0, 5.0], [0, 0, 10.0]])
    dip = np.zeros((2, 3, 2))
    dip[:, 2, 0] = 1
    dip[:, 0, 1] = 1
    tot, rad, _ = L.dipolestatlayer_decayrate0(pts, dip, 600.0, 1.0, 2.25, 0.0, nv, ar)
    assert np.array_equal(tot, rad) and np.all(rad > 0)
    # lossy metal substrate: with the image field in the upper medium the
    # absorbed part is positive and falls as height^-3; the reference's
    # permittivity of the image point's (lower) medium makes it negative
    tot, rad, _ = L.dipolestatlayer_decayrate0(pts, dip, 600.0, 1.0, -10 + 1j, 0.0, nv, ar, image_eps="upper")
    nr = tot - rad
    assert np.all(nr > 0)
    assert np.allclose(nr[0] / nr[1], 8.0, rtol=1e-12)
    assert np.allclose(nr[:, 0] / nr[:, 1], 2.0, rtol=1e-12)  # perpendicular vs parallel image field
    tot, rad, _ = L.dipolestatlayer_decayrate0(pts, dip, 600.0, 1.0, -10 + 1j, 0.0, nv, ar)
    assert np.all(tot - rad < 0)
    tot, rad, _ = L.dipolestatlayer_decayrate0(pts, dip, 600.0, 1.3, 1.3, 0.0, nv, ar)
    assert np.allclose(rad, 1.0, rtol=2e-3) and np.array_equal(tot, rad)
