"""Pins for the CPU oracle: closed forms, invariants, brute force and Mie.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign, an index swap or a transposed operand) fails it.
"""
import math

import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import mnpbem_oracle as O
from oracle.mie import mie_cross_sections


# ---------------------------------------------------------------- quadrature
def test_rule18_exact_to_degree_11():
    x, y, w = O.triangle_rule18()
    assert len(w) == 28 and abs(w.sum() - 1) < 1e-14
    for a in range(12):
        for b in range(12 - a):
            exact = 2.0 * math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(np.sum(w * x ** a * y ** b) - exact) < 1e-13, (a, b)


def test_lglnodes_match_legendre():
    from numpy.polynomial import legendre as L
    for n in (5, 7):
        x, w = O.lglnodes(n)
        c = np.zeros(n + 1)
        c[-1] = 1
        interior = np.sort(L.legroots(L.legder(c)))
        assert np.allclose(np.sort(x), np.concatenate([[-1], interior, [1]]), atol=1e-14)
        assert np.allclose(w, 2 / (n * (n + 1) * L.legval(x, c) ** 2), atol=1e-14)
        for deg in range(2 * n):  # LGL is exact to degree 2n-1
            assert abs(np.sum(w * x ** deg) - (2 / (deg + 1) if deg % 2 == 0 else 0)) < 1e-13


def _inv_r_triangle(p, v):
    """Closed form of the integral of 1/|x-p| over a flat triangle, p in its plane."""
    tot = 0.0
    n = np.cross(v[1] - v[0], v[2] - v[0])
    n /= np.linalg.norm(n)
    for e in range(3):
        a, b = v[e], v[(e + 1) % 3]
        t = (b - a) / np.linalg.norm(b - a)
        m = np.cross(t, n)  # outward in-plane edge normal for ccw vertices
        dist = np.dot(a - p, m)
        sa, sb = np.dot(a - p, t), np.dot(b - p, t)
        ra, rb = np.linalg.norm(a - p), np.linalg.norm(b - p)
        tot += dist * np.log((rb + sb) / (ra + sa))
    return tot


def test_polar_rule_self_integral_closed_form():
    """The polar rule integrates 1/r from the centroid of near-equilateral faces.

    Its 6-point angular LGL rule integrates 1/|sin phi| only to ~1e-3, so the
    tolerance is that of the rule itself; a wrong mapping or weight is O(1).
    """
    x3, y3, w3 = O.polar_rule()
    assert len(w3) == 144 and abs(w3.sum() - 1) < 1e-14
    # rho = (x+1+1e-6)/2 reaches 5e-7 past the edges
    assert np.all(w3 > 0) and np.all(x3 > -1e-6) and np.all(y3 > -1e-6) and np.all(x3 + y3 < 1 + 1e-6)
    v, f = shapes.ellipsoid(2, (6.0, 6.0, 9.0))
    for i in range(0, len(f), 7):
        t = v[f[i]]
        c = t.mean(axis=0)
        pts = x3[:, None] * t[0] + y3[:, None] * t[1] + (1 - x3 - y3)[:, None] * t[2]
        area = 0.5 * np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0]))
        num = np.sum(w3 * area / np.linalg.norm(pts - c, axis=1))
        assert abs(num / _inv_r_triangle(c, t) - 1) < 2e-3


# ------------------------------------------------------------------ geometry
def test_norm_flat_basis_and_orientation():
    v, f = shapes.icosphere(2, 3.0)
    pos, t1, t2, n, a = O.norm_flat(v, f)
    for u in (t1, t2, n):
        assert np.allclose(np.linalg.norm(u, axis=1), 1, atol=1e-14)
    assert np.allclose(np.sum(t1 * t2, 1), 0, atol=1e-14)
    assert np.allclose(np.sum(t1 * n, 1), 0, atol=1e-14)
    assert np.all(np.sum(n * pos, 1) > 0)  # outward
    # Heron's formula
    A, B, C = (np.linalg.norm(v[f[:, i]] - v[f[:, (i + 1) % 3]], axis=1) for i in range(3))
    s = (A + B + C) / 2
    assert np.allclose(a, np.sqrt(s * (s - A) * (s - B) * (s - C)), rtol=1e-12)


def test_refinematrix_brute_force():
    v, f = shapes.ellipsoid(1, (3.0, 4.0, 6.0))
    pos, _, _, _, _ = O.norm_flat(v, f)
    rad = O.bradius(v, f, pos)
    rows, cols, kind = O.refinematrix(pos, rad, 3.0, 0.0)
    got = {(r, c): k for r, c, k in zip(rows, cols, kind)}
    n = len(f)
    for j in range(n):
        for i in range(n):
            d = np.linalg.norm(pos[i] - pos[j])
            if i == j:
                assert got[(i, j)] == 2
            elif (d - rad[j]) / rad[i] < 3.0 or d - rad[j] < 0:
                assert got[(i, j)] == 1
            else:
                assert (i, j) not in got
    # column-major order
    lin = cols * n + rows
    assert np.all(np.diff(lin) > 0)


# ----------------------------------------------------------- Green function
@pytest.fixture(scope="module")
def green_small():
    v, f = shapes.ellipsoid(2, (6.0, 6.0, 9.0))
    return O.GreenRet(v, f, closed=False), v, f


def test_taylor_tables_match_direct_quadrature(green_small):
    """sum_n g_n (ik)^n e^{ik r0} equals rule-18 quadrature of e^{ikr}/r up to O((k dr)^6)."""
    g, v, f = green_small
    k = 0.02 + 0.005j
    sel = np.nonzero(g.kind == 1)[0][::37]
    i, j = g.rows[sel], g.cols[sel]
    pts, w = O.quad_flat(v, f, j)
    x = g.pos[i][:, None, :] - pts
    r = np.linalg.norm(x, axis=2)
    r0 = np.linalg.norm(g.pos[i] - g.pos[j], axis=1)
    fr = (1j * k) ** np.arange(g.order + 1)
    Gt = (g.g[sel] @ fr) * np.exp(1j * k * r0)
    Gq = np.sum(w * np.exp(1j * k * r) / r, axis=1)
    assert np.max(np.abs(Gt / Gq - 1)) < 1e-9
    Ft = np.einsum("pco,o->pc", g.f[sel], fr) * np.exp(1j * k * r0)[:, None]
    Fq = np.sum((w * (1j * k - 1 / r) * np.exp(1j * k * r) / r ** 2)[:, :, None] * x, axis=1)
    assert np.max(np.abs(Ft - Fq)) / np.max(np.abs(Fq)) < 1e-9


def test_self_terms(green_small):
    """g_0 = closed-form integral of 1/r over the face; higher orders follow e^{ikr}/r."""
    g, v, f = green_small
    sd = np.nonzero(g.kind == 2)[0]
    assert np.all(g.rows[sd] == g.cols[sd]) and len(sd) == g.n
    for s in sd[::29]:
        i = g.rows[s]
        ref = _inv_r_triangle(g.pos[i], v[f[i]])
        assert abs(g.g[s, 0] / ref - 1) < 2e-3
    k = 0.05
    pts, w = O.quadpol_flat(v, f, g.area, g.rows[sd])
    r = np.linalg.norm(g.pos[g.rows[sd]][:, None, :] - pts, axis=2)
    Gq = np.sum(w * np.exp(1j * k * r) / r, axis=1)
    Gt = g.g[sd] @ ((1j * k) ** np.arange(g.order + 1))
    assert np.max(np.abs(Gt / Gq - 1)) < 1e-10


def test_static_sum_rule_before_and_after_correction():
    """Gauss: sum_row area_row F_row,i / area_i -> -2 pi for a closed surface.

    Before the closed-surface correction the discrete column sums converge to
    -2 pi at first order in the face size (a sign or orientation slip gives
    +2 pi or 0); after it they equal -2 pi to rounding.
    """
    err = []
    for lev in (1, 2):
        v, f = shapes.icosphere(lev, 5.0)
        g = O.GreenRet(v, f, closed=False)
        s = np.sum(g.area[:, None] * g._F_static() / g.area[None, :], axis=0)
        err.append(np.max(np.abs(s + 2 * np.pi)))
    assert err[1] < 0.1 * 2 * np.pi and err[1] < 0.6 * err[0]
    g2 = O.GreenRet(v, f, closed=True)
    s2 = np.sum(g2.area[:, None] * g2._F_static() / g2.area[None, :], axis=0)
    assert np.allclose(s2, -2 * np.pi, atol=1e-11)


def test_kernel_closed_form_and_reciprocity(green_small):
    g, _, _ = green_small
    k = 0.07 + 0.01j
    m = g.eval(k, ("G", "F", "H1", "H2"))
    ref = np.zeros((g.n, g.n), bool)
    ref[g.rows, g.cols] = True
    i, j = np.nonzero(~ref)
    d = np.linalg.norm(g.pos[i] - g.pos[j], axis=1)
    Gc = g.area[j] * np.exp(1j * k * d) / d
    Fc = (np.sum(g.nvec[i] * (g.pos[i] - g.pos[j]), 1) * (1j * k - 1 / d) / d ** 2
          * g.area[j] * np.exp(1j * k * d))
    assert np.allclose(m["G"][i, j], Gc, rtol=1e-13)
    assert np.allclose(m["F"][i, j], Fc, rtol=1e-12, atol=1e-15 * np.abs(Fc).max())
    sym = ~ref & ~ref.T
    a, b = np.nonzero(sym)
    assert np.allclose(m["G"][a, b] / g.area[b], m["G"][b, a] / g.area[a], rtol=1e-13)
    assert np.allclose(m["H1"] - m["F"], 2 * np.pi * np.eye(g.n))
    assert np.allclose(m["H2"] - m["F"], -2 * np.pi * np.eye(g.n))


# ------------------------------------------------------------------- solver
def test_chain_solution_satisfies_exact_operator():
    """mldivide (the Sigma/Delta chain) solves the equations afun applies."""
    v, f = shapes.ellipsoid(2, (5.0, 5.0, 12.0))
    g = O.GreenRet(v, f)
    enei = 600.0
    e1 = O.eps_drude(enei, 10.0, 9.065, 0.0708)
    bem = O.BemRet(g).initmat(enei, e1, 1.77)
    pol = np.array([[1.0, 0, 0], [0, 0, 1.0]])
    dirs = np.array([[0, 0, 1.0], [1.0, 0, 0]])
    a2, a2p = O.planewave_potential(g.pos, g.nvec, pol, dirs, enei, 1.77)
    exc = {"a2": a2, "a2p": a2p}
    sig = bem.solve(exc)
    phi, a, alpha, De = bem.excitation(exc)
    r_phi, r_a, r_De, r_alpha = bem.afun(sig)
    assert np.max(np.abs(r_phi)) < 1e-12 * np.max(np.abs(a))
    for got, want in ((r_a, a), (r_De, De), (r_alpha, alpha)):
        assert np.linalg.norm(got - want) < 1e-11 * np.linalg.norm(want)


# ---------------------------------------------------------------------- Mie
def test_mie_rayleigh_limit_and_energy():
    a = 1.0
    eps = -2.5 + 0.8j
    lam = np.array([2000.0, 4000.0])
    ext, sca, ab = mie_cross_sections(lam, eps, 1.0, 2 * a)
    k = 2 * np.pi / lam
    al = (eps - 1) / (eps + 2)
    assert np.allclose(ab, 4 * np.pi * k * a ** 3 * np.imag(al), rtol=2e-3)
    assert np.allclose(sca, 8 * np.pi / 3 * k ** 4 * a ** 6 * np.abs(al) ** 2, rtol=2e-3)
    e2, s2, a2 = mie_cross_sections([500.0], 2.25, 1.0, 200.0)
    assert abs(a2[0]) < 1e-10 * e2[0]


def test_bem_converges_to_mie(pinfty):
    pn, pa = pinfty
    pol, dirs = np.array([[1.0, 0, 0]]), np.array([[0, 0, 1.0]])
    enei = 520.0
    e1 = O.eps_drude(enei, 10.0, 9.065, 0.0708)
    me, ms, _ = mie_cross_sections(enei, e1, 1.0, 20.0)
    errs = []
    for lev in (1, 2):
        v, f = shapes.icosphere(lev, 10.0)
        g = O.GreenRet(v, f)
        bem = O.BemRet(g).initmat(enei, e1, 1.0)
        a2, a2p = O.planewave_potential(g.pos, g.nvec, pol, dirs, enei, 1.0)
        sig = bem.solve({"a2": a2, "a2p": a2p})
        ext = O.extinction(g.pos, g.area, sig, pol, dirs, enei, 1.0)[0]
        sca = O.scattering(g.pos, g.area, sig, enei, 1.0, pn, pa)[0]
        errs.append((abs(ext / me[0] - 1), abs(sca / ms[0] - 1)))
    assert errs[1][0] < 0.1 and errs[1][1] < 0.2
    assert errs[1][0] < 0.5 * errs[0][0] and errs[1][1] < 0.5 * errs[0][1]
