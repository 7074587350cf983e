"""CPU pin of the oracle's point Green function as used for potentials at
points: the gradient of phi = G sig by central differences in the point
equals -(-Gp sig), i.e. Gp is the gradient of G, in the far branch and in the
refined (rule-18 Taylor) branch near the surface alike."""
import numpy as np

from mnpbem_amd import shapes
from oracle import dipole_oracle as D
from oracle import mnpbem_oracle as O


def test_point_green_gradient_is_gp():
    v, f = shapes.ellipsoid(2, (10.0, 10.0, 15.0))
    g = O.GreenRet(v, f)
    rng = np.random.default_rng(3)
    sig = rng.normal(size=(g.n, 2)) + 1j * rng.normal(size=(g.n, 2))
    pts = np.array([[0.0, 0.0, 20.0], [12.0, 3.0, -4.0], [0.0, 0.0, 15.8], [1.0, -2.0, 3.0]])
    k = 2 * np.pi / 600.0 * np.sqrt(1.77)
    G, Gp = D.point_green(g, pts, k)
    assert np.any(((np.linalg.norm(pts[2] - g.pos, axis=1) - g.rad) / g.rad) < 3.0)  # near branch in use
    h = 1e-4
    for c in range(3):
        dx = np.zeros(3)
        dx[c] = h
        Gu, _ = D.point_green(g, pts + dx, k)
        Gd, _ = D.point_green(g, pts - dx, k)
        fd = (Gu - Gd) @ sig / (2 * h)
        want = Gp[:, c] @ sig
        assert np.abs(fd - want).max() < 1e-6 * np.abs(want).max()
