"""Pins the oracle's adjoint of the plane-wave extinction with respect to the
vertex positions (oracle/adjoint_oracle.py) against central differences of
the oracle's own forward solve, on seeded directions through all vertex
coordinates of a 320-face ellipsoid in eps 1.77 (Drude gold, two
polarisations with unequal weights).

The step 5e-4 nm sits between the round-off of the differences (below
1e-4 nm) and the displacement that flips a pair of the refinement selection
(about 5e-3 nm here); at it the differences agree with the adjoint to about
1e-7.  A plausible slip in the derivation (a dropped sum-rule term, a sign of
the r0 derivative, the wrong vertex of a face) moves the result by far more
than the 1e-5 allowed.
"""
import numpy as np

from mnpbem_amd import shapes
from oracle import adjoint_oracle as A
from oracle import mnpbem_oracle as O

ENEI = 600.0
POL = np.array([[1.0, 0, 0], [0, 0, 1.0]])
DIRS = np.array([[0, 0, 1.0], [1.0, 0, 0]])
W = np.array([1.0, 0.7])


def test_adjoint_gradient_matches_central_differences():
    v, f = shapes.ellipsoid(2, (10.0, 10.0, 25.0))
    eps1, eps2 = O.eps_drude(ENEI, 10.0, 9.065, 0.0708), 1.77
    grad, ext = A.extinction_gradient(v, f, ENEI, eps1, eps2, POL, DIRS, W)
    assert grad.shape == v.shape and ext > 0
    rng = np.random.default_rng(11)
    h = 5e-4
    for _ in range(3):
        u = rng.normal(size=v.shape)
        u /= np.linalg.norm(u)
        ep = A.weighted_extinction(v + h * u, f, ENEI, eps1, eps2, POL, DIRS, W)
        em = A.weighted_extinction(v - h * u, f, ENEI, eps1, eps2, POL, DIRS, W)
        fd = (ep - em) / (2 * h)
        assert abs(np.sum(grad * u) - fd) <= 1e-5 * abs(fd)


def test_adjoint_gradient_invariances():
    """A rigid translation moves nothing (sum of the gradient over the
    vertices is zero), and scaling the weights scales the gradient."""
    v, f = shapes.ellipsoid(1, (10.0, 10.0, 25.0))
    eps1, eps2 = O.eps_drude(ENEI, 10.0, 9.065, 0.0708), 1.77
    g1, _ = A.extinction_gradient(v, f, ENEI, eps1, eps2, POL, DIRS, W)
    g2, _ = A.extinction_gradient(v, f, ENEI, eps1, eps2, POL, DIRS, 2 * W)
    assert np.allclose(g2, 2 * g1, rtol=1e-12, atol=0)
    assert np.all(np.abs(g1.sum(axis=0)) <= 1e-9 * np.abs(g1).sum(axis=0))
