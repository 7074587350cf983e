"""Flat-triangle surface meshes used by the retarded BEM path.

Face convention (MNPBEM ``Particles/@particle/private/norm_flat.m:36-38``):
a face ``(v1, v2, v3)`` has normal ``(v1 - v2) x (v3 - v2)``; every mesh built
here orders its faces so that this normal points out of the particle.
Faces are 0-based ``int64`` (n, 3); vertices are ``float64`` (m, 3) in nm.
"""
import os

import numpy as np

_ICO_T = (1.0 + 5.0 ** 0.5) / 2.0
_ICO_VERTS = np.array([
    [-1, _ICO_T, 0], [1, _ICO_T, 0], [-1, -_ICO_T, 0], [1, -_ICO_T, 0],
    [0, -1, _ICO_T], [0, 1, _ICO_T], [0, -1, -_ICO_T], [0, 1, -_ICO_T],
    [_ICO_T, 0, -1], [_ICO_T, 0, 1], [-_ICO_T, 0, -1], [-_ICO_T, 0, 1]],
    dtype=np.float64)
_ICO_FACES = np.array([
    [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
    [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
    [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
    [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)


def orient_outward(verts, faces, center=None):
    """Reorder each face so that ``(v1-v2) x (v3-v2)`` points away from ``center``."""
    faces = np.array(faces, dtype=np.int64, copy=True)
    c = verts.mean(axis=0) if center is None else np.asarray(center, float)
    v1, v2, v3 = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    n = np.cross(v1 - v2, v3 - v2)
    cen = (v1 + v2 + v3) / 3.0 - c
    flip = np.einsum("ij,ij->i", n, cen) < 0
    faces[flip] = faces[flip][:, ::-1]
    return faces


def icosphere(level, radius=1.0):
    """Icosahedron subdivided ``level`` times (20 * 4**level faces), vertices on a sphere.

    Each subdivision splits a triangle into four through its edge midpoints; the
    midpoints are projected onto the sphere after every level.
    """
    verts = _ICO_VERTS / np.linalg.norm(_ICO_VERTS, axis=1, keepdims=True)
    faces = _ICO_FACES.copy()
    for _ in range(int(level)):
        vlist = list(verts)
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            idx = cache.get(key)
            if idx is None:
                m = 0.5 * (vlist[a] + vlist[b])
                vlist.append(m / np.linalg.norm(m))
                idx = len(vlist) - 1
                cache[key] = idx
            return idx

        nf = np.empty((4 * len(faces), 3), dtype=np.int64)
        for t, (a, b, c) in enumerate(faces.tolist()):
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf[4 * t:4 * t + 4] = [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        verts = np.asarray(vlist)
        faces = nf
    verts = verts * float(radius)
    return verts, orient_outward(verts, faces, center=np.zeros(3))


def ellipsoid(level, semi_axes):
    """Icosphere of unit radius scaled per axis to the given semi-axes (nm)."""
    verts, faces = icosphere(level, 1.0)
    verts = verts * np.asarray(semi_axes, dtype=np.float64)[None, :]
    return verts, orient_outward(verts, faces, center=np.zeros(3))


def sphere256_points():
    """The 256-point unit-sphere point set behind MNPBEM's ``trisphere(256)`` (float32)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "sphere256.npy")
    return np.load(path, allow_pickle=False)


def trisphere256(diameter=2.0):
    """Flat triangulation of the 256-point sphere, as ``trisphere(256, diameter)``.

    ``Particles/particleshapes/trisphere.m:38-61``: the float32 points are
    promoted to double, scaled, then renormalised to radius ``diameter/2``
    (the curved-particle vertex projection); the faces are the convex hull
    (``misc/sphtriangulate.m``, a Delaunay triangulation of the stereographic
    projection closed at the pole), oriented outward.
    """
    from scipy.spatial import ConvexHull

    v = sphere256_points().astype(np.float64)
    v = 0.5 * v * diameter
    v = 0.5 * diameter * (v / np.sqrt(np.sum(v * v, axis=1))[:, None])
    faces = ConvexHull(v).simplices.astype(np.int64)
    return v, orient_outward(v, faces, center=np.zeros(3))
