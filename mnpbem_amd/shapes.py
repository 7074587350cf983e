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


def ellipsoid_segment(nphi, ntheta, semi_axes, phi_min=0.0, phi_max=np.pi / 2):
    """Flat triangulation of the ellipsoid surface segment phi_min <= phi <=
    phi_max (all theta): (a sin t cos p, b sin t sin p, c cos t) on a grid of
    nphi azimuthal and ntheta polar intervals, quads split along one diagonal
    and fans at the poles, 2 nphi (ntheta - 1) faces, oriented outward.
    Grid lines with cos(phi) = 0 or sin(phi) = 0 get exact zeros, so that the
    segments phi in [0, pi/2] ('xy'), [0, pi] ('y') and [-pi/2, pi/2] ('x')
    are the parts that the mirror symmetries complete to the closed surface
    (cf. Particles/particleshapes/trispheresegment.m, which grids phi and
    theta the same way)."""
    a, b, c = (float(x) for x in semi_axes)
    ph = np.linspace(float(phi_min), float(phi_max), int(nphi) + 1)
    th = np.linspace(0.0, np.pi, int(ntheta) + 1)
    cp, sp = np.cos(ph), np.sin(ph)
    cp[np.abs(cp) < 1e-12] = 0.0
    sp[np.abs(sp) < 1e-12] = 0.0
    ring = np.concatenate([np.column_stack([a * np.sin(t) * cp, b * np.sin(t) * sp, np.full_like(ph, c * np.cos(t))])
                           for t in th[1:-1]])
    verts = np.vstack([[0.0, 0.0, c], ring, [0.0, 0.0, -c]])
    m = len(ph)
    top, bot = 0, len(verts) - 1
    idx = lambda r, j: 1 + r * m + j
    faces = []
    for j in range(int(nphi)):
        faces.append([top, idx(0, j), idx(0, j + 1)])
        for r in range(int(ntheta) - 2):
            faces.append([idx(r, j), idx(r + 1, j), idx(r + 1, j + 1)])
            faces.append([idx(r, j), idx(r + 1, j + 1), idx(r, j + 1)])
        faces.append([idx(int(ntheta) - 2, j), bot, idx(int(ntheta) - 2, j + 1)])
    return verts, orient_outward(verts, np.array(faces, dtype=np.int64), center=np.zeros(3))


_OCTANTS = [(sx, sy, sz) for sx in (1, -1) for sy in (1, -1) for sz in (1, -1)]


def ellipsoid_octants(m, semi_axes, sym=None):
    """Ellipsoid meshed octant by octant: each octant's spherical triangle,
    corners (+-1, 0, 0), (0, +-1, 0), (0, 0, +-1), is cut into m^2 flat
    triangles on the barycentric grid of step 1/m, projected radially onto
    the unit sphere and scaled to the semi-axes (nm).  The octant edges lie
    in the coordinate planes, so the part x, y >= 0 (sym 'xy', 2 m^2
    faces), x >= 0 ('x') or y >= 0 ('y', 4 m^2 faces each) is what the
    mirror copies complete to the closed surface; sym None gives the whole
    surface (8 m^2 faces).  Faces oriented outward."""
    keep = {None: lambda o: True, "xy": lambda o: o[0] > 0 and o[1] > 0, "x": lambda o: o[0] > 0,
            "y": lambda o: o[1] > 0}[sym]
    m = int(m)
    verts, faces = [], []
    for o in _OCTANTS:
        if not keep(o):
            continue
        A, Bv, Cv = np.diag(np.asarray(o, dtype=np.float64))
        index = {}
        for i in range(m + 1):
            for j in range(m + 1 - i):
                pnt = (i * A + j * Bv + (m - i - j) * Cv) / m
                index[(i, j)] = len(verts)
                verts.append(pnt / np.linalg.norm(pnt))
        for i in range(m):
            for j in range(m - i):
                faces.append([index[(i, j)], index[(i + 1, j)], index[(i, j + 1)]])
                if i + j < m - 1:
                    faces.append([index[(i + 1, j)], index[(i + 1, j + 1)], index[(i, j + 1)]])
    verts = np.asarray(verts) * np.asarray(semi_axes, dtype=np.float64)[None, :]
    return verts, orient_outward(verts, np.array(faces, dtype=np.int64), center=np.zeros(3))


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
