"""Writes tests/golden/oracle_small.npz from the oracle.

The reference ships no golden vectors (it has no tests), so these fixtures
record the pinned oracle's outputs on small meshes.  They guard the oracle
against regressions and let GPU tests check results without recomputing.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mnpbem_amd import shapes  # noqa: E402
from oracle import mnpbem_oracle as O  # noqa: E402

DRUDE = (10.0, 9.065, 0.0708)
CASES = {
    # name: (mesh, eps_out, wavelengths, pol, dir)
    "sphere80": (lambda: shapes.icosphere(1, 10.0), 1.0, [400.0, 520.0, 800.0],
                 [[1.0, 0, 0]], [[0, 0, 1.0]]),
    "ellipsoid320": (lambda: shapes.ellipsoid(2, (10.0, 10.0, 40.0)), 1.77, [600.0, 900.0],
                     [[1.0, 0, 0], [0, 0, 1.0]], [[0, 0, 1.0], [1.0, 0, 0]]),
}


def compute():
    v, f = shapes.trisphere256(2.0)
    _, _, _, pn, pa = O.norm_flat(v, f)
    out = {}
    for name, (mesh, eps_out, lams, pol, dirs) in CASES.items():
        verts, faces = mesh()
        g = O.GreenRet(verts, faces)
        pol, dirs = np.array(pol), np.array(dirs)
        sel = np.arange(0, len(g.rows), 97)
        out[f"{name}/pairs"] = np.stack([g.rows[sel], g.cols[sel]])
        out[f"{name}/g"] = g.g[sel]
        out[f"{name}/f"] = g.f[sel]
        out[f"{name}/nnz"] = np.array(len(g.rows))
        G = g.eval(0.1 + 0.02j, ("G", "H1"))
        out[f"{name}/G_colsum"] = G["G"].sum(axis=0)
        out[f"{name}/H1_rowsum"] = G["H1"].sum(axis=1)
        ext, sca = [], []
        for lam in lams:
            bem = O.BemRet(g).initmat(lam, O.eps_drude(lam, *DRUDE), eps_out)
            a2, a2p = O.planewave_potential(g.pos, g.nvec, pol, dirs, lam, eps_out)
            sig = bem.solve({"a2": a2, "a2p": a2p})
            ext.append(O.extinction(g.pos, g.area, sig, pol, dirs, lam, eps_out))
            sca.append(O.scattering(g.pos, g.area, sig, lam, eps_out, pn, pa))
        out[f"{name}/enei"] = np.array(lams)
        out[f"{name}/ext"] = np.array(ext)
        out[f"{name}/sca"] = np.array(sca)
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_small.npz")
    np.savez_compressed(path, **compute())
    print("wrote", path, os.path.getsize(path), "bytes")
