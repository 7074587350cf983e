"""Write ``sphere256.npy``: the 256-point minimum-energy unit-sphere point set.

MNPBEM integrates the scattered Poynting flux over ``trisphere(256, 2)``
(``Simulation/retarded/@planewaveret/init.m:29``), a flat triangulation of the
point set ``sphere256`` stored in ``Particles/particleshapes/trisphere.mat``.
The points are data, not code; this script copies them (float32, 256x3, as
stored) into ``sphere256.npy`` so that the package does not need the
reference tree at run time.  It reads the ``.mat`` file with ``scipy.io``
(no pickle involved) and is run once, in a checkout that has the reference:

    python mnpbem_amd/data/extract_sphere256.py /path/to/MNPBEM
"""
import os
import sys

import numpy as np
import scipy.io as sio


def main(root):
    mat = os.path.join(root, "Particles", "particleshapes", "trisphere.mat")
    s = sio.loadmat(mat, variable_names=["sphere256"])["sphere256"]
    pts = np.stack([s["x"][0, 0].ravel(), s["y"][0, 0].ravel(),
                    s["z"][0, 0].ravel()], axis=1)
    assert pts.shape == (256, 3) and pts.dtype == np.float32
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sphere256.npy")
    np.save(out, pts, allow_pickle=False)
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
