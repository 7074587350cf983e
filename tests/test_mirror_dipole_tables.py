"""Host tables of dipolestatmirror/dipoleretmirror (the mirror dipole sets
and their signed combinations with symmetry keys), pinned on the CPU: the
potential of each combination on the full particle has, on copy j, the
symmetry value of its z-key (scalar) times its value on the part, and the
combinations average back to the unit dipoles at the points."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from mnpbem_amd.bem import _SYMVALUE, dipolestatmirror
from oracle import dipole_oracle as D
from oracle import mirror_oracle as M
from oracle import mnpbem_oracle as O

SEG = {"xy": (0.0, np.pi / 2), "y": (0.0, np.pi), "x": (-np.pi / 2, np.pi / 2)}


@pytest.mark.parametrize("sym", ["x", "y", "xy"])
def test_mirror_dipole_combinations_have_their_symmetry(sym):
    v, f = shapes.ellipsoid_segment(4 if sym == "xy" else 8, 8, (10.0, 12.0, 20.0), *SEG[sym])
    V, F = M.full_mesh(v, f, sym)
    g = O.GreenRet(V, F)
    n, J = len(f), len(M.SYMTABLE[sym])
    pts = np.array([[3.0, 4.0, 24.0], [14.0, 2.0, 3.0]])
    phis = []
    for axes, sgn in dipolestatmirror._SETS[sym]:
        p = pts.copy()
        for a in axes:
            p[:, a] = -p[:, a]
        dip = np.repeat(np.diag(np.asarray(sgn, float))[None, :, :].transpose(0, 2, 1), len(pts), axis=0)
        phis.append(D.dipolestat_potential(g.nvec, D.dipolestat_field(g.pos, p, dip, 1.0)))
    total = 0
    for signs, keys in dipolestatmirror._COMB[sym]:
        phi = sum(s * ph for s, ph in zip(signs, phis))
        zval = _SYMVALUE[keys[-1]]
        for j in range(J):
            assert np.allclose(phi[j * n:(j + 1) * n], zval[j] * phi[:n], rtol=1e-12, atol=1e-12 * np.abs(phi).max())
        total = total + phi
    assert np.allclose(total / J, phis[0], rtol=1e-13, atol=1e-15)
