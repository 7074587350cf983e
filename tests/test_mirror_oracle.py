"""Pins for the mirror-symmetry oracle: the full particle's Green matrices
are equivariant under the mirror group, and the mirror solver's expanded
solution equals the dense solution on the full particle."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import mirror_oracle as M
from oracle import mnpbem_oracle as O

DRUDE = (10.0, 9.065, 0.0708)
SEGMENT = {"xy": (0.0, np.pi / 2), "y": (0.0, np.pi), "x": (-np.pi / 2, np.pi / 2)}


def segment(sym, nphi, ntheta, axes=(10.0, 12.0, 20.0)):
    lo, hi = SEGMENT[sym]
    return shapes.ellipsoid_segment(nphi, ntheta, axes, lo, hi)


@pytest.mark.parametrize("sym", ["x", "y", "xy"])
def test_full_mesh_closed_and_equivariant(sym):
    v, f = segment(sym, 3 if sym == "xy" else 6, 6)
    V, F = M.full_mesh(v, f, sym)
    J, n = len(M.SYMTABLE[sym]), len(f)
    assert len(F) == J * n
    g = O.GreenRet(V, F)
    # closed surface: the normals integrate to zero, the enclosed volume is positive
    assert np.abs(g.area @ g.nvec).max() < 1e-9 * g.area.sum()
    assert np.sum(g.area * np.sum(g.pos * g.nvec, axis=1)) > 0
    bits = {"x": [0, 1], "y": [0, 2], "xy": [0, 1, 2, 3]}[sym]  # copy -> (x flip, y flip) bits
    comp = {b: i for i, b in enumerate(bits)}
    for k in (0.31 + 0.02j, 0.0):
        mats = g.eval(k, ("G", "F"))
        for key, A in mats.items():
            for a in range(J):
                for b in range(J):
                    blk = A[a * n:(a + 1) * n, b * n:(b + 1) * n]
                    ref = A[:n, comp[bits[a] ^ bits[b]] * n:(comp[bits[a] ^ bits[b]] + 1) * n]
                    assert np.abs(blk - ref).max() <= 1e-10 * np.abs(ref).max(), (key, a, b)


@pytest.mark.parametrize("sym", ["x", "y", "xy"])
def test_mirror_solution_equals_full_solution(sym):
    v, f = segment(sym, 3 if sym == "xy" else 6, 6)
    V, F = M.full_mesh(v, f, sym)
    n = len(f)
    g = O.GreenRet(V, F)
    enei, eps_out = 560.0, 1.33
    e1 = O.eps_drude(enei, *DRUDE)
    pol = np.array([[1.0, 0, 0], [0, 1.0, 0], [0.6, 0.8, 0]])
    dirs = np.array([[0, 0, 1.0], [0, 0, -1.0], [0, 0, -1.0]])
    a2, a2p = O.planewave_potential(g.pos, g.nvec, pol, dirs, enei, eps_out)
    want = O.BemRet(g).initmat(enei, e1, eps_out).solve({"a2": a2, "a2p": a2p})
    mb = M.BemRetMirror(g, n, sym).initmat(enei, e1, eps_out)
    pw_dirs = np.array([[0, 0, 1.0], [0, 0, -1.0]])
    symvals = M.planewave_symvals(sym)
    sols = []
    for i, sv in enumerate(symvals):
        p2 = np.zeros((2, 3))
        p2[:, i] = 1.0
        b2, b2p = O.planewave_potential(g.pos[:n], g.nvec[:n], p2, pw_dirs, enei, eps_out)
        sx = [M.symindex(sym, key) for key in M.PLANEWAVE_KEYS[sym][i]]
        sols.append(mb.solve({"a2": b2, "a2p": b2p}, sx))
    got = M.full_solution(sols, symvals, pol, dirs)
    for key in ("sig1", "sig2", "h1", "h2"):
        err = np.linalg.norm(got[key] - want[key]) / np.linalg.norm(want[key])
        assert err < 1e-9, (key, err)
    # a plausible slip: the pol-x solution expanded with the pol-y symmetry values
    bad = M.expand(sols[0]["sig2"], symvals[1])[:, 0]
    assert np.linalg.norm(bad - want["sig2"][:, 0]) > 1e-3 * np.linalg.norm(want["sig2"][:, 0])


@pytest.mark.parametrize("sym", ["x", "y", "xy"])
def test_stat_mirror_solution_equals_full_solution(sym):
    """Quasistatic: the expanded mirror charge equals bemstat's on the full
    particle, for pols in the xy plane."""
    v, f = segment(sym, 3 if sym == "xy" else 6, 6)
    V, F = M.full_mesh(v, f, sym)
    n = len(f)
    g = O.GreenRet(V, F)
    enei, eps_out = 600.0, 1.77
    e1 = O.eps_drude(enei, *DRUDE)
    pol = np.array([[1.0, 0, 0], [0, 1.0, 0], [0.6, 0.8, 0]])
    want = O.BemStat(g).initmat(enei, e1, eps_out).solve(O.planewavestat_potential(g.nvec, pol))
    mb = M.BemStatMirror(g, n, sym).initmat(enei, e1, eps_out)
    symvals = M.planewavestat_symvals(sym)
    sols = []
    for i, sv in enumerate(symvals):
        p1 = np.zeros((1, 3))
        p1[0, i] = 1.0
        s = M.symindex(sym, M.PLANEWAVE_KEYS[sym][i][-1])
        sols.append(mb.solve(O.planewavestat_potential(g.nvec[:n], p1), s))
    got = M.full_stat(sols, symvals, pol)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 1e-10, err
    # the reference's z-key '-' for the pol-y excitation of sym 'x' does not
    # reproduce the full particle's charge
    if sym == "x":
        s_bad = M.symindex(sym, "-")
        p1 = np.array([[0, 1.0, 0]])
        bad = M.expand(mb.solve(O.planewavestat_potential(g.nvec[:n], p1), s_bad), symvals[1])
        assert np.linalg.norm(bad[:, 0] - want[:, 1]) > 1e-2 * np.linalg.norm(want[:, 1])
