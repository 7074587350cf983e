"""CPU pins of the oracle's compound-particle restatement
(oracle/compound_oracle.py): coated-sphere Mie theory reduces to the
homogeneous sphere; a one-part compound is the single-particle solver; a
dimer far apart extinguishes twice one particle; a core-shell sphere
converges to coated-sphere Mie theory, and only with the full L1, L2
matrices of the nested-media branch."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import compound_oracle as CO
from oracle import mie as M
from oracle import mnpbem_oracle as O

DRUDE = (10.0, 9.065, 0.0708)
POL, DIR = np.array([[1.0, 0, 0]]), np.array([[0, 0, 1.0]])
EPS_OUT, EPS_SILICA = 1.77, 2.04


@pytest.fixture(scope="module")
def pinf():
    v, f = shapes.trisphere256(2.0)
    _, _, _, n, a = O.norm_flat(v, f)
    return n, a


def test_coated_mie_reductions():
    lam = 600.0
    eg = O.eps_drude(lam, *DRUDE)
    # core and shell of one material: the homogeneous sphere of the outer radius
    a = CO.coated_mie_cross_sections(lam, eg, eg, EPS_OUT, 6.0, 10.0, 30)
    b = M.mie_cross_sections(lam, eg, EPS_OUT, 20.0, 30)
    assert np.allclose(a, [x[0] for x in b], rtol=1e-12)
    # a shell of the medium: the homogeneous core sphere
    a = CO.coated_mie_cross_sections(lam, eg, EPS_OUT, EPS_OUT, 6.0, 10.0, 30)
    b = M.mie_cross_sections(lam, eg, EPS_OUT, 12.0, 30)
    assert np.allclose(a, [x[0] for x in b], rtol=1e-12)
    # a vanishing core: the homogeneous shell-material sphere
    a = CO.coated_mie_cross_sections(lam, EPS_SILICA, eg, EPS_OUT, 1e-4, 10.0, 30)
    b = M.mie_cross_sections(lam, eg, EPS_OUT, 20.0, 30)
    assert np.allclose(a, [x[0] for x in b], rtol=1e-9)


def test_one_part_compound_is_the_single_particle_solver():
    v, f = shapes.icosphere(2, 5.0)
    lam = 520.0
    eg = O.eps_drude(lam, *DRUDE)
    g = CO.CompGreenRet([(v, f)], [[2, 1]], [1])
    sig = CO.BemRetCompound(g).initmat(lam, [EPS_OUT, eg]).solve(
        CO.planewave_excitation(g, POL, DIR, lam, EPS_OUT))
    g0 = O.GreenRet(v, f)
    a2, a2p = O.planewave_potential(g0.pos, g0.nvec, POL, DIR, lam, EPS_OUT)
    ref = O.BemRet(g0).initmat(lam, eg, EPS_OUT).solve({"a2": a2, "a2p": a2p})
    for k in ref:
        assert np.abs(sig[k] - ref[k]).max() < 1e-12 * np.abs(ref[k]).max()


def test_dimer_far_apart_is_two_particles(pinf):
    v, f = shapes.icosphere(2, 5.0)
    lam = 520.0
    eg = O.eps_drude(lam, *DRUDE)
    single = CO.CompGreenRet([(v, f)], [[2, 1]], [1])
    s1 = CO.BemRetCompound(single).initmat(lam, [EPS_OUT, eg]).solve(
        CO.planewave_excitation(single, POL, DIR, lam, EPS_OUT))
    e1 = CO.cross_sections(single, s1, POL, DIR, lam, EPS_OUT, *pinf)[0][0]
    res = {}
    for inout, eps in (([[2, 1], [2, 1]], [EPS_OUT, eg]), ([[2, 1], [3, 1]], [EPS_OUT, eg, eg])):
        g = CO.CompGreenRet([(v, f), (v + [300.0, 0, 0], f)], inout, [1, 2])
        bem = CO.BemRetCompound(g)
        assert not bem.full_L
        sig = bem.initmat(lam, eps).solve(CO.planewave_excitation(g, POL, DIR, lam, EPS_OUT))
        res[str(inout)] = CO.cross_sections(g, sig, POL, DIR, lam, EPS_OUT, *pinf)[0][0]
        assert abs(res[str(inout)] / (2 * e1) - 1) < 2e-3
    # one medium for both insides, or one each: the same physics
    a, b = res.values()
    assert abs(a / b - 1) < 1e-4


def test_core_shell_converges_to_coated_mie(pinf):
    """Silica core (r 5 nm) in a gold shell (r 10 nm) in water at 550 nm: the
    extinction approaches coated-sphere Mie theory with mesh refinement; the
    nested-media branch needs the full L matrices (diagonal ones are off by
    more than half)."""
    lam = 550.0
    eg = O.eps_drude(lam, *DRUDE)
    eps = [EPS_OUT, eg, EPS_SILICA]
    mie = CO.coated_mie_cross_sections(lam, EPS_SILICA, eg, EPS_OUT, 5.0, 10.0)
    err = {}
    for lev in (2, 3):
        vc, fc = shapes.icosphere(lev, 5.0)
        vs, fs = shapes.icosphere(lev, 10.0)
        g = CO.CompGreenRet([(vc, fc), (vs, fs)], [[3, 2], [2, 1]], [1, 2])
        for full in ((True, False) if lev == 2 else (True,)):
            bem = CO.BemRetCompound(g)
            assert bem.full_L
            bem.full_L = full
            sig = bem.initmat(lam, eps).solve(CO.planewave_excitation(g, POL, DIR, lam, EPS_OUT))
            ext, sca, _ = CO.cross_sections(g, sig, POL, DIR, lam, EPS_OUT, *pinf)
            err[(lev, full)] = (ext[0] / mie[0] - 1, sca[0] / mie[1] - 1)
    assert abs(err[(2, False)][0]) > 0.5
    assert abs(err[(3, True)][0]) < 0.04 and abs(err[(3, True)][1]) < 0.07
    assert abs(err[(3, True)][0]) < abs(err[(2, True)][0])


@pytest.mark.parametrize("nested", [True, False])
def test_solver_inverts_its_operator(nested):
    """mldivide's solution satisfies the operator the mixed solvers refine
    with: afun.m's form (eps on the potentials) without nesting, eps on the
    charges with the full L1, L2 of nested media (the row form leaves a
    residual of order one there)."""
    vc, fc = shapes.icosphere(2, 5.0)
    vs, fs = shapes.icosphere(2, 10.0)
    parts = [(vc, fc), (vs, fs)] if nested else [(vc, fc), (vc + [12.0, 0, 0], fc)]
    inout = [[3, 2], [2, 1]] if nested else [[2, 1], [3, 1]]
    g = CO.CompGreenRet(parts, inout, [1, 2])
    bem = CO.BemRetCompound(g).initmat(520.0, [1.0, 2.0, 3.0])
    assert bem.full_L == nested
    exc = CO.planewave_excitation(g, POL, DIR, 520.0, 1.0)
    sig = bem.solve(exc)
    phi, a, alpha, De = bem.excitation(exc)
    got = bem.afun(sig)
    for x, want in zip(got[1:], (a, De, alpha)):
        assert np.abs(x - want).max() < 1e-11 * np.abs(want).max()
    if nested:
        bem.full_L = False  # the row form
        assert np.abs(bem.afun(sig)[2] - De).max() > 1e-2 * np.abs(De).max()
