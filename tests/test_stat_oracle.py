"""Pins for the quasistatic oracle: BEM vs the quasistatic Mie sphere with
mesh refinement, and the retarded solver's small-particle limit."""
import numpy as np

from mnpbem_amd import shapes
from oracle import mnpbem_oracle as O
from oracle.mie import miestat_cross_sections

DRUDE = (10.0, 9.065, 0.0708)


def _stat(level, radius, enei, eps_out=1.0, pol=((1.0, 0, 0),)):
    v, f = shapes.icosphere(level, radius)
    g = O.GreenRet(v, f)
    eps_in = O.eps_drude(enei, *DRUDE)
    bem = O.BemStat(g).initmat(enei, eps_in, eps_out)
    phip = O.planewavestat_potential(g.nvec, pol)
    sig = bem.solve(phip)
    return g, sig, O.stat_cross_sections(g.pos, g.area, sig, pol, enei, eps_out), eps_in


def test_bemstat_converges_to_miestat():
    enei, eps_out = 520.0, 1.77
    errs = []
    for level in (1, 2, 3):
        _, _, (ext, sca, ab), eps_in = _stat(level, 10.0, enei, eps_out)
        want = miestat_cross_sections(enei, eps_in, eps_out, 20.0)
        errs.append(abs(ext[0] / want[0][0] - 1))
        assert np.isclose(ext[0], sca[0] + ab[0])
    assert errs[2] < errs[1] < errs[0] and errs[2] < 0.05, errs


def test_retarded_and_quasistatic_meet_for_small_particles():
    """A 2 nm sphere at 600 nm ((k a)^2 ~ 1e-4): the retarded and quasistatic
    BEM extinctions differ only by their discretisation errors, so the gap
    shrinks with the mesh (3.7%, 3.1%, 1.5% at levels 1-3)."""
    enei, eps_out, pol, dirs = 600.0, 1.0, np.array([[1.0, 0, 0]]), np.array([[0, 0, 1.0]])
    gaps = []
    for level in (1, 2, 3):
        g, _, (ext_s, _, _), eps_in = _stat(level, 1.0, enei, eps_out, pol)
        bem = O.BemRet(g).initmat(enei, eps_in, eps_out)
        a2, a2p = O.planewave_potential(g.pos, g.nvec, pol, dirs, enei, eps_out)
        sig = bem.solve({"a2": a2, "a2p": a2p})
        ext_r = O.extinction(g.pos, g.area, sig, pol, dirs, enei, eps_out)
        gaps.append(abs(ext_r[0] / ext_s[0] - 1))
    assert gaps[2] < gaps[1] < gaps[0] and gaps[2] < 0.02, gaps


def test_static_F_is_retarded_F_at_k0():
    v, f = shapes.icosphere(1, 10.0)
    g = O.GreenRet(v, f)
    F = O.BemStat(g).F
    # closed-surface sum rule of the static operator: sum_row area_row F_ri / area_i = -2 pi
    s = (g.area[:, None] * F).sum(axis=0) / g.area
    assert np.allclose(s, -2 * np.pi, rtol=1e-12)
    assert np.allclose(F.imag, 0)
