"""GPU parity of the quasistatic solver (FP64 drop-in and FP32 + FP64
refinement) and the quasistatic plane wave against the oracle and the
quasistatic Mie sphere, through the C ABI."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import mnpbem_oracle as O
from oracle.mie import miestat_cross_sections

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)


def _particle(mesh, eps_out):
    from mnpbem_amd import bem as B
    op = B.bemoptions(sim="stat")
    return B, B.comparticle([B.epsconst(eps_out), B.epsdrude(*DRUDE)], [mesh], [2, 1], 1, op), op


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["sphere_l2", "ellipsoid_l3"])
def test_bemstat_matches_oracle(case, precision):
    from mnpbem_amd import device as dev
    mesh = shapes.icosphere(2, 10.0) if case == "sphere_l2" else shapes.ellipsoid(3, (10.0, 10.0, 30.0))
    eps_out, enei = 1.77, 560.0
    pol = np.array([[1.0, 0, 0], [0, 0, 1.0]])
    print(f"[bemstat parity] {case} {precision} n={len(mesh[1])}", flush=True)
    B, p, op = _particle(mesh, eps_out)
    bem = B.bemsolver(p, op, precision=precision)
    assert isinstance(bem, B.bemstat)
    exc = B.planewave(pol, op=op)
    pot = exc(p, enei)
    g = O.GreenRet(*mesh)
    phip = O.planewavestat_potential(g.nvec, pol)
    assert np.allclose(dev.to_host(pot.phip), phip, rtol=1e-14, atol=1e-15)
    sig = bem.solve(pot)
    ob = O.BemStat(g).initmat(enei, O.eps_drude(enei, *DRUDE), eps_out)
    want = ob.solve(phip)
    got = dev.to_host(sig.sig)
    err = np.abs(got - want).max() / np.abs(want).max()
    cs = exc.cross_sections(sig)
    cs_ref = O.stat_cross_sections(g.pos, g.area, want, pol, enei, eps_out)
    cerr = max(np.abs(a / b - 1).max() for a, b in zip(cs, cs_ref))
    print(f"  sig rel err {err:.2e}, cross sections rel err {cerr:.2e}, stats {bem.stats()}", flush=True)
    assert err < 1e-10 and cerr < 1e-10


def test_bemstat_sphere_vs_miestat():
    """Quasistatic sphere, 1,280 faces: within the discretisation error of the
    quasistatic Mie result; the 20,480-face mesh exercises the 256-block path."""
    enei, eps_out = 520.0, 1.0
    errs = []
    for level in (3, 5):
        B, p, op = _particle(shapes.icosphere(level, 10.0), eps_out)
        bem = B.bemsolver(p, op, precision="mixed")
        exc = B.planewave([[0, 1.0, 0]], op=op)
        ext, sca, ab = exc.cross_sections(bem.solve(exc(p, enei)))
        want = miestat_cross_sections(enei, O.eps_drude(enei, *DRUDE), eps_out, 20.0)
        errs.append(abs(ext[0] / want[0][0] - 1))
        print(f"[bemstat vs miestat] level {level}: ext {ext[0]:.6g} vs {want[0][0]:.6g}, sweeps {bem.stats()}",
              flush=True)
    assert errs[1] < errs[0] < 0.05


def test_bemstat_mixed_bf16_matches_fp64_5120_faces():
    """Quasistatic ellipsoid, 5,120 faces: the split-bf16 factorisation with
    FP64 refinement reproduces the FP64 solver's cross sections."""
    B, p, op = _particle(shapes.ellipsoid(4, (10.0, 10.0, 40.0)), 1.77)
    exc = B.planewave([[1.0, 0, 0], [0, 0, 1.0]], op=op)
    b64 = B.bemsolver(p, op, precision="fp64")
    bbf = B.bemsolver(p, op, precision="mixed_bf16")
    for lam in (550.0, 800.0):
        print(f"[bemstat mixed_bf16] enei {lam}", flush=True)
        cb = exc.cross_sections(bbf.solve(exc(p, lam)))
        c64 = exc.cross_sections(b64.solve(exc(p, lam)))
        it, rr = bbf.stats()
        print(f"  sweeps {it}, relres {rr:.2e}", flush=True)
        assert rr < 1e-12 and it <= 6
        for a, b in zip(cb, c64):
            assert np.allclose(a, b, rtol=1e-10)
