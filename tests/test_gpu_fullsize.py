"""Parity at the workloads' real sizes.

- 20,480-face prolate ellipsoid (10 x 10 x 40 nm, Drude gold in eps 1.77, pols
  x along z and z along x): the mixed solver (FP32 chain + FP64 refinement)
  against the FP64 drop-in at two wavelengths, cross sections to 1e-6 (the
  precision contract of the mixed path; measured about 1e-11).
- 81,920-face sphere of diameter 150 nm (Drude gold in eps 1, a dipole 5 nm
  above the surface, seeded direction and moment): the H-matrix iterative
  solver (cleaf 32, htol 1e-6, GMRES tol 1e-6, FP32 H-matrix storage) against
  Mie theory at three wavelengths.  A sphere's decay rates are the radial and
  tangential ones weighted by cos^2 and sin^2 of the angle between moment and
  position; the tolerances are the discretisation error of the faceted sphere
  (total rate) and of the 256-point sphere the radiated power is integrated
  over (radiative rate).
"""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import dipole_oracle as D
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)


def test_mixed_matches_fp64_at_20480_faces():
    import torch

    from mnpbem_amd import bem as B
    v, f = shapes.ellipsoid(5, (10.0, 10.0, 40.0))
    op = B.bemoptions(sim="ret")
    p = B.comparticle([B.epsconst(1.77), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, op)
    exc = B.planewave([[1.0, 0, 0], [0, 0, 1.0]], [[0, 0, 1.0], [1.0, 0, 0]], op)
    lams = (560.0, 910.0)
    res = {}
    for prec in ("fp64", "mixed"):
        bem = B.bemsolver(p, op, precision=prec)
        res[prec] = []
        for lam in lams:
            print(f"[fullsize 20480] {prec} enei {lam}", flush=True)
            res[prec].append(np.array(exc.cross_sections(bem.solve(exc(p, lam)))))
            if prec == "mixed":
                it, rr = bem.stats()
                print(f"  refinement sweeps {it}, relres {rr:.2e}", flush=True)
                assert rr < 1e-12
        del bem
        torch.cuda.empty_cache()
    for a, b, lam in zip(res["mixed"], res["fp64"], lams):
        err = np.abs(a[:2] - b[:2]) / np.abs(b[:2])
        print(f"[fullsize 20480] {lam} nm: ext {b[0]}, sca {b[1]}; mixed rel err {err.max():.2e}", flush=True)
        assert err.max() < 1e-6
        assert np.all(b[0] > 0) and np.all(b[1] > 0)


def test_hmatrix_decay_rates_at_81920_faces_against_mie():
    from mnpbem_amd import bem as B
    v, f = shapes.icosphere(6, 75.0)
    rng = np.random.default_rng(2024)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    dip = rng.normal(size=3)
    dip /= np.linalg.norm(dip)
    op = B.bemoptions(sim="ret", iter=dict(cleaf=32, htol=1e-6, kmax=[4, 100], tol=1e-6, hprecision="fp32"))
    p = B.comparticle([B.epsconst(1.0), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, op)
    it = B.bemsolver(p, op)
    exc = B.dipoleret(80.0 * u[None], dip[None])
    c2 = float(np.dot(u, dip)) ** 2
    for lam in (500.0, 600.0, 700.0):
        print(f"[fullsize 81920] enei {lam}", flush=True)
        sig = it.solve(exc(p, lam))
        st = it.stats()
        tot, rad, _ = exc.decayrate(sig)
        # 120 multipoles: the series falls off like (75/80)^(2l) (20, the
        # reference's mieret default, leaves the total rate 40% short)
        tm, rm = D.mie_decayrate(lam, O.eps_drude(lam, *DRUDE), 1.0, 150.0, [80.0], lmax=120)
        mt = tm[0, 1] * c2 + tm[0, 0] * (1 - c2)  # z = radial, x = tangential
        mr = rm[0, 1] * c2 + rm[0, 0] * (1 - c2)
        et, er = abs(tot[0, 0] / mt - 1), abs(rad[0, 0] / mr - 1)
        print(f"  flag {st['flag']}, iterations {st['iter']}, relres {st['relres']:.2e}; tot {tot[0, 0]:.5g} "
              f"(Mie {mt:.5g}, rel diff {et:.2e}), rad {rad[0, 0]:.5g} (Mie {mr:.5g}, rel diff {er:.2e})",
              flush=True)
        assert st["flag"] == 0
        assert et < 0.01 and er < 0.02  # rad: the 256-point sphere of the far-field integral
