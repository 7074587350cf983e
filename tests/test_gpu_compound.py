"""GPU parity of compound particles (several parts with their own media) in
the dense retarded solver: dimers with one or two inside media, a core-shell
sphere (nested media: the full L1, L2 matrices), against the oracle's
restatement of compgreenret's connectivity and bemret's general branch; a
sphere split into two parts with one closed surface against the single
particle; mixed precision against FP64; the core-shell sphere against
coated-sphere Mie theory at 2,560 faces."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import compound_oracle as CO
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)
POL = np.array([[1.0, 0, 0], [0, 0, 1.0]])
DIR = np.array([[0, 0, 1.0], [1.0, 0, 0]])
EPS_OUT, EPS_SILICA = 1.77, 2.04


def _host(sig):
    from mnpbem_amd import device as dev
    return {k: dev.to_host(sig[k]) for k in ("sig1", "sig2", "h1", "h2")}


def _relerr(a, b):
    return max(np.abs(a[k] - b[k]).max() / np.abs(b[k]).max() for k in b)


def _cases():
    v1, f1 = shapes.icosphere(2, 5.0)
    vc, fc = shapes.icosphere(2, 5.0)
    vs, fs = shapes.icosphere(2, 10.0)
    return {
        "dimer, one inside medium": ([(v1, f1), (v1 + [12.0, 0, 0], f1)], [[2, 1], [2, 1]], (1, 2), ["out", "au"]),
        "dimer, two inside media": ([(v1, f1), (v1 + [12.0, 0, 0], f1)], [[2, 1], [3, 1]], (1, 2),
                                    ["out", "au", "silica"]),
        "core-shell (silica in gold)": ([(vc, fc), (vs, fs)], [[3, 2], [2, 1]], (1, 2), ["out", "au", "silica"]),
    }


def _eps_list(names, lam):
    table = {"out": EPS_OUT, "au": O.eps_drude(lam, *DRUDE), "silica": EPS_SILICA}
    return [table[nm] for nm in names]


def _device_particle(B, parts, inout, closed, names):
    mats = {"out": B.epsconst(EPS_OUT), "au": B.epsdrude(*DRUDE), "silica": B.epsconst(EPS_SILICA)}
    return B.comparticle([mats[nm] for nm in names], parts, inout, *closed, B.bemoptions(sim="ret"))


@pytest.mark.parametrize("case", list(_cases()))
def test_compound_matches_oracle(case, pinfty):
    from mnpbem_amd import bem as B
    parts, inout, closed, names = _cases()[case]
    p = _device_particle(B, parts, inout, closed, names)
    exc = B.planewave(POL, DIR)
    g = CO.CompGreenRet(parts, inout, closed)
    ob = CO.BemRetCompound(g)
    assert ob.full_L == ("core-shell" in case)
    b64 = B.bemret(p, precision="fp64")
    bmx = B.bemret(p, precision="mixed")
    for lam in (520.0, 700.0):
        ref = ob.initmat(lam, _eps_list(names, lam)).solve(CO.planewave_excitation(g, POL, DIR, lam, EPS_OUT))
        cref = CO.cross_sections(g, ref, POL, DIR, lam, EPS_OUT, *pinfty)
        s64 = b64.solve(exc(p, lam))
        c64 = exc.cross_sections(s64)
        e_sig = _relerr(_host(s64), ref)
        e_cs = max(np.max(np.abs(a / b - 1)) for a, b in zip(c64[:2], cref[:2]))
        print(f"[compound] {case} {lam:.0f} nm: fp64 sig vs oracle {e_sig:.2e}, cross sections {e_cs:.2e}",
              flush=True)
        assert e_sig < 1e-10 and e_cs < 1e-10
        cmx = exc.cross_sections(bmx.solve(exc(p, lam)))
        e_mx = max(np.max(np.abs(a / b - 1)) for a, b in zip(cmx[:2], c64[:2]))
        print(f"[compound] {case} {lam:.0f} nm: mixed vs fp64 {e_mx:.2e}, refinement {bmx.stats()}", flush=True)
        assert e_mx < 1e-8


def test_split_sphere_with_one_closed_surface_is_the_sphere():
    """A sphere meshed as two parts (the faces above and below z = 0) with
    the same media and one closed surface [1, 2] is the single particle; with
    a closed surface per part the sum rule covers only each half and the
    solution changes."""
    from mnpbem_amd import bem as B
    v, f = shapes.icosphere(2, 8.0)
    up = v[f].mean(axis=1)[:, 2] >= 0
    parts = [(v, f[up]), (v, f[~up])]
    mats = [B.epsconst(EPS_OUT), B.epsdrude(*DRUDE)]
    single = B.comparticle(mats, [(v, np.vstack([f[up], f[~up]]))], [2, 1], 1)
    whole = B.comparticle(mats, parts, [[2, 1], [2, 1]], [1, 2])
    halves = B.comparticle(mats, parts, [[2, 1], [2, 1]], 1, 2)
    exc = B.planewave(POL, DIR)
    lam = 560.0
    ref = _host(B.bemret(single, precision="fp64").solve(exc(single, lam)))
    got = _host(B.bemret(whole, precision="fp64").solve(exc(whole, lam)))
    other = _host(B.bemret(halves, precision="fp64").solve(exc(halves, lam)))
    e1, e2 = _relerr(got, ref), _relerr(other, ref)
    print(f"[compound] split sphere: one closed surface {e1:.2e}, one per part {e2:.2e}", flush=True)
    assert e1 < 1e-12 and e2 > 1e-4


def test_core_shell_against_coated_mie(pinfty):
    """Silica core (r 5 nm) in a gold shell (r 10 nm), 1,280 faces each, in
    water: extinction and scattering within the discretisation error of
    coated-sphere Mie theory (the oracle shows 3% and 5.4% at 550 nm)."""
    from mnpbem_amd import bem as B
    vc, fc = shapes.icosphere(3, 5.0)
    vs, fs = shapes.icosphere(3, 10.0)
    p = _device_particle(B, [(vc, fc), (vs, fs)], [[3, 2], [2, 1]], (1, 2), ["out", "au", "silica"])
    exc = B.planewave(POL[:1], DIR[:1])
    bem = B.bemret(p, precision="mixed")
    for lam in (450.0, 550.0, 650.0):
        ext, sca, _ = exc.cross_sections(bem.solve(exc(p, lam)))
        mie = CO.coated_mie_cross_sections(lam, EPS_SILICA, O.eps_drude(lam, *DRUDE), EPS_OUT, 5.0, 10.0)
        print(f"[compound] core-shell 2560 faces {lam:.0f} nm: ext {ext[0]:.4g} (Mie {mie[0]:.4g}), "
              f"sca {sca[0]:.4g} (Mie {mie[1]:.4g})", flush=True)
        assert abs(ext[0] / mie[0] - 1) < 0.04 and abs(sca[0] / mie[1] - 1) < 0.07
