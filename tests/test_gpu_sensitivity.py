"""GPU checks of the extinction derivatives with respect to the inside
dielectric function and the Drude parameters (one tangent solve with the
forward factors): against central finite differences of the FP64 solver,
of the oracle, and between the FP64 and mixed solvers."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)
POL = np.array([[1.0, 0, 0], [0, 0, 1.0]])
DIRS = np.array([[0, 0, 1.0], [1.0, 0, 0]])


def _setup(mesh, inside, eps_out, precision):
    from mnpbem_amd import bem as B
    op = B.bemoptions(sim="ret")
    p = B.comparticle([B.epsconst(eps_out), inside], [mesh], [2, 1], 1, op)
    return B, p, B.bemsolver(p, op, precision=precision), B.planewave(POL, DIRS, op)


def _ext(B, p, bem, exc, enei):
    bem.enei = None
    return np.asarray(exc.cross_sections(bem.solve(exc(p, enei)))[0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_drude_gradient_matches_finite_differences(precision):
    """1,280-face ellipsoid: d ext / d (eps0, wp, gammad) against central
    differences of the FP64 solver (relative step 1e-6), to 1e-5."""
    from mnpbem_amd.sensitivity import drude_extinction_gradient
    mesh = shapes.ellipsoid(3, (10.0, 10.0, 20.0))
    enei = 620.0
    from mnpbem_amd import bem as Bm
    B, p, bem, exc = _setup(mesh, Bm.epsdrude(*DRUDE), 1.77, precision)
    ext, grad = drude_extinction_gradient(exc, bem, enei)
    print(f"[sensitivity {precision}] ext {ext}, grad {grad.tolist()}", flush=True)
    _, pf, b64, exf = _setup(mesh, B.epsdrude(*DRUDE), 1.77, "fp64")
    for j in range(3):
        h = 1e-6 * DRUDE[j]
        vals = []
        for s in (1, -1):
            par = list(DRUDE)
            par[j] += s * h
            pf.eps[1] = B.epsdrude(*par)
            vals.append(_ext(B, pf, b64, exf, enei))
        fd = (vals[0] - vals[1]) / (2 * h)
        err = np.abs(grad[:, j] - fd) / np.abs(fd).max()
        print(f"  parameter {j}: analytic {grad[:, j]}, FD {fd}, rel diff {err.max():.2e}", flush=True)
        assert err.max() < 1e-5


def test_eps_derivative_matches_oracle_differences():
    """1,280-face sphere with a constant inside eps: d ext / d Re eps1 and
    d ext / d Im eps1 against central differences of the oracle's dense
    solver (step 1e-6), to 1e-5."""
    from mnpbem_amd import bem as B
    from mnpbem_amd.sensitivity import extinction_eps1_derivative
    mesh = shapes.icosphere(3, 10.0)
    enei, eps_in, eps_out = 530.0, -4.2 + 0.9j, 1.0
    _, p, bem, exc = _setup(mesh, B.epsconst(eps_in), eps_out, "fp64")
    sig = bem.solve(exc(p, enei))
    dr, di = extinction_eps1_derivative(exc, bem, sig)
    g = O.GreenRet(*mesh)
    a2, a2p = O.planewave_potential(g.pos, g.nvec, POL, DIRS, enei, eps_out)

    def ext(e):
        s = O.BemRet(g).initmat(enei, e, eps_out).solve({"a2": a2, "a2p": a2p})
        return O.extinction(g.pos, g.area, s, POL, DIRS, enei, eps_out)
    h = 1e-6
    fd_r = (ext(eps_in + h) - ext(eps_in - h)) / (2 * h)
    fd_i = (ext(eps_in + 1j * h) - ext(eps_in - 1j * h)) / (2 * h)
    print(f"[sensitivity eps] d/dRe {dr} vs {fd_r}; d/dIm {di} vs {fd_i}", flush=True)
    scale = max(np.abs(fd_r).max(), np.abs(fd_i).max())
    assert np.all(np.abs(dr - fd_r) <= 1e-5 * scale) and np.all(np.abs(di - fd_i) <= 1e-5 * scale)


def test_autograd_function_and_5120_faces():
    """Config-5 size (5,120-face ellipsoid, 650 nm): the torch autograd
    interface returns the same gradient as drude_extinction_gradient, and the
    mixed solver's gradient matches the FP64 one to 1e-8."""
    import torch
    from mnpbem_amd import bem as B
    from mnpbem_amd.sensitivity import DrudeExtinction, drude_extinction_gradient
    mesh = shapes.ellipsoid(4, (10.0, 10.0, 40.0))
    enei = 650.0
    _, p64, b64, exc = _setup(mesh, B.epsdrude(*DRUDE), 1.77, "fp64")
    ext64, g64 = drude_extinction_gradient(exc, b64, enei)
    _, pmx, bmx, exm = _setup(mesh, B.epsdrude(*DRUDE), 1.77, "mixed")
    extmx, gmx = drude_extinction_gradient(exm, bmx, enei)
    print(f"[sensitivity 5120] ext {ext64}; grad fp64 {g64.tolist()}; mixed rel diff "
          f"{np.abs(gmx - g64).max() / np.abs(g64).max():.2e}", flush=True)
    assert np.allclose(extmx, ext64, rtol=1e-10)
    assert np.abs(gmx - g64).max() <= 1e-8 * np.abs(g64).max()
    params = torch.tensor(DRUDE, dtype=torch.float64, requires_grad=True)
    ext_t = DrudeExtinction.apply(params, bmx, exm, enei)
    w = torch.tensor([0.3, -1.1], dtype=torch.float64)
    (ext_t * w).sum().backward()
    want = w.numpy() @ gmx
    assert np.allclose(params.grad.numpy(), want, rtol=1e-12, atol=0)
