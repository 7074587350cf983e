"""GPU checks of the adjoint: the transposed solve against a dense solve of
A^T lam = c from the oracle's matrices, and the gradient of the weighted
extinction with respect to the vertex positions against the oracle's
adjoint (oracle/adjoint_oracle.py, itself pinned by central differences),
central differences of the GPU FP64 solver (5,120-face ellipsoid of
semi-axes 10 x 10 x 40 nm in eps 1.77 at 650 nm, 64 seeded directions) and
of the oracle (1,280 faces); the mixed solver's gradient against the FP64
one; the eps1 derivative of the adjoint against the tangent solve; and the
autograd interface."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import adjoint_oracle as A
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

DRUDE = (10.0, 9.065, 0.0708)
POL = np.array([[1.0, 0, 0], [0, 0, 1.0]])
DIRS = np.array([[0, 0, 1.0], [1.0, 0, 0]])


def _solver(v, f, eps_out, precision):
    from mnpbem_amd import bem as B
    op = B.bemoptions(sim="ret")
    p = B.comparticle([B.epsconst(eps_out), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, op)
    return B, p, B.bemsolver(p, op, precision=precision), B.planewave(POL, DIRS, op)


def _pairs(p):
    import ctypes as C
    from mnpbem_amd import _lib
    rows = np.zeros(p.green.nnz, np.int64)
    cols = np.zeros(p.green.nnz, np.int64)
    _lib.call("mnp_greenret_refined", p.green.handle, rows.ctypes.data_as(C.c_void_p),
              cols.ctypes.data_as(C.c_void_p), None, None)
    return rows, cols


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_transposed_solve_matches_dense_oracle(precision):
    """320-face ellipsoid, random right-hand sides: lam of the transposed solve
    satisfies the oracle's dense A^T lam = c to 1e-9 (FP64) / 1e-9 after
    refinement (mixed)."""
    from mnpbem_amd import device as dev
    from mnpbem_amd.sensitivity import solve_transposed
    v, f = shapes.ellipsoid(2, (10.0, 10.0, 25.0))
    enei, eps_out = 600.0, 1.77
    B, p, bem, exc = _solver(v, f, eps_out, precision)
    bem(enei)
    n, P = p.n, 2
    rng = np.random.default_rng(5)
    c = rng.normal(size=(n, 8 * P)) + 1j * rng.normal(size=(n, 8 * P))
    lam = dev.to_host(solve_transposed(bem, dev.to_device_colmajor(c)))
    g = O.GreenRet(v, f)
    ob = O.BemRet(g).initmat(enei, O.eps_drude(enei, *DRUDE), eps_out)
    At = A.system_matrix(ob).T
    xcol = lambda q: [q] + [P + t + 3 * q for t in range(3)] + [4 * P + q] + [5 * P + t + 3 * q for t in range(3)]
    for q in range(P):
        cq = np.concatenate([c[:, j] for j in xcol(q)])
        want = np.linalg.solve(At, cq)
        got = np.concatenate([lam[:, j] for j in xcol(q)])  # the layout of b has the same column pattern
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"[transposed solve {precision}] pol {q}: max rel err {err:.2e}", flush=True)
        assert err < 1e-9
    if precision == "mixed":
        it, rr = bem.stats()
        print(f"  refinement sweeps {it}, relres {rr:.2e}", flush=True)
        assert rr < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gradient_matches_oracle_adjoint(precision):
    """320-face ellipsoid: the GPU vertex gradient equals the oracle's adjoint
    to 1e-8 of its largest entry."""
    from mnpbem_amd.sensitivity import extinction_gradient
    v, f = shapes.ellipsoid(2, (10.0, 10.0, 25.0))
    enei, eps_out = 600.0, 1.77
    w = np.array([1.0, 0.7])
    B, p, bem, exc = _solver(v, f, eps_out, precision)
    sig = bem.solve(exc(p, enei))
    dv, _ = extinction_gradient(exc, bem, sig, w)
    want, ext = A.extinction_gradient(v, f, enei, O.eps_drude(enei, *DRUDE), eps_out, POL, DIRS, w)
    err = np.abs(dv - want).max() / np.abs(want).max()
    print(f"[gradient vs oracle adjoint {precision}] |grad| {np.abs(want).max():.4e}, max rel err {err:.2e}",
          flush=True)
    assert err < 1e-8


def test_gradient_matches_central_differences_5120():
    """5,120-face ellipsoid 10 x 10 x 40 nm in eps 1.77, 650 nm, weights
    (1, 0.7): 64 seeded directions through all vertex coordinates (an RMS
    displacement of 1 per vertex), central differences of the GPU FP64 solver
    with a step of 1e-4 nm agree with the adjoint to 1e-5 of the largest
    directional derivative.  The moved meshes keep the refined pairs of the
    unmoved one (comparticle.moved): at this mesh's 588,775 pairs a
    displacement of 1e-5 nm already moves some across the cutoff, a jump that
    the derivative of the solver at fixed selection does not see."""
    from mnpbem_amd.sensitivity import extinction_gradient
    v, f = shapes.ellipsoid(4, (10.0, 10.0, 40.0))
    enei, eps_out = 650.0, 1.77
    w = np.array([1.0, 0.7])
    B, p, bem, exc = _solver(v, f, eps_out, "fp64")
    sig = bem.solve(exc(p, enei))
    dv, _ = extinction_gradient(exc, bem, sig, w)
    rows0, cols0 = _pairs(p)

    def ext_moved(verts):
        q = p.moved(verts)
        b = B.bemsolver(q, q.op, precision="fp64")
        r, c = _pairs(q)
        assert np.array_equal(r, rows0) and np.array_equal(c, cols0)
        return float(np.dot(w, exc.ext(b.solve(exc(q, enei)))))
    rng = np.random.default_rng(2025)
    h = 1e-4
    an, fd = [], []
    for j in range(64):
        u = rng.normal(size=v.shape)
        u *= np.sqrt(len(v)) / np.linalg.norm(u)
        fd.append((ext_moved(v + h * u) - ext_moved(v - h * u)) / (2 * h))
        an.append(float(np.sum(dv * u)))
    an, fd = np.array(an), np.array(fd)
    err = np.abs(an - fd) / np.abs(an).max()
    print(f"[gradient vs FD 5120] max |d| {np.abs(an).max():.4e}, max err {err.max():.2e}, "
          f"median {np.median(err):.2e}", flush=True)
    assert err.max() < 1e-5


def test_gradient_matches_oracle_differences_1280():
    """1,280-face ellipsoid: the GPU FP64 gradient along seeded directions
    against central differences of the oracle's forward solve, to 1e-6."""
    from mnpbem_amd.sensitivity import extinction_gradient
    v, f = shapes.ellipsoid(3, (10.0, 10.0, 40.0))
    enei, eps_out = 650.0, 1.77
    w = np.array([1.0, 0.7])
    B, p, bem, exc = _solver(v, f, eps_out, "fp64")
    dv, _ = extinction_gradient(exc, bem, bem.solve(exc(p, enei)), w)
    eps1 = O.eps_drude(enei, *DRUDE)
    rng = np.random.default_rng(77)
    h = 5e-4
    for j in range(3):
        u = rng.normal(size=v.shape)
        u /= np.linalg.norm(u)
        ep = A.weighted_extinction(v + h * u, f, enei, eps1, eps_out, POL, DIRS, w)
        em = A.weighted_extinction(v - h * u, f, enei, eps1, eps_out, POL, DIRS, w)
        fd = (ep - em) / (2 * h)
        an = float(np.sum(dv * u))
        print(f"[gradient vs oracle FD 1280] direction {j}: adjoint {an:.10e}, oracle FD {fd:.10e}", flush=True)
        assert abs(an - fd) <= 1e-6 * abs(fd)


def test_mixed_gradient_and_eps_derivative_5120():
    """5,120 faces, 650 nm: the mixed solver's vertex gradient equals the FP64
    one to 1e-6; the adjoint's eps1 derivative equals the tangent solve's."""
    from mnpbem_amd.sensitivity import extinction_eps1_derivative, extinction_gradient
    v, f = shapes.ellipsoid(4, (10.0, 10.0, 40.0))
    enei, eps_out = 650.0, 1.77
    w = np.array([1.0, 0.7])
    res = {}
    for prec in ("fp64", "mixed"):
        B, p, bem, exc = _solver(v, f, eps_out, prec)
        sig = bem.solve(exc(p, enei))
        dv, de = extinction_gradient(exc, bem, sig, w)
        dr, di = extinction_eps1_derivative(exc, bem, sig)
        res[prec] = (dv, de, np.array([np.dot(w, dr), np.dot(w, di)]))
    dv64, de64, tan64 = res["fp64"]
    err = np.abs(res["mixed"][0] - dv64).max() / np.abs(dv64).max()
    print(f"[mixed gradient 5120] max rel diff vs FP64 {err:.2e}; eps1 adjoint {de64}, tangent {tan64}", flush=True)
    assert err < 1e-6
    assert np.allclose(de64, tan64, rtol=1e-8)
    assert np.allclose(res["mixed"][1], tan64, rtol=1e-8)


def test_autograd_interface():
    """ExtinctionAdjoint: the forward is the extinction, the backward the
    weighted Jacobians (vertices and Drude parameters); the Drude part equals
    drude_extinction_gradient's."""
    import torch

    from mnpbem_amd import bem as B
    from mnpbem_amd.sensitivity import ExtinctionAdjoint, drude_extinction_gradient, extinction_gradient
    v, f = shapes.ellipsoid(2, (10.0, 10.0, 25.0))
    enei, eps_out = 600.0, 1.77
    verts = torch.tensor(v, dtype=torch.float64, requires_grad=True)
    params = torch.tensor(DRUDE, dtype=torch.float64, requires_grad=True)
    ext = ExtinctionAdjoint.apply(verts, params, f, POL, DIRS, enei, eps_out, "fp64")
    g = torch.tensor([0.4, -1.3], dtype=torch.float64)
    (ext * g).sum().backward()
    _, p, bem, exc = _solver(v, f, eps_out, "fp64")
    sig = bem.solve(exc(p, enei))
    assert np.allclose(ext.detach().numpy(), exc.ext(sig), rtol=1e-12)
    dv, _ = extinction_gradient(exc, bem, sig, g.numpy())
    assert np.allclose(verts.grad.numpy(), dv, rtol=1e-10, atol=1e-12 * np.abs(dv).max())
    _, gd = drude_extinction_gradient(exc, bem, enei)
    assert np.allclose(params.grad.numpy(), g.numpy() @ gd, rtol=1e-7)
