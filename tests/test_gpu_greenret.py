"""GPU parity of the Green-function stage (geometry, refinement, tables, G/F/H)
against the oracle, through the C ABI."""
import numpy as np
import pytest

from mnpbem_amd import shapes
from oracle import mnpbem_oracle as O

pytestmark = pytest.mark.gpu

MESHES = {
    "ico1": lambda: shapes.icosphere(1, 10.0),
    "ell2": lambda: shapes.ellipsoid(2, (10.0, 10.0, 40.0)),
    "ico3": lambda: shapes.icosphere(3, 10.0),
}


@pytest.fixture(scope="module", params=list(MESHES))
def pair(request):
    from mnpbem_amd.greenfun import CompGreenRet
    v, f = MESHES[request.param]()
    print(f"[greenret parity] mesh {request.param} n={len(f)}", flush=True)
    return O.GreenRet(v, f), CompGreenRet(v, f)


def test_geometry(pair):
    ref, gpu = pair
    geo = gpu.geometry()
    for key, want in (("pos", ref.pos), ("nvec", ref.nvec), ("tvec1", ref.tvec1),
                      ("tvec2", ref.tvec2)):
        assert np.allclose(geo[key], want, rtol=0, atol=1e-13 * np.abs(want).max()), key
    assert np.allclose(geo["area"], ref.area, rtol=1e-14)


def _report(name, got, want, kind, rows, cols):
    err = np.abs(got - want).reshape(len(kind), -1).max(axis=1)
    rel = err / np.maximum(np.abs(want).reshape(len(kind), -1).max(axis=1), 1e-300)
    for k in (1, 2):
        sel = kind == k
        if sel.any():
            w = np.argmax(np.where(sel, rel, -1))
            print(f"  {name} kind {k}: max rel {rel[sel].max():.3e} at pair {w} "
                  f"({rows[w]},{cols[w]}) got {np.ravel(got[w])[:3]} want {np.ravel(want[w])[:3]}",
                  flush=True)


def test_refined_pairs_and_tables(pair):
    ref, gpu = pair
    rows, cols, g, f = gpu.refined()
    assert gpu.nnz == len(ref.rows)
    assert np.array_equal(rows, ref.rows) and np.array_equal(cols, ref.cols)
    for o in range(ref.order + 1):
        _report(f"g[{o}]", g[:, o], ref.g[:, o], ref.kind, rows, cols)
        _report(f"f[{o}]", f[:, :, o], ref.f[:, :, o], ref.kind, rows, cols)
    # Self terms (kind 2) integrate from the centroid with a polar rule whose
    # innermost ring lies 5e-7 face radii from it: |pos - pt| cancels about
    # nine digits, so 1-ulp differences of the rule's points (SciPy/NumPy vs
    # glibc cos/sin) reach 1e-10 relative.  Near-field pairs are well
    # conditioned.
    self_ = ref.kind == 2
    for o in range(ref.order + 1):
        for sel, rt in ((~self_, 1e-12), (self_, 1e-9)):
            scale = np.abs(ref.g[:, o]).max()
            assert np.allclose(g[sel, o], ref.g[sel, o], rtol=rt, atol=rt * 0.1 * scale), o
            scale = np.abs(ref.f[:, :, o]).max()
            assert np.allclose(f[sel, :, o], ref.f[sel, :, o], rtol=10 * rt, atol=rt * scale), o


@pytest.mark.parametrize("k", [0.08 + 0.0j, 0.12 + 0.031j, 0.3 + 0.9j])
def test_dense_matrices(pair, k):
    from mnpbem_amd import device as dev
    ref, gpu = pair
    want = ref.eval(k, ("G", "F", "H1", "H2"))
    for key in ("G", "F", "H1", "H2"):
        got = dev.to_host(gpu.eval(k, key))
        w = want[key]
        err = np.abs(got - w)
        i, j = np.unravel_index(np.argmax(err / (np.abs(w) + 1e-300)), w.shape)
        print(f"  {key}: max abs err {err.max():.3e}, worst rel at ({i},{j}) "
              f"got {got[i, j]} want {w[i, j]}", flush=True)
        # the diagonal carries the self terms (conditioning: see above)
        rt = np.where(np.eye(len(w), dtype=bool), 1e-9, 1e-12)
        assert np.all(err <= rt * np.abs(w) + 1e-13 * np.abs(w).max()), (key, err.max())
        got32 = dev.to_host(gpu.eval(k, key, precision="fp32"))
        assert np.allclose(got32, w, rtol=2e-6, atol=2e-6 * np.abs(w).max()), key
