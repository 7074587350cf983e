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


def test_refined_pairs_and_tables(pair):
    ref, gpu = pair
    rows, cols, g, f = gpu.refined()
    assert gpu.nnz == len(ref.rows)
    assert np.array_equal(rows, ref.rows) and np.array_equal(cols, ref.cols)
    for o in range(ref.order + 1):
        scale = np.abs(ref.g[:, o]).max()
        assert np.allclose(g[:, o], ref.g[:, o], rtol=1e-12, atol=1e-13 * scale), o
        scale = np.abs(ref.f[:, :, o]).max()
        assert np.allclose(f[:, :, o], ref.f[:, :, o], rtol=1e-11, atol=1e-12 * scale), o


@pytest.mark.parametrize("k", [0.08 + 0.0j, 0.12 + 0.031j, 0.3 + 0.9j])
def test_dense_matrices(pair, k):
    from mnpbem_amd import device as dev
    ref, gpu = pair
    want = ref.eval(k, ("G", "F", "H1", "H2"))
    for key in ("G", "F", "H1", "H2"):
        got = dev.to_host(gpu.eval(k, key))
        w = want[key]
        err = np.abs(got - w)
        assert np.all(err <= 1e-12 * np.abs(w) + 1e-13 * np.abs(w).max()), (key, err.max())
        got32 = dev.to_host(gpu.eval(k, key, precision="fp32"))
        assert np.allclose(got32, w, rtol=2e-6, atol=2e-6 * np.abs(w).max()), key
