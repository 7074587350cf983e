"""Host logic of meshfield: the mesh expansion of meshfield/init.m."""
import numpy as np

from mnpbem_amd.meshfield import _expand


def test_expand_rules():
    x, z = np.meshgrid(np.linspace(-2, 2, 5), np.linspace(-3, 3, 7))
    a, b, c = _expand(x, 0.5, z)
    assert a.shape == b.shape == c.shape == (7, 5) and np.all(b == 0.5) and np.array_equal(c, z)
    x, y = np.meshgrid(np.linspace(-2, 2, 5), np.linspace(-3, 3, 7))
    a, b, c = _expand(x, y, np.array([1.0, 2.0, 3.0]))
    assert a.shape == (7, 5, 3) and np.array_equal(c[2, 1], [1.0, 2.0, 3.0]) and np.array_equal(a[..., 2], x)
    y2, z2 = np.meshgrid(np.linspace(0, 1, 4), np.linspace(0, 1, 6))
    a, b, c = _expand(np.array([5.0]), y2, z2)
    assert a.shape == (6, 4) and np.all(a == 5.0) and np.array_equal(b, y2)
