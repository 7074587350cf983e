import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def pinfty():
    """Unit-sphere integration surface trisphere(256, 2): nvec and area."""
    from mnpbem_amd import shapes
    from oracle import mnpbem_oracle as O
    v, f = shapes.trisphere256(2.0)
    _, _, _, n, a = O.norm_flat(v, f)
    return n, a
