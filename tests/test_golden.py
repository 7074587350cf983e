"""The oracle reproduces its committed fixtures (tests/golden/oracle_small.npz)."""
import os

import numpy as np

from tests.golden.make_golden import compute

HERE = os.path.dirname(os.path.abspath(__file__))


def test_oracle_matches_fixtures():
    want = np.load(os.path.join(HERE, "golden", "oracle_small.npz"), allow_pickle=False)
    got = compute()
    assert sorted(got) == sorted(want.files)
    for key in want.files:
        w, g = want[key], got[key]
        if np.issubdtype(w.dtype, np.integer):
            assert np.array_equal(g, w), key
        else:
            assert np.allclose(g, w, rtol=1e-12, atol=1e-14 * np.abs(w).max()), key
