"""Material tables (epstable) against an independent spline: tabulated data
of the reference's Material/@epstable (gold.dat, silver.dat, Johnson and
Christy) are fixtures under tests/golden/materials."""
import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from mnpbem_amd.bem import EV2NM, epstable

HERE = os.path.join(os.path.dirname(__file__), "golden", "materials")


@pytest.mark.parametrize("name", ["gold.dat", "silver.dat"])
def test_epstable_matches_notaknot_spline(name):
    tab = epstable(os.path.join(HERE, name))
    data = np.loadtxt(os.path.join(HERE, name), comments="%")
    ene, n, k = data.T
    # exact at the tabulated wavelengths
    for e, nn, kk in zip(ene[::7], n[::7], k[::7]):
        eps, kw = tab(EV2NM / e)
        assert abs(eps - (nn + 1j * kk) ** 2) <= 1e-12 * abs(eps)
        assert abs(kw - 2 * np.pi / (EV2NM / e) * np.sqrt(eps)) <= 1e-12 * abs(kw)
    x = EV2NM / ene
    o = np.argsort(x)
    sn = CubicSpline(x[o], n[o], bc_type="not-a-knot")
    sk = CubicSpline(x[o], k[o], bc_type="not-a-knot")
    mid = 0.5 * (x[o][1:] + x[o][:-1])
    got = np.array([tab.nk(w) for w in mid])
    assert np.allclose(got, sn(mid) + 1j * sk(mid), rtol=1e-12, atol=1e-12)
    # natural end conditions differ near the ends
    nat = CubicSpline(x[o], n[o], bc_type="natural")
    assert abs(nat(mid[0]) - got[0].real) > 1e-6
    with pytest.raises(ValueError):
        tab(x.min() - 1.0)
