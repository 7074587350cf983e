"""Host-side logic of the Python mirror of MNPBEM's API (no GPU calls)."""
import numpy as np
import pytest

from mnpbem_amd import bem as B, shapes


def test_drude_matches_closed_form():
    d = B.epsdrude(10.0, 9.065, 0.0708)
    for lam in (400.0, 800.0):
        w = 1239.8414059345282 / lam
        eps, k = d(lam)
        assert abs(eps - (10 - 9.065 ** 2 / (w * (w + 0.0708j)))) < 1e-13
        assert abs(k - 2 * np.pi / lam * np.sqrt(eps)) < 1e-15
    # survey example values at 400 and 800 nm
    assert abs(d(400.0)[0] - (1.4514 + 0.1953j)) < 2e-4
    assert abs(d(800.0)[0] - (-24.141 + 1.5597j)) < 2e-3


def test_eps_const_and_fun():
    c = B.epsconst(1.77)
    assert c(500.0)[0] == 1.77 and abs(c(500.0)[1] - 2 * np.pi / 500 * np.sqrt(1.77)) < 1e-15
    f = B.epsfun(lambda lam: 2.0 + 0.1j)
    assert f(600.0)[0] == 2.0 + 0.1j


def test_bemoptions_defaults_and_override():
    op = B.bemoptions()
    assert op["sim"] == "ret" and op["RelCutoff"] == 3.0 and op["order"] == 5
    assert B.bemoptions(order=2)["order"] == 2


def test_comparticle_media():
    v, f = shapes.icosphere(1, 5.0)
    p = B.comparticle([B.epsconst(1.0), B.epsdrude(10.0, 9.065, 0.0708)], [(v, f)], [2, 1], 1)
    assert p.n == 80
    assert p.eps2(500.0) == 1.0
    assert p.eps1(500.0) == B.epsdrude(10.0, 9.065, 0.0708)(500.0)[0]
    with pytest.raises(ValueError):  # one inout row per part
        B.comparticle([B.epsconst(1.0)], [(v, f), (v, f)], [2, 1], 1)
    with pytest.raises(NotImplementedError):
        B.comparticle([B.epsconst(1.0)], [(v, f)], [2, 1], 1, B.bemoptions(interp="curv"))


def test_compound_comparticle_host():
    """Two parts (core-shell): concatenated mesh, face media, closed masks as
    comparticle's closed arguments give them (Particles/@comparticle/closed.m)."""
    v1, f1 = shapes.icosphere(1, 3.0)
    v2, f2 = shapes.icosphere(1, 5.0)
    eps = [B.epsconst(1.0), B.epsdrude(10.0, 9.065, 0.0708), B.epsconst(2.0)]
    p = B.comparticle(eps, [(v1, f1), (v2, f2)], [[3, 2], [2, 1]], 1, 2, B.bemoptions())
    assert p.compound and p.n == 160 and p.part_off == [0, 80, 160]
    assert np.array_equal(p.verts[len(v1):], v2) and np.array_equal(p.faces[80:], f2 + len(v1))
    fio = p.face_inout
    assert np.all(fio[:80] == [3, 2]) and np.all(fio[80:] == [2, 1])
    assert p.closed_mask.tolist() == [1, 2]
    q = B.comparticle(eps, [(v1, f1), (v2, f2)], [[3, 2], [2, 1]], [1, 2])
    assert q.closed_mask.tolist() == [3, 3]
    assert [e for e in p.eps_media(500.0)][2] == 2.0
    with pytest.raises(NotImplementedError):
        p.eps1(500.0)
    with pytest.raises(NotImplementedError):
        B.bemsolver(p, B.bemoptions(sim="stat"))


def test_planewave_rejects_parallel_polarisation():
    with pytest.raises(ValueError):
        B.planewave([[0, 0, 1.0]], [[0, 0, 1.0]])


def test_meshes_outward_and_sizes():
    for lev, nf in ((0, 20), (2, 320), (3, 1280)):
        v, f = shapes.icosphere(lev, 2.0)
        assert len(f) == nf
        n = np.cross(v[f[:, 0]] - v[f[:, 1]], v[f[:, 2]] - v[f[:, 1]])
        c = v[f].mean(axis=1)
        assert np.all(np.sum(n * c, axis=1) > 0)
        assert np.allclose(np.linalg.norm(v, axis=1), 2.0)
    v, f = shapes.ellipsoid(2, (10, 10, 40))
    assert np.allclose((v[:, 0] / 10) ** 2 + (v[:, 1] / 10) ** 2 + (v[:, 2] / 40) ** 2, 1.0)
    pv, pf = shapes.trisphere256()
    assert pv.shape == (256, 3) and len(pf) == 2 * 256 - 4
    assert np.allclose(np.linalg.norm(pv, axis=1), 1.0)


def test_initmat_many_forgets_wavelengths_when_the_call_fails(monkeypatch):
    """mnp_bemret_initmat_many initialises the solvers group by group; when it
    fails part-way, no solver may keep a cached wavelength (it could then be
    taken to hold matrices it no longer holds).  The C call is replaced by one
    that fails, so this runs without a GPU."""
    import ctypes as C
    from mnpbem_amd import _lib

    class FakeP:
        def eps1(self, enei):
            return -2.0 + 0.3j

        def eps2(self, enei):
            return 1.0

    class FakeBem:
        def __init__(self, enei):
            self.p, self.enei, self.handle = FakeP(), enei, C.c_void_p(0x1000)

    bems = [FakeBem(500.0) for _ in range(6)]

    def fail(name, *args):
        assert name == "mnp_bemret_initmat_many"
        raise _lib.MnpError("initmat_many failed in its second group")

    monkeypatch.setattr(_lib, "call", fail)
    with pytest.raises(_lib.MnpError):
        B.initmat_many(bems, 650.0)
    assert all(b.enei is None for b in bems)


def test_solver_forgets_wavelength_when_initmat_fails(monkeypatch):
    """bem(enei) of the dense solver: a failed initmat leaves no cached
    wavelength behind (bemret.__call__)."""
    import ctypes as C
    from mnpbem_amd import _lib

    bem = B.bemret.__new__(B.bemret)
    bem.handle, bem.enei = C.c_void_p(0), 500.0

    class FakeP:
        def eps1(self, enei):
            return -2.0 + 0.3j

        def eps2(self, enei):
            return 1.0
    bem.p = FakeP()

    def fail(name, *args):
        raise _lib.MnpError("initmat failed")

    monkeypatch.setattr(_lib, "call", fail)
    with pytest.raises(_lib.MnpError):
        bem(650.0)
    assert bem.enei is None
