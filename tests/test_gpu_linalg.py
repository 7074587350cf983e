"""GPU checks of the dense complex building blocks (GEMM on MFMA, LU with
partial pivoting, triangular solves, inverse) against NumPy/SciPy LAPACK."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(rng, m, n, dtype=np.complex128):
    return (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(dtype)


def _dev(a, dtype):
    import torch
    from mnpbem_amd import device as dev
    return dev.to_device_colmajor(a, dtype)


def _prec(dtype):
    from mnpbem_amd import _lib
    return _lib.MNP_FP64 if dtype == np.complex128 else _lib.MNP_FP32


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("mnk", [(64, 64, 16), (130, 77, 45), (1, 300, 257), (257, 3, 1000),
                                 (2100, 2300, 700), (2049, 2177, 129), (2051, 2103, 64), (2311, 2067, 75),
                                 # few right-hand sides: the tall-skinny kernel, direct and split over k
                                 (64, 16, 256), (3000, 8, 300), (12000, 16, 5000), (777, 13, 9999)])
def test_gemm(dtype, mnk):
    import torch
    from mnpbem_amd import _lib, device as dev
    dev.ensure()
    print(f"[gemm] {dtype.__name__} {mnk}", flush=True)
    m, n, k = mnk
    rng = np.random.default_rng(1)
    A, B, C0 = _rand(rng, m, k), _rand(rng, k, n), _rand(rng, m, n)
    alpha, beta = 0.7 - 0.2j, -1.1 + 0.4j
    tdt = torch.complex128 if dtype == np.complex128 else torch.complex64
    dA, dB, dC = _dev(A, tdt), _dev(B, tdt), _dev(C0, tdt)
    _lib.call("mnp_gemm", _prec(dtype), m, n, k, _lib.MnpComplex.of(alpha), dev.ptr(dev.base(dA)), m,
              dev.ptr(dev.base(dB)), k, _lib.MnpComplex.of(beta), dev.ptr(dev.base(dC)), m)
    want = alpha * (A @ B) + beta * C0
    got = dev.to_host(dC)
    tol = 1e-13 if dtype == np.complex128 else 2e-5
    assert np.max(np.abs(got - want)) <= tol * np.sqrt(k) * (np.abs(want).max() + 1)
    # beta = 0 must ignore C's contents
    dC.fill_(float("nan"))
    _lib.call("mnp_gemm", _prec(dtype), m, n, k, _lib.MnpComplex.of(1.0), dev.ptr(dev.base(dA)), m,
              dev.ptr(dev.base(dB)), k, _lib.MnpComplex.of(0.0), dev.ptr(dev.base(dC)), m)
    got = dev.to_host(dC)
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - A @ B)) <= tol * np.sqrt(k) * (np.abs(A @ B).max() + 1)


def _bf16_rn(x):
    """float32 -> nearest bf16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _split_bf16_product(A, B):
    """The split-bf16 complex product in exact arithmetic: x = hi + lo with
    hi = rn(x), lo = rn(x - hi); each real product X Y -> Xh Yh + Xh Yl + Xl Yh;
    Cr = Ar Br - Ai Bi, Ci = Ar Bi + Ai Br."""
    def split(x):
        h = _bf16_rn(x)
        return h.astype(np.float64), _bf16_rn(x - h).astype(np.float64)

    def prod(X, Y):
        (xh, xl), (yh, yl) = split(X), split(Y)
        return xh @ yh + xh @ yl + xl @ yh

    Ar, Ai = A.real.astype(np.float32), A.imag.astype(np.float32)
    Br, Bi = B.real.astype(np.float32), B.imag.astype(np.float32)
    return (prod(Ar, Br) - prod(Ai, Bi)) + 1j * (prod(Ar, Bi) + prod(Ai, Br))


@pytest.mark.parametrize("mnk", [(64, 64, 16), (130, 77, 45), (1, 300, 257), (257, 3, 1000),
                                 (2100, 2300, 700), (2049, 2177, 129)])
def test_gemm_split_bf16(mnk):
    """The split-bf16 kernel computes the documented scheme: it agrees with its
    exact-arithmetic emulation to FP32 accumulation error (a dropped cross term
    would be 2^-9 off), and the scheme is within 2^-15 of the exact product."""
    import torch
    from mnpbem_amd import _lib, device as dev
    dev.ensure()
    print(f"[gemm split-bf16] {mnk}", flush=True)
    m, n, k = mnk
    rng = np.random.default_rng(2)
    A, B, C0 = (_rand(rng, m, k, np.complex64), _rand(rng, k, n, np.complex64), _rand(rng, m, n, np.complex64))
    alpha, beta = 0.75 - 0.25j, -1.0 + 0.5j  # exact in FP32
    dA, dB, dC = (_dev(x, torch.complex64) for x in (A, B, C0))
    _lib.call("mnp_gemm", _lib.MNP_FP32_SPLIT_BF16, m, n, k, _lib.MnpComplex.of(alpha), dev.ptr(dev.base(dA)), m,
              dev.ptr(dev.base(dB)), k, _lib.MnpComplex.of(beta), dev.ptr(dev.base(dC)), m)
    got = dev.to_host(dC).astype(np.complex128)
    emu = _split_bf16_product(A, B)
    exact = A.astype(np.complex128) @ B.astype(np.complex128)
    scale = np.sqrt(k) * 2.0  # rms size of one entry of A B (unit-variance complex normals)
    want = alpha * emu + beta * C0.astype(np.complex128)
    assert np.max(np.abs(got - want)) <= 2.0 ** -18 * scale * 4
    assert np.max(np.abs(emu - exact)) <= 2.0 ** -15 * scale * 4
    # beta = 0 ignores C
    dC.fill_(float("nan"))
    _lib.call("mnp_gemm", _lib.MNP_FP32_SPLIT_BF16, m, n, k, _lib.MnpComplex.of(1.0), dev.ptr(dev.base(dA)), m,
              dev.ptr(dev.base(dB)), k, _lib.MnpComplex.of(0.0), dev.ptr(dev.base(dC)), m)
    got = dev.to_host(dC).astype(np.complex128)
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - emu)) <= 2.0 ** -18 * scale * 4


def _lu_dev(A, dtype):
    import torch
    from mnpbem_amd import _lib, device as dev
    n = A.shape[0]
    tdt = torch.complex128 if dtype == np.complex128 else torch.complex64
    dA = _dev(A, tdt)
    piv = torch.empty(n, dtype=torch.int32, device="cuda")
    info = C.c_int(-1)
    _lib.call("mnp_getrf", _prec(dtype), n, n, dev.ptr(dev.base(dA)), n, dev.ptr(piv), C.byref(info))
    return dA, piv, info.value


@pytest.mark.parametrize("n", [7, 33, 200, 517, 1100])
def test_getrf_matches_lapack(n):
    import scipy.linalg as sl
    from mnpbem_amd import device as dev
    dev.ensure()
    print(f"[getrf] n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = _rand(rng, n, n)
    dA, piv, info = _lu_dev(A, np.complex128)
    assert info == 0
    LU = dev.to_host(dA)
    ipiv = piv.cpu().numpy()
    lu_ref, piv_ref = sl.lu_factor(A)
    assert np.array_equal(ipiv, piv_ref)
    assert np.allclose(LU, lu_ref, rtol=0, atol=1e-11 * np.abs(lu_ref).max())
    # P A = L U
    PA = A.copy()
    for j, p in enumerate(ipiv):
        PA[[j, p]] = PA[[p, j]]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    assert np.max(np.abs(L @ U - PA)) < 1e-12 * np.abs(A).max() * n


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_solves_and_inverse(dtype):
    import torch
    from mnpbem_amd import _lib, device as dev
    dev.ensure()
    n, m, nrhs = 300, 150, 9
    print(f"[getrs/getri] {dtype.__name__}", flush=True)
    rng = np.random.default_rng(5)
    A = _rand(rng, n, n) + 4 * np.eye(n)
    dA, piv, info = _lu_dev(A.astype(dtype), dtype)
    assert info == 0
    tdt = torch.complex128 if dtype == np.complex128 else torch.complex64
    tol = 1e-11 if dtype == np.complex128 else 5e-4
    B = _rand(rng, n, nrhs)
    dB = _dev(B, tdt)
    _lib.call("mnp_getrs", _prec(dtype), n, nrhs, dev.ptr(dev.base(dA)), n, dev.ptr(piv),
              dev.ptr(dev.base(dB)), n)
    X = np.linalg.solve(A, B)
    assert np.max(np.abs(dev.to_host(dB) - X)) < tol * np.abs(X).max()
    R = _rand(rng, m, n)
    dR = _dev(R, tdt)
    _lib.call("mnp_getrs_right", _prec(dtype), m, n, dev.ptr(dev.base(dA)), n, dev.ptr(piv),
              dev.ptr(dev.base(dR)), m)
    Y = R @ np.linalg.inv(A)
    assert np.max(np.abs(dev.to_host(dR) - Y)) < tol * np.abs(Y).max()
    dX = dev.colmajor((n, n), tdt)
    _lib.call("mnp_getri", _prec(dtype), n, dev.ptr(dev.base(dA)), n, dev.ptr(piv),
              dev.ptr(dev.base(dX)), n)
    Ai = np.linalg.inv(A)
    assert np.max(np.abs(dev.to_host(dX) - Ai)) < tol * np.abs(Ai).max()


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [45, 300, 700])
def test_getrf_nopiv(dtype, n):
    """Unpivoted LU: equals LAPACK's factors when partial pivoting would not
    exchange rows, and reports it when it would."""
    import scipy.linalg as sl
    import torch
    from mnpbem_amd import _lib, device as dev
    dev.ensure()
    print(f"[getrf_nopiv] {dtype.__name__} n={n}", flush=True)
    rng = np.random.default_rng(n + 1)
    tdt = torch.complex128 if dtype == np.complex128 else torch.complex64
    A = _rand(rng, n, n) + (2.5 * n) * np.eye(n)  # diagonally dominant: no swaps
    lu_ref, piv_ref = sl.lu_factor(A)
    assert np.array_equal(piv_ref, np.arange(n))
    dA = _dev(A, tdt)
    info, sw = C.c_int(-1), C.c_int(-1)
    _lib.call("mnp_getrf_nopiv", _prec(dtype), n, n, dev.ptr(dev.base(dA)), n, C.byref(info),
              C.byref(sw))
    assert info.value == 0 and sw.value == 0
    tol = 1e-12 if dtype == np.complex128 else 2e-5
    assert np.allclose(dev.to_host(dA), lu_ref, rtol=0, atol=tol * np.abs(lu_ref).max())
    B = _rand(rng, n, n)  # a general matrix needs exchanges
    dB = _dev(B, tdt)
    _lib.call("mnp_getrf_nopiv", _prec(dtype), n, n, dev.ptr(dev.base(dB)), n, C.byref(info),
              C.byref(sw))
    assert sw.value == 1


def test_singular_pivot_reported():
    from mnpbem_amd import device as dev
    dev.ensure()
    rng = np.random.default_rng(9)
    A = _rand(rng, 40, 40)
    A[:, 17] = 0
    _, _, info = _lu_dev(A, np.complex128)
    assert info == 18
