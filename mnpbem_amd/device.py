"""Device plumbing: torch tensors as HBM buffers handed to the C ABI.

torch only allocates memory and orders work on its current stream; every
computation runs in libmnpbem_hip.so.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib

_initialised = False


def init(device=0):
    """Selects the device and binds the library to torch's current stream."""
    global _initialised
    if not torch.cuda.is_available():
        raise _lib.MnpError("no GPU visible: the MI355X path has no CPU fallback")
    torch.cuda.set_device(device)
    _lib.call("mnp_set_device", device)
    _lib.call("mnp_set_stream", C.c_void_p(torch.cuda.current_stream().cuda_stream))
    _initialised = True


def ensure():
    if not _initialised:
        init(torch.cuda.current_device() if torch.cuda.is_available() else 0)


def ptr(t):
    """Device address of a tensor (None for None)."""
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous()
    return C.c_void_p(t.data_ptr())


def colmajor(shape, dtype=torch.complex128):
    """Uninitialised column-major device array of `shape` (returned as a
    transposed view so that indexing is t[i, j, ...] in the logical shape)."""
    t = torch.empty(tuple(reversed(shape)), dtype=dtype, device="cuda")
    return t.permute(*reversed(range(len(shape))))


def base(t):
    """Contiguous storage behind a colmajor() view."""
    b = t.permute(*reversed(range(t.dim())))
    assert b.is_contiguous()
    return b


def to_device_colmajor(a, dtype=torch.complex128):
    """Host array -> device column-major array (logical shape preserved)."""
    a = np.asarray(a)
    t = torch.from_numpy(np.ascontiguousarray(a.transpose(*reversed(range(a.ndim)))))
    t = t.to(device="cuda", dtype=dtype)
    return t.permute(*reversed(range(a.ndim)))


def to_host(t):
    return t.detach().cpu().numpy()
