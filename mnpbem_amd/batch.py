"""Batched plane-wave spectra of many particles (no counterpart in the
reference: its spectra are a loop of bem \\ exc(p, enei) and ext/sca/abs per
wavelength, Demo/planewave/ret/demospecret1.m:34-42, run once per particle).

``spectra(particles, wavelengths, exc, precision)`` solves every (particle,
wavelength) pair and returns ext, sca, abs of shape (particles, wavelengths,
npol).  Pairs are grouped by face count; each group is solved in batches of
``capacity`` members by mnp_bemret_batch_cross_sections: one launch per step
of every stage for the whole batch, captured as a hipGraph and replayed per
batch, with the panel width and product tile picked by timing the
candidates once per (faces, capacity, precision).  A member the batch could
not solve (zero pivot, or residual above 1e-12 after the refinement sweeps;
status != 0) is solved by the dense solver (bemret), as are the members of a
group when ``mode='dense'`` or when the dense solver is the faster of the
two for that face count (``mode='auto'``, timed on the first batch).
"""
import ctypes as C
import time

import numpy as np
import torch

from . import _lib
from . import device as dev

_SOLVER = {"fp64": _lib.MNP_SOLVER_FP64, "mixed": _lib.MNP_SOLVER_MIXED}


def default_capacity(nfaces, precision, budget_bytes=48 << 30, limit=128):
    """Members per batch within a device-memory budget (the batch's matrices:
    six n x n in the chain precision per member, plus the FP64 operator)."""
    per = nfaces * nfaces * (160 if precision == "fp64" else 112)
    return int(max(1, min(limit, budget_bytes // per)))


class bemretbatch:
    """mnp_bemret_batch_t for members of ``nfaces`` faces and the plane waves
    of ``exc`` (a planewaveret; its pinfty surface gives sca)."""

    def __init__(self, nfaces, capacity, precision, exc, use_graph=True, variant=-1, sweeps=-1):
        if precision not in _SOLVER:
            raise ValueError(f"bemretbatch: precision must be one of {sorted(_SOLVER)}")
        dev.ensure()
        self.n, self.cap, self.precision, self.exc = int(nfaces), int(capacity), precision, exc
        self.npol = len(exc.pol)
        h = C.c_void_p()
        _lib.call("mnp_bemret_batch_create", self.n, self.cap, _SOLVER[precision], self.npol,
                  exc.pinf_area.numel(), C.byref(h))
        self.handle = h
        _lib.call("mnp_bemret_batch_set_options", h, int(use_graph), int(variant), int(sweeps))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib.can_release():
            _lib.lib().mnp_bemret_batch_destroy(h)
            self.handle = None

    def cross_sections(self, members):
        """members: list of (comparticle, enei) with count <= capacity.
        Returns ext, sca, abs (count x npol), status (count) and relres."""
        m = len(members)
        if not 1 <= m <= self.cap:
            raise ValueError("bemretbatch: 1 <= members <= capacity")
        greens = (C.c_void_p * m)(*[p.green.handle.value for p, _ in members])
        lam = np.array([float(e) for _, e in members])
        e1 = (_lib.MnpComplex * m)(*[_lib.MnpComplex.of(p.eps1(e)) for p, e in members])
        e2 = (_lib.MnpComplex * m)(*[_lib.MnpComplex.of(p.eps2(e)) for p, e in members])
        for p, _ in members:
            if p.inout[0, 1] != self.exc.medium:
                raise NotImplementedError("the particle's outside must be the embedding medium")
        out = [np.zeros((m, self.npol)) for _ in range(3)]
        status = np.zeros(m, dtype=np.int32)
        relres = np.zeros(m)
        pol = np.asfortranarray(self.exc.pol)
        dr = np.asfortranarray(self.exc.dir)
        _lib.call("mnp_bemret_batch_cross_sections", self.handle, m, greens, lam.ctypes.data_as(C.c_void_p),
                  e1, e2, pol.ctypes.data_as(C.c_void_p), dr.ctypes.data_as(C.c_void_p),
                  dev.ptr(dev.base(self.exc.pinf_nvec)), dev.ptr(self.exc.pinf_area),
                  *[o.ctypes.data_as(C.c_void_p) for o in out], status.ctypes.data_as(C.c_void_p),
                  relres.ctypes.data_as(C.c_void_p))
        return out[0], out[1], out[2], status, relres

    def info(self):
        """(variant in use, tuner timings per variant in ms, graph nodes)."""
        t, n = C.c_int(), C.c_int64()
        ms = np.zeros(6)
        _lib.call("mnp_bemret_batch_info", self.handle, C.byref(t), ms.ctypes.data_as(C.c_void_p), C.byref(n))
        return t.value, ms, n.value


def _dense(p, enei, exc, precision, cache):
    from . import bem as B
    key = id(p)
    if key not in cache:
        cache.clear()  # one dense solver at a time: its matrices are large
        cache[key] = B.bemret(p, precision=precision)
    return exc.cross_sections(cache[key].solve(exc(p, enei)))


def plan_capacity(particles, nlam, precision):
    """Batch size per face count for a sweep of nlam wavelengths: the memory
    budget's, at most the number of members."""
    counts = {}
    for p in particles:
        counts[p.n] = counts.get(p.n, 0) + nlam
    return {n: min(default_capacity(n, precision), c) for n, c in counts.items()}


def spectra(particles, wavelengths, exc, precision="mixed", capacity=None, mode="batch", stats=None):
    """ext, sca, abs (len(particles) x len(wavelengths) x npol) of every
    (particle, wavelength) pair; see the module docstring.  ``capacity``: the
    batch size (an int, at most the members of a face count), or a dict of
    fixed batch sizes per face count (e.g. plan_capacity's, so that a short
    sweep tunes the variants a longer one then uses).  ``stats`` (a dict)
    receives per face count the batch size, variant, graph size, members solved
    by the dense solver and the timings behind mode 'auto'."""
    lams = np.asarray(wavelengths, dtype=np.float64)
    npol = len(exc.pol)
    out = np.zeros((3, len(particles), len(lams), npol))
    groups = {}
    for ip, p in enumerate(particles):
        groups.setdefault(p.n, []).append(ip)
    for n, ips in sorted(groups.items()):
        members = [(ip, il) for ip in ips for il in range(len(lams))]
        st = {"members": len(members), "dense_fallbacks": 0}
        cache = {}
        use_dense = mode == "dense"
        if not use_dense:
            if isinstance(capacity, dict):
                cap = int(capacity[n])  # a fixed batch size (slots beyond the members repeat the last one)
            else:
                cap = min(capacity or default_capacity(n, precision), len(members))
            bb = bemretbatch(n, cap, precision, exc)
            st["capacity"] = cap
            start = 0
            if mode == "auto":
                # time one batch against the dense solver on one member
                first = members[:cap]
                torch.cuda.synchronize()
                bb.cross_sections([(particles[ip], lams[il]) for ip, il in first])  # tunes and captures
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                res = bb.cross_sections([(particles[ip], lams[il]) for ip, il in first])
                torch.cuda.synchronize()
                t_batch = (time.perf_counter() - t0) / len(first)
                ip0, il0 = first[0]
                _dense(particles[ip0], lams[il0], exc, precision, cache)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                _dense(particles[ip0], lams[il0] * (1 + 1e-9), exc, precision, cache)
                torch.cuda.synchronize()
                t_dense = time.perf_counter() - t0
                st["s_per_member_batch"], st["s_per_member_dense"] = t_batch, t_dense
                if t_dense < t_batch:
                    use_dense = True
                else:
                    _store(out, first, res, particles, lams, exc, precision, cache, st)
                    start = cap
            if not use_dense:
                for a in range(start, len(members), cap):
                    chunk = members[a:a + cap]
                    res = bb.cross_sections([(particles[ip], lams[il]) for ip, il in chunk])
                    _store(out, chunk, res, particles, lams, exc, precision, cache, st)
                t, ms, nodes = bb.info()
                st.update(variant=t, variant_ms=ms.tolist(), graph_nodes=nodes)
            del bb
        if use_dense:
            st["dense"] = True
            for ip, il in members:
                c = _dense(particles[ip], lams[il], exc, precision, cache)
                for k in range(3):
                    out[k, ip, il] = c[k]
        cache.clear()
        if stats is not None:
            stats[n] = st
    return out[0], out[1], out[2]


def _store(out, chunk, res, particles, lams, exc, precision, cache, st):
    ext, sca, ab, status, _ = res
    for j, (ip, il) in enumerate(chunk):
        if status[j] != 0:
            st["dense_fallbacks"] += 1
            c = _dense(particles[ip], lams[il], exc, precision, cache)
            ext[j], sca[j], ab[j] = c
        out[0, ip, il], out[1, ip, il], out[2, ip, il] = ext[j], sca[j], ab[j]
