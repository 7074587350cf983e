"""Host-side mirror of MNPBEM's retarded BEM operator API.

The names and argument meanings follow MNPBEM:

    op   = bemoptions(sim='ret')                       # Misc/bemoptions.m
    p    = comparticle([epsconst(1), epsdrude(...)], [(verts, faces)], [2, 1], 1, op)
    bem  = bemsolver(p, op)                            # Base/bemsolver.m -> bemret
    exc  = planewave([[1, 0, 0]], [[0, 0, 1]], op)     # Base/planewave.m -> planewaveret
    sig  = bem.solve(exc(p, enei))                     # sig = bem \\ exc(p, enei)
    ext, sca, ab = exc.ext(sig), exc.sca(sig), exc.abs(sig)

Everything numerical runs in libmnpbem_hip.so on the GPU; this module only
keeps handles, allocates device buffers through torch, and converts the
Matlab-style options.  Supported here: particles of flat triangles, one
closed particle in one medium (inout = [inside, outside]) for every solver,
and compounds of several parts with their own media (dimers, core-shell
particles: comparticle(eps, [p1, p2], [[3, 2], [2, 1]], 1, 2, op)) for the
dense retarded solver with plane waves.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from . import device as dev
from .greenfun import CompGreenRet

EV2NM = 1.0 / 8.0655477e-4  # Misc/units.m:2


# ---------------------------------------------------------------- materials
class epsconst:
    """Constant dielectric function (Material/@epsconst)."""

    def __init__(self, eps):
        self.eps = complex(eps)

    def __call__(self, enei):
        return self.eps, 2 * np.pi / enei * np.sqrt(self.eps)


class epsdrude:
    """Drude dielectric function eps0 - wp^2 / (w (w + i gammad)), w = eV2nm / enei
    (Material/@epsdrude/subsref.m:16-28); parameters in eV."""

    def __init__(self, eps0, wp, gammad):
        self.eps0, self.wp, self.gammad = float(eps0), float(wp), float(gammad)

    def __call__(self, enei):
        w = EV2NM / enei
        eps = self.eps0 - self.wp ** 2 / (w * (w + 1j * self.gammad))
        return eps, 2 * np.pi / enei * np.sqrt(eps)


class epsfun:
    """User function of the wavelength in nm (Material/epsfun.m)."""

    def __init__(self, fun, key="nm"):
        self.fun, self.key = fun, key

    def __call__(self, enei):
        eps = complex(self.fun(enei if self.key == "nm" else EV2NM / enei))
        return eps, 2 * np.pi / enei * np.sqrt(eps)


def _notaknot(x, y):
    """Cubic spline through (x, y), x increasing, with not-a-knot ends (the
    end conditions of Matlab's spline for four or more points): returns the
    second derivatives at the knots."""
    n = len(x)
    if n < 4:
        raise ValueError("epstable: at least four tabulated points are needed")
    h = np.diff(x)
    d = np.diff(y) / h
    A = np.zeros((n, n))
    r = np.zeros(n, dtype=y.dtype)
    for i in range(1, n - 1):
        A[i, i - 1:i + 2] = h[i - 1], 2 * (h[i - 1] + h[i]), h[i]
        r[i] = 6 * (d[i] - d[i - 1])
    # third derivative continuous at the second and the second-last knot
    A[0, :3] = h[1], -(h[0] + h[1]), h[0]
    A[-1, -3:] = h[-1], -(h[-2] + h[-1]), h[-2]
    return np.linalg.solve(A, r)


class epstable:
    """Tabulated dielectric function (Material/@epstable): a file of lines
    'energy (eV)  n  k' ('%' starts a comment); n and k are interpolated by
    not-a-knot cubic splines in the wavelength eV2nm / energy (epstable.m:
    19-26) and eps = (n + i k)^2 (subsref.m:22-33).  Wavelengths outside the
    table raise, as the reference asserts."""

    def __init__(self, path):
        rows = []
        with open(path) as fh:
            for line in fh:
                line = line.split("%", 1)[0].split()
                if line:
                    rows.append([float(v) for v in line[:3]])
        ene, n, k = np.asarray(rows).T
        order = np.argsort(EV2NM / ene)
        self.enei = (EV2NM / ene)[order]
        self.ni, self.ki = n[order], k[order]
        self._m = _notaknot(self.enei, self.ni + 1j * self.ki)

    def nk(self, enei):
        """n + i k at the wavelength (the splines of n and k)."""
        x, m = self.enei, self._m
        if not (x[0] <= enei <= x[-1]):
            raise ValueError(f"epstable: wavelength {enei} nm outside the table [{x[0]}, {x[-1]}]")
        i = min(max(int(np.searchsorted(x, enei)) - 1, 0), len(x) - 2)
        h = x[i + 1] - x[i]
        a, b = (x[i + 1] - enei) / h, (enei - x[i]) / h
        y = self.ni + 1j * self.ki
        return a * y[i] + b * y[i + 1] + ((a ** 3 - a) * m[i] + (b ** 3 - b) * m[i + 1]) * h * h / 6

    def __call__(self, enei):
        eps = complex(self.nk(enei)) ** 2
        return eps, 2 * np.pi / enei * np.sqrt(eps)


# ------------------------------------------------------------------ options
def bemoptions(**kw):
    """bemoptions('sim','ret', ...) defaults (Misc/bemoptions.m:18-24)."""
    op = dict(sim="ret", waitbar=0, RelCutoff=3.0, order=5, interp="flat")
    op.update(kw)
    return op


# --------------------------------------------------------------- particles
class comparticle:
    """comparticle(eps, {p1, p2, ...}, inout, closed..., op).

    ``eps`` is the table of dielectric functions, ``inout`` (one row per part)
    the 1-based [inside, outside] media of each part
    (Particles/@comparticle/comparticle.m:11-28).  The arguments after inout
    are closed arguments (Particles/@comparticle/closed.m: an index or a list
    of part indices; every listed part not set before gets the closed surface
    made of the listed parts) followed by the option dict; with one part the
    legacy form ``comparticle(eps, [p], inout, closed=1, op)`` is the same.
    Several parts make a compound (dimers, core-shell particles, ...), solved
    by the dense solver.
    """

    def __init__(self, eps, parts, inout, *args, closed=None, op=None):
        closed_args = []
        for a in args:
            if isinstance(a, dict):
                op = a
            else:
                closed_args.append(a)
        if closed is not None:
            closed_args.append(closed)
        op = bemoptions() if op is None else op
        if op.get("interp", "flat") != "flat":
            raise NotImplementedError("only flat boundary elements are supported")
        self.eps = list(eps)
        self.parts = [(np.asarray(v), np.asarray(f)) for v, f in parts]
        if not self.parts:
            raise ValueError("comparticle: no particle")
        self.inout = np.atleast_2d(np.asarray(inout, dtype=int))
        if self.inout.shape != (len(self.parts), 2):
            raise ValueError("comparticle: inout needs one [inside, outside] row per particle")
        self.op = op
        self._green = None
        if len(self.parts) == 1:
            self.verts, self.faces = self.parts[0]
            # a closed argument 1 (or True) closes the particle; 0, False or none leaves it open
            self.closed = any(c is not False and np.all(np.asarray(c) != 0) for c in closed_args)
            self.closed_mask = np.array([1 if self.closed else 0], dtype=np.uint32)
        else:
            verts, faces, off = [], [], 0
            self.part_off = [0]
            for v, f in self.parts:
                verts.append(v)
                faces.append(f + off)
                off += len(v)
                self.part_off.append(self.part_off[-1] + len(f))
            self.verts, self.faces = np.vstack(verts), np.vstack(faces)
            mask = [0] * len(self.parts)
            done = [False] * len(self.parts)
            for c in closed_args:
                idx = [int(i) for i in np.atleast_1d(c)]
                if any(i <= 0 or i > len(self.parts) for i in idx):
                    raise NotImplementedError("closed arguments: part indices 1..nparts (no flipped parts)")
                bits = sum(1 << (i - 1) for i in idx)
                for i in idx:
                    if not done[i - 1]:
                        mask[i - 1], done[i - 1] = bits, True
            self.closed_mask = np.array(mask, dtype=np.uint32)
            self.closed = bool(np.any(self.closed_mask))

    @property
    def compound(self):
        return len(self.parts) > 1

    @property
    def n(self):
        return len(self.faces)

    @property
    def face_inout(self):
        """n x 2 media (1-based) of every face: its part's inout row."""
        if not self.compound:
            return np.broadcast_to(self.inout[0], (self.n, 2))
        return np.repeat(self.inout, np.diff(self.part_off), axis=0)

    def eps1(self, enei):
        if self.compound:
            raise NotImplementedError("eps1 of a compound: use eps_media")
        return self.eps[self.inout[0, 0] - 1](enei)[0]

    def eps2(self, enei):
        if self.compound:
            raise NotImplementedError("eps2 of a compound: use eps_media")
        return self.eps[self.inout[0, 1] - 1](enei)[0]

    def eps_media(self, enei):
        """Dielectric function of every medium of the table at enei."""
        return [e(enei)[0] for e in self.eps]

    @property
    def green(self):
        """The particle's compgreenret(p, p, op), built once per mesh."""
        if self._green is None:
            self._green = self._make_green()
        return self._green

    def _make_green(self, like=None):
        op = self.op
        kw = dict(order=op.get("order", 5), rel_cutoff=op.get("RelCutoff", 0.0), abs_cutoff=op.get("AbsCutoff", 0.0),
                  npol=tuple(np.broadcast_to(op.get("npol", (7, 5)), (2,))), closed=self.closed, like=like)
        if self.compound:
            if like is not None:
                raise NotImplementedError("moved compound particles")
            kw["compound"] = (np.asarray(self.part_off, np.int64), self.inout, self.closed_mask)
        return CompGreenRet(self.verts, self.faces, **kw)

    def moved(self, verts):
        """The same particle with vertices ``verts`` and the refined pairs of
        this one (no new refinement selection: the Green function's elements
        stay continuous in the vertices)."""
        if self.compound:
            raise NotImplementedError("moved compound particles")
        q = comparticle(self.eps, [(verts, self.faces)], self.inout, closed=1 if self.closed else 0, op=self.op)
        q._green = q._make_green(like=self.green)
        return q

    def geometry(self):
        return self.green.geometry()


# ---------------------------------------------------------- result struct
class compstruct(dict):
    """Fields of a BEM quantity on the particle surface (Particles/@compstruct).

    Field values are device tensors: scalars n x npol, vectors n x 3 x npol.
    """

    def __init__(self, p, enei, **fields):
        super().__init__(**fields)
        self.p, self.enei = p, enei

    def __getattr__(self, key):
        try:
            return self[key]
        except KeyError as e:
            raise AttributeError(key) from e


# ------------------------------------------------------------------ solvers
class bemret:
    """Dense retarded BEM solver (BEM/@bemret), FP64 or FP32 + FP64 refinement.

    ``precision='fp64'`` is the drop-in of bemret; ``precision='mixed'``
    builds the chain in complex FP32 and refines with the exact operator in
    FP64 (the structure of bemretiter with precond 'full');
    ``precision='mixed_bf16'`` forms the chain's large products from hi/lo
    bf16 splits on bf16 MFMA, with the same refinement.
    """

    _SOLVERS = {"fp64": _lib.MNP_SOLVER_FP64, "mixed": _lib.MNP_SOLVER_MIXED,
                "mixed_bf16": _lib.MNP_SOLVER_MIXED_BF16}

    def __init__(self, p, op=None, precision=None):
        op = p.op if op is None else op
        self.p = p
        self.precision = precision or op.get("precision", "fp64")
        if self.precision not in self._SOLVERS:
            raise ValueError(f"bemret: unknown precision {self.precision!r}")
        solver = self._SOLVERS[self.precision]
        h = C.c_void_p()
        _lib.call("mnp_bemret_create", p.green.handle, solver, C.byref(h))
        self.handle = h
        if "refine_tol" in op or "refine_maxit" in op:
            _lib.call("mnp_bemret_set_refinement", h, float(op.get("refine_tol", 1e-12)),
                      int(op.get("refine_maxit", 10)))
        self.enei = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib.can_release():
            _lib.lib().mnp_bemret_destroy(h)
            self.handle = None

    def __call__(self, enei):
        """bem(enei): initialise the matrices for one wavelength (initmat.m)."""
        if self.enei != enei:
            self.enei = None  # stays unset if initmat fails part-way
            if getattr(self.p, "compound", False):
                em = self.p.eps_media(enei)
                arr = (_lib.MnpComplex * len(em))(*[_lib.MnpComplex.of(e) for e in em])
                _lib.call("mnp_bemret_initmat_media", self.handle, float(enei), len(em), arr)
            else:
                e1, e2 = self.p.eps1(enei), self.p.eps2(enei)
                _lib.call("mnp_bemret_initmat", self.handle, float(enei), _lib.MnpComplex.of(e1),
                          _lib.MnpComplex.of(e2))
            self.enei = enei
        return self

    def solve(self, exc):
        """sig = bem \\ exc  (BEM/@bemret/mldivide.m)."""
        self(exc.enei)
        n = self.p.n
        npol = next(v.shape[-1] for v in exc.values() if v is not None)
        args = []
        for key in ("phi1", "phi1p", "a1", "a1p", "phi2", "phi2p", "a2", "a2p"):
            v = exc.get(key)
            args.append(None if v is None else dev.ptr(dev.base(v)))
        sig1 = dev.colmajor((n, npol))
        sig2 = dev.colmajor((n, npol))
        h1 = dev.colmajor((n, 3, npol))
        h2 = dev.colmajor((n, 3, npol))
        _lib.call("mnp_bemret_solve", self.handle, npol, *args,
                  *[dev.ptr(dev.base(t)) for t in (sig1, sig2, h1, h2)])
        return compstruct(self.p, exc.enei, sig1=sig1, sig2=sig2, h1=h1, h2=h2)

    mldivide = solve
    __truediv__ = None

    def stats(self):
        """(refinement sweeps, final relative residual) of the last solve."""
        it, rr = C.c_int(), C.c_double()
        _lib.call("mnp_bemret_stats", self.handle, C.byref(it), C.byref(rr))
        return it.value, rr.value


def initmat_many(bems, enei):
    """bem(enei) for several dense retarded solvers at once (one per particle of
    a sweep; ``enei`` a scalar or one wavelength per solver): their assemblies
    and factorisations run four at a time on concurrent streams
    (mnp_bemret_initmat_many)."""
    bems = list(bems)
    lam = np.broadcast_to(np.asarray(enei, dtype=np.float64), (len(bems),)).copy()
    todo = [i for i, b in enumerate(bems) if b.enei != lam[i]]
    if not todo:
        return
    if len({id(bems[i]) for i in todo}) != len(todo):
        raise ValueError("initmat_many: a solver appears twice")
    # The C call initialises the solvers group by group; should a later group
    # fail, earlier ones already hold the new wavelength.  Forget every cached
    # wavelength first, so that no solver is later taken to be at its old one.
    for i in todo:
        bems[i].enei = None
    hs = (C.c_void_p * len(todo))(*[bems[i].handle.value for i in todo])
    l = np.ascontiguousarray(lam[todo])
    e1 = (_lib.MnpComplex * len(todo))(*[_lib.MnpComplex.of(bems[i].p.eps1(lam[i])) for i in todo])
    e2 = (_lib.MnpComplex * len(todo))(*[_lib.MnpComplex.of(bems[i].p.eps2(lam[i])) for i in todo])
    _lib.call("mnp_bemret_initmat_many", len(todo), hs, l.ctypes.data_as(C.POINTER(C.c_double)), e1, e2)
    for i in todo:
        bems[i].enei = lam[i]


def bemiter_options(**kw):
    """bemiter.options defaults (BEM/@bemiter/options.m:18-44)."""
    op = dict(solver="gmres", tol=1e-6, maxit=100, restart=None, precond="hmat", output=0,
              cleaf=200, htol=1e-6, kmax=[4, 100], fadmiss=None, hprecision="fp64")
    op.update(kw)
    return op


class bemretiter:
    """Iterative retarded BEM solver with H-matrices (BEM/@bemretiter).

    ``op['iter']`` holds bemiter options (cleaf, htol, kmax, tol, maxit,
    precond) and ``hprecision`` ('fp64', the reference's H-matrices, or
    'fp32' storage).  GMRES runs on the GPU with the exact operator (four
    H-matvecs) and the preconditioner ``precond`` ('hmat': two levels, a
    coarse bemret chain on aggregated H-matrices and block Jacobi; 'block':
    block Jacobi; 'none'); see mnp_bemretiter_* in include/mnpbem_hip.h.
    """

    _PREFIX = "mnp_bemretiter"

    def __init__(self, p, op):
        from .hmatrix import admissibility, block_ranges, clustertree
        it = bemiter_options(**op.get("iter", {}))
        if it["solver"] != "gmres":
            raise NotImplementedError("only GMRES is implemented")
        self.p, self.iter = p, it
        geo = p.geometry()
        self.tree = clustertree(geo["pos"], it["cleaf"])
        full, low = admissibility(self.tree, it["fadmiss"])
        fr = np.ascontiguousarray(block_ranges(self.tree, full).T)
        lr = np.ascontiguousarray(block_ranges(self.tree, low).T)
        perm = np.ascontiguousarray(self.tree.perm, dtype=np.int64)
        kmax = int(np.max(np.atleast_1d(it["kmax"])))
        htol = float(np.min(np.atleast_1d(it["htol"])))
        h = C.c_void_p()
        prec = {"fp64": _lib.MNP_FP64, "fp32": _lib.MNP_FP32}[it["hprecision"]]
        self._create(prec, perm, fr, lr, htol, kmax, h)
        self.handle = h
        _lib.call(self._PREFIX + "_set_options", h, float(it["tol"]), int(it["maxit"]),
                  self._precond_code(it["precond"]))
        if "coarse" in it and self._PREFIX == "mnp_bemretiter":
            # row tiles per aggregate of the two-level preconditioner's coarse level
            _lib.call("mnp_bemretiter_set_coarse", h, int(it["coarse"]))
        self.enei = None

    def _create(self, prec, perm, fr, lr, htol, kmax, h):
        _lib.call(self._PREFIX + "_create", self.p.green.handle, prec, perm.ctypes.data_as(C.c_void_p), len(fr),
                  fr.ctypes.data_as(C.c_void_p), len(lr), lr.ctypes.data_as(C.c_void_p), htol, kmax,
                  C.byref(h))

    # precond: 'hmat' (the bemiter default) couples the whole surface: the
    # reference's rank-4 H-matrix chain, here the two-level preconditioner
    # (coarse bemret chain on the H-matrices' row tiles + block Jacobi);
    # 'block' the block-Jacobi smoother alone; None / 'none' none.
    _PRECOND = {"hmat": 2, "block": 1, "none": 0}

    def _precond_code(self, name):
        if name in (None, False):
            return 0
        if name not in self._PRECOND:
            raise ValueError(f"unknown precond {name!r}: 'hmat', 'block' or 'none'")
        return self._PRECOND[name]

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib.can_release():
            getattr(_lib.lib(), self._PREFIX + "_destroy")(h)
            self.handle = None

    def __call__(self, enei):
        if self.enei != enei:
            self.enei = None  # stays unset if initmat fails part-way
            e1, e2 = self.p.eps1(enei), self.p.eps2(enei)
            _lib.call(self._PREFIX + "_initmat", self.handle, float(enei), _lib.MnpComplex.of(e1),
                      _lib.MnpComplex.of(e2))
            self.enei = enei
        return self

    def solve(self, exc):
        """[sig, bem] = solve(bem, exc)  (BEM/@bemretiter/solve.m)."""
        self(exc.enei)
        n = self.p.n
        npol = next(v.shape[-1] for v in exc.values() if v is not None)
        args = [None if exc.get(k) is None else dev.ptr(dev.base(exc.get(k)))
                for k in ("phi1", "phi1p", "a1", "a1p", "phi2", "phi2p", "a2", "a2p")]
        out = dict(sig1=dev.colmajor((n, npol)), sig2=dev.colmajor((n, npol)),
                   h1=dev.colmajor((n, 3, npol)), h2=dev.colmajor((n, 3, npol)))
        _lib.call("mnp_bemretiter_solve", self.handle, npol, *args,
                  *[dev.ptr(dev.base(out[k])) for k in ("sig1", "sig2", "h1", "h2")])
        return compstruct(self.p, exc.enei, **out)

    mldivide = solve

    def stats(self):
        """dict(flag, relres, iter, true_relres, compression) of the last solve."""
        f, rr, it, tr, cp = C.c_int(), C.c_double(), C.c_int(), C.c_double(), C.c_double()
        _lib.call(self._PREFIX + "_stats", self.handle, C.byref(f), C.byref(rr), C.byref(it),
                  C.byref(tr), C.byref(cp))
        return dict(flag=f.value, relres=rr.value, iter=it.value, true_relres=tr.value,
                    compression=cp.value)


class bemstatiter(bemretiter):
    """Iterative quasistatic BEM solver with an H-matrix of F (BEM/@bemstatiter):
    GMRES on -(F + Lambda) sig = phi' with a block preconditioner; options as
    for bemretiter (op['iter'])."""

    _PREFIX = "mnp_bemstatiter"
    # the quasistatic solver has the block preconditioner only ('hmat' maps to it)
    _PRECOND = {"hmat": 1, "block": 1, "none": 0}

    def solve(self, exc):
        """[sig, bem] = solve(bem, exc)  (BEM/@bemstatiter/solve.m), exc with field phip."""
        self(exc.enei)
        n, npol = self.p.n, exc.phip.shape[-1]
        sig = dev.colmajor((n, npol))
        _lib.call("mnp_bemstatiter_solve", self.handle, npol, dev.ptr(dev.base(exc.phip)), dev.ptr(dev.base(sig)))
        return compstruct(self.p, exc.enei, sig=sig)

    mldivide = solve


class bemstat:
    """Quasistatic BEM solver (BEM/@bemstat): sig = -(Lambda + F)^{-1} phi'.

    ``precision='fp64'`` is the drop-in; ``'mixed'`` factorises in complex
    FP32 and refines in FP64 against F x + Lambda x; ``'mixed_bf16'`` forms
    the factorisation's large products from hi/lo bf16 splits.
    """

    def __init__(self, p, op=None, precision=None):
        op = p.op if op is None else op
        self.p = p
        self.precision = precision or op.get("precision", "fp64")
        if self.precision not in bemret._SOLVERS:
            raise ValueError(f"bemstat: unknown precision {self.precision!r}")
        solver = bemret._SOLVERS[self.precision]
        h = C.c_void_p()
        _lib.call("mnp_bemstat_create", p.green.handle, solver, C.byref(h))
        self.handle = h
        self.enei = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib.can_release():
            _lib.lib().mnp_bemstat_destroy(h)
            self.handle = None

    def __call__(self, enei):
        if self.enei != enei:
            self.enei = None  # stays unset if initmat fails part-way
            _lib.call("mnp_bemstat_initmat", self.handle, float(enei), _lib.MnpComplex.of(self.p.eps1(enei)),
                      _lib.MnpComplex.of(self.p.eps2(enei)))
            self.enei = enei
        return self

    def solve(self, exc):
        """sig = bem \\ exc  (BEM/@bemstat/mldivide.m:13-15), exc with field phip."""
        self(exc.enei)
        n, npol = self.p.n, exc.phip.shape[-1]
        sig = dev.colmajor((n, npol))
        _lib.call("mnp_bemstat_solve", self.handle, npol, dev.ptr(dev.base(exc.phip)), dev.ptr(dev.base(sig)))
        return compstruct(self.p, exc.enei, sig=sig)

    mldivide = solve

    def stats(self):
        it, rr = C.c_int(), C.c_double()
        _lib.call("mnp_bemstat_stats", self.handle, C.byref(it), C.byref(rr))
        return it.value, rr.value


# ------------------------------------------------------- mirror symmetry
# Particles/@comparticlemirror/private/init.m:29-37 and symvalue.m:14-31
_SYMTABLE = {
    "x": np.array([[1, 1], [1, -1]]),
    "y": np.array([[1, 1], [1, -1]]),
    "xy": np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, 1, -1], [1, -1, -1, 1]]),
}
_SYMVALUE = {"+": (1, 1), "-": (1, -1), "++": (1, 1, 1, 1), "+-": (1, 1, -1, -1),
             "-+": (1, -1, 1, -1), "--": (1, -1, -1, 1)}


def _mirror_full_mesh(verts, faces, sym):
    """The full particle of comparticlemirror/private/init.m:42-54: copies
    [p, flip x] ('x'), [p, flip y] ('y') or [p, flip x, flip y, flip y of
    flip x] ('xy'); a flip negates one coordinate and reverses the vertex
    order of each face (Particles/@particle/flip.m, flipfaces.m)."""
    parts = [(np.asarray(verts, np.float64), np.asarray(faces, np.int64))]
    for k, keys in ((0, ("x", "xy")), (1, ("y", "xy"))):
        if sym in keys:
            new = []
            for v, f in parts:
                v = v.copy()
                v[:, k] = -v[:, k]
                new.append((v, f[:, ::-1].copy()))
            parts += new
    off = np.cumsum([0] + [len(v) for v, _ in parts])
    return (np.vstack([v for v, _ in parts]), np.vstack([f + off[i] for i, (_, f) in enumerate(parts)]))


class comparticlemirror(comparticle):
    """comparticlemirror(eps, {p}, inout, closed, op) with op['sym'] = 'x',
    'y' or 'xy' (Particles/@comparticlemirror): ``parts`` is the symmetry
    part, which the mirror copies complete to the closed particle
    ``full()``.  ``n`` counts the faces of the part; the Green function is
    the full particle's."""

    def __init__(self, eps, parts, inout, closed=1, op=None):
        op = bemoptions() if op is None else op
        sym = op.get("sym")
        if sym not in _SYMTABLE:
            raise ValueError("comparticlemirror: op['sym'] must be 'x', 'y' or 'xy'")
        super().__init__(eps, parts, inout, closed, op)
        self.sym = sym
        self.symtable = _SYMTABLE[sym]
        V, F = _mirror_full_mesh(self.verts, self.faces, sym)
        self.pfull = comparticle(eps, [(V, F)], inout, closed, {k: v for k, v in op.items() if k != "sym"})

    def full(self):
        return self.pfull

    @property
    def green(self):
        return self.pfull.green

    def geometry(self):
        g = self.pfull.geometry()
        return {k: v[:self.n] for k, v in g.items()}

    def symvalue(self, key):
        return np.array([_SYMVALUE[k] for k in key]) if isinstance(key, (list, tuple)) else np.array(_SYMVALUE[key])

    def symindex(self, tab):
        return int(np.flatnonzero(np.all(self.symtable == np.asarray(tab), axis=1))[0])


class compstructmirror:
    """Particles/@compstructmirror: ``val`` is a list of compstructs on the
    symmetry part, each with ``symval`` (3 x J: symmetry values of the x, y
    and z components; scalars take the z row); ``fun`` maps it to the full
    particle."""

    def __init__(self, p, enei, fun=None):
        self.p, self.enei, self.fun, self.val = p, enei, fun, []

    def full(self):
        return self.fun(self)


class bemretmirror:
    """Dense retarded BEM solver with mirror symmetry (BEM/@bemretmirror),
    precision 'fp64', 'mixed' or 'mixed_bf16' as for bemret."""

    def __init__(self, p, op=None, precision=None):
        op = p.op if op is None else op
        if not isinstance(p, comparticlemirror):
            raise TypeError("bemretmirror needs a comparticlemirror")
        self.p = p
        self.precision = precision or op.get("precision", "fp64")
        if self.precision not in bemret._SOLVERS:
            raise ValueError(f"bemretmirror: unknown precision {self.precision!r}")
        tab = np.ascontiguousarray(p.symtable, dtype=np.int32)
        h = C.c_void_p()
        _lib.call("mnp_bemretmirror_create", p.green.handle, bemret._SOLVERS[self.precision], len(tab),
                  tab.ctypes.data_as(C.POINTER(C.c_int)), C.byref(h))
        self.handle = h
        if "refine_tol" in op or "refine_maxit" in op:
            _lib.call("mnp_bemretmirror_set_refinement", h, float(op.get("refine_tol", 1e-12)),
                      int(op.get("refine_maxit", 10)))
        self.enei = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib.can_release():
            _lib.lib().mnp_bemretmirror_destroy(h)
            self.handle = None

    def __call__(self, enei):
        if self.enei != enei:
            self.enei = None  # stays unset if initmat fails part-way
            e1, e2 = self.p.eps1(enei), self.p.eps2(enei)
            _lib.call("mnp_bemretmirror_initmat", self.handle, float(enei), _lib.MnpComplex.of(e1),
                      _lib.MnpComplex.of(e2))
            self.enei = enei
        return self

    def solve(self, exc):
        """sig = bem \\ exc  (BEM/@bemretmirror/mldivide.m) for a compstructmirror."""
        self(exc.enei)
        n = self.p.n
        sig = compstructmirror(self.p, exc.enei, exc.fun)
        self.val_stats = []
        for val in exc.val:
            npol = next(v.shape[-1] for k, v in val.items() if v is not None and k != "symval")
            sx = [self.p.symindex(row) for row in val.symval]
            args = [None if val.get(k) is None else dev.ptr(dev.base(val[k]))
                    for k in ("phi1", "phi1p", "a1", "a1p", "phi2", "phi2p", "a2", "a2p")]
            out = dict(sig1=dev.colmajor((n, npol)), sig2=dev.colmajor((n, npol)),
                       h1=dev.colmajor((n, 3, npol)), h2=dev.colmajor((n, 3, npol)))
            _lib.call("mnp_bemretmirror_solve", self.handle, *sx, npol, *args,
                      *[dev.ptr(dev.base(out[k])) for k in ("sig1", "sig2", "h1", "h2")])
            sig.val.append(compstruct(self.p, exc.enei, symval=val.symval, **out))
            self.val_stats.append(self.stats())
        return sig

    mldivide = solve

    def stats(self):
        """(refinement sweeps, final relative residual) of the last excitation
        solved; ``val_stats`` holds them for each excitation of the last solve."""
        it, rr = C.c_int(), C.c_double()
        _lib.call("mnp_bemretmirror_stats", self.handle, C.byref(it), C.byref(rr))
        return it.value, rr.value


class bemstatmirror:
    """Dense quasistatic BEM solver with mirror symmetry (BEM/@bemstatmirror),
    precision 'fp64', 'mixed' or 'mixed_bf16' as for bemstat.  One solver
    per symmetry value; an excitation goes through the one of its z-key
    (mldivide.m:14-22)."""

    def __init__(self, p, op=None, precision=None):
        op = p.op if op is None else op
        if not isinstance(p, comparticlemirror):
            raise TypeError("bemstatmirror needs a comparticlemirror")
        self.p = p
        self.precision = precision or op.get("precision", "fp64")
        if self.precision not in bemret._SOLVERS:
            raise ValueError(f"bemstatmirror: unknown precision {self.precision!r}")
        tab = np.ascontiguousarray(p.symtable, dtype=np.int32)
        h = C.c_void_p()
        _lib.call("mnp_bemstatmirror_create", p.green.handle, bemret._SOLVERS[self.precision], len(tab),
                  tab.ctypes.data_as(C.POINTER(C.c_int)), C.byref(h))
        self.handle = h
        if "refine_tol" in op or "refine_maxit" in op:
            _lib.call("mnp_bemstatmirror_set_refinement", h, float(op.get("refine_tol", 1e-12)),
                      int(op.get("refine_maxit", 10)))
        self.enei = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib.can_release():
            _lib.lib().mnp_bemstatmirror_destroy(h)
            self.handle = None

    def __call__(self, enei):
        if self.enei != enei:
            self.enei = None  # stays unset if initmat fails part-way
            e1, e2 = self.p.eps1(enei), self.p.eps2(enei)
            _lib.call("mnp_bemstatmirror_initmat", self.handle, float(enei), _lib.MnpComplex.of(e1),
                      _lib.MnpComplex.of(e2))
            self.enei = enei
        return self

    def solve(self, exc):
        """sig = bem \\ exc  (BEM/@bemstatmirror/mldivide.m) for a compstructmirror."""
        self(exc.enei)
        sig = compstructmirror(self.p, exc.enei, exc.fun)
        self.val_stats = []
        for val in exc.val:
            s = self.p.symindex(val.symval[-1])
            npol = val.phip.shape[-1]
            out = dev.colmajor((self.p.n, npol))
            _lib.call("mnp_bemstatmirror_solve", self.handle, s, npol, dev.ptr(dev.base(val.phip)),
                      dev.ptr(dev.base(out)))
            sig.val.append(compstruct(self.p, exc.enei, symval=val.symval, sig=out))
            it, rr = C.c_int(), C.c_double()
            _lib.call("mnp_bemstatmirror_stats", self.handle, s, C.byref(it), C.byref(rr))
            self.val_stats.append((it.value, rr.value))
        return sig

    mldivide = solve

    def stats(self):
        """(refinement sweeps, final relative residual) of the last excitation solved."""
        return self.val_stats[-1] if getattr(self, "val_stats", None) else (0, 0.0)


def bemsolver(p, op=None, **kw):
    """Base/bemsolver.m: bemret for op.sim == 'ret', bemstat for op.sim ==
    'stat'; bemretiter / bemstatiter when op.iter is set; bemretmirror for a
    comparticlemirror; bemstatlayer / bemretlayer with op.layer."""
    op = dict(p.op if op is None else op)
    op.update(kw)
    sim = op.get("sim", "ret")
    if getattr(p, "compound", False) and (sim != "ret" or op.get("iter") or op.get("layer") is not None
                                          or op.get("sym")):
        raise NotImplementedError("compound particles: only the dense retarded solver (bemret) is built")
    if isinstance(p, comparticlemirror) or op.get("sym"):
        if op.get("iter") or op.get("layer") is not None:
            raise NotImplementedError("mirror symmetry: only the dense solvers without layers are built")
        if sim == "stat":
            return bemstatmirror(p, op)
        if sim != "ret":
            raise NotImplementedError(f"unknown simulation type {sim!r}")
        return bemretmirror(p, op)
    if sim == "stat":
        if op.get("layer") is not None:
            if op.get("iter"):
                raise NotImplementedError("substrate: only the dense quasistatic solver is built")
            from .layer import bemstatlayer
            return bemstatlayer(p, op)
        return bemstatiter(p, op) if op.get("iter") else bemstat(p, op)
    if op.get("layer") is not None:
        if op.get("iter") and sim == "ret":
            from .layer import bemretlayeriter
            return bemretlayeriter(p, op)
        if op.get("iter") or sim != "ret":
            raise NotImplementedError("substrate: the dense quasistatic and retarded solvers are built")
        from .layer import bemretlayer
        return bemretlayer(p, op)
    if sim != "ret":
        raise NotImplementedError(f"unknown simulation type {sim!r}")
    if op.get("iter"):
        return bemretiter(p, op)
    return bemret(p, op)


# ------------------------------------------------------------ excitations
def _side_masks(p, medium):
    """Device 0/1 masks of the faces whose inside (m1) or outside (m2) is medium."""
    fio = p.face_inout
    m1 = torch.as_tensor((fio[:, 0] == medium).astype(np.float64), device="cuda")
    m2 = torch.as_tensor((fio[:, 1] == medium).astype(np.float64), device="cuda")
    return m1, m2


def _radiating(sig, p, medium):
    """compstruct whose sig2, h2 are the charges and currents facing medium:
    sig1, h1 on faces whose inside is it plus sig2, h2 on faces whose outside is it."""
    m1, m2 = _side_masks(p, medium)
    s2 = dev.colmajor(tuple(sig.sig2.shape))
    h2 = dev.colmajor(tuple(sig.h2.shape))
    s2.copy_(sig.sig1 * m1[:, None] + sig.sig2 * m2[:, None])
    h2.copy_(sig.h1 * m1[:, None, None] + sig.h2 * m2[:, None, None])
    return compstruct(p, sig.enei, sig1=sig.sig1, h1=sig.h1, sig2=s2, h2=h2)


class planewaveret:
    """Plane-wave excitation (Simulation/retarded/@planewaveret).

    ``pol`` and ``dir`` are npol x 3; ``medium`` is the 1-based index of the
    embedding medium.  Scattering is integrated over trisphere(256, 2).
    """

    def __init__(self, pol, dir, op=None, medium=1, pinfty=None):
        self.pol = np.atleast_2d(np.asarray(pol, dtype=np.float64))
        self.dir = np.atleast_2d(np.asarray(dir, dtype=np.float64))
        if np.any(np.sum(self.pol * self.dir, axis=1) != 0):
            raise ValueError("polarisation must be perpendicular to the propagation direction")
        self.medium = medium
        if pinfty is None:
            from . import shapes
            v, f = shapes.trisphere256(2.0)
            nvec, area = _flat_normals(v, f)
        else:
            nvec, area = pinfty
        dev.ensure()
        self.pinf_nvec = dev.to_device_colmajor(nvec, torch.float64)
        self.pinf_area = torch.as_tensor(np.asarray(area, np.float64), device="cuda")

    def _eps_med(self, p, enei):
        return p.eps[self.medium - 1](enei)[0]

    def __call__(self, p, enei):
        """exc(p, enei): potentials a2, a2p on the faces (potential.m); for a
        compound a1, a1p on the faces whose inside is the medium and a2, a2p
        on those whose outside is it (potential.m:47-88, the two inout
        columns)."""
        compound = getattr(p, "compound", False)
        if not compound and p.inout[0, 1] != self.medium:
            raise NotImplementedError("the particle's outside must be the embedding medium")
        n, npol = p.n, len(self.pol)
        a2 = dev.colmajor((n, 3, npol))
        a2p = dev.colmajor((n, 3, npol))
        pol = np.asfortranarray(self.pol)
        dr = np.asfortranarray(self.dir)
        _lib.call("mnp_planewave_potential", p.green.handle, float(enei),
                  _lib.MnpComplex.of(self._eps_med(p, enei)), npol,
                  pol.ctypes.data_as(C.c_void_p), dr.ctypes.data_as(C.c_void_p),
                  dev.ptr(dev.base(a2)), dev.ptr(dev.base(a2p)))
        if not compound:
            return compstruct(p, enei, a2=a2, a2p=a2p)
        m1, m2 = _side_masks(p, self.medium)
        out = {}
        for key, v in (("a1", a2), ("a1p", a2p), ("a2", a2), ("a2p", a2p)):
            t = dev.colmajor((n, 3, npol))
            t.copy_(v * (m1 if key.startswith("a1") else m2)[:, None, None])
            out[key] = t
        return compstruct(p, enei, **out)

    potential = __call__

    def cross_sections(self, sig):
        """(ext, sca, abs) per polarisation, FP64.  For a compound the far field
        sums both sides: sig1, h1 of faces whose inside is the medium and sig2,
        h2 of faces whose outside is it (@spectrumret/farfield.m:30-60)."""
        p, enei = sig.p, sig.enei
        npol = len(self.pol)
        if getattr(p, "compound", False):
            sig = _radiating(sig, p, self.medium)
        out = [np.zeros(npol) for _ in range(3)]
        pol = np.asfortranarray(self.pol)
        dr = np.asfortranarray(self.dir)
        _lib.call("mnp_planewave_cross_sections", p.green.handle, float(enei),
                  _lib.MnpComplex.of(self._eps_med(p, enei)), npol,
                  pol.ctypes.data_as(C.c_void_p), dr.ctypes.data_as(C.c_void_p),
                  self.pinf_area.numel(), dev.ptr(dev.base(self.pinf_nvec)),
                  dev.ptr(self.pinf_area), dev.ptr(dev.base(sig.sig2)),
                  dev.ptr(dev.base(sig.h2)),
                  *[o.ctypes.data_as(C.c_void_p) for o in out])
        return tuple(out)

    def ext(self, sig):
        return self.cross_sections(sig)[0]

    def sca(self, sig):
        return self.cross_sections(sig)[1]

    def abs(self, sig):
        return self.cross_sections(sig)[2]

    extinction, scattering, absorption = ext, sca, abs


class planewavestat:
    """Quasistatic plane wave (Simulation/static/@planewavestat): phi' = -nvec . pol;
    cross sections from the induced dipole moment."""

    def __init__(self, pol, op=None, medium=1):
        self.pol = np.atleast_2d(np.asarray(pol, dtype=np.float64))
        self.medium = medium

    def __call__(self, p, enei):
        n, npol = p.n, len(self.pol)
        phip = dev.colmajor((n, npol))
        pol = np.asfortranarray(self.pol)
        _lib.call("mnp_planewavestat_potential", p.green.handle, npol, pol.ctypes.data_as(C.c_void_p),
                  dev.ptr(dev.base(phip)))
        return compstruct(p, enei, phip=phip)

    potential = __call__

    def cross_sections(self, sig):
        p, enei = sig.p, sig.enei
        npol = len(self.pol)
        out = [np.zeros(npol) for _ in range(3)]
        pol = np.asfortranarray(self.pol)
        _lib.call("mnp_planewavestat_cross_sections", p.green.handle, float(enei),
                  _lib.MnpComplex.of(p.eps[self.medium - 1](enei)[0]), npol, pol.ctypes.data_as(C.c_void_p),
                  dev.ptr(dev.base(sig.sig)), *[o.ctypes.data_as(C.c_void_p) for o in out])
        return tuple(out)

    def ext(self, sig):
        return self.cross_sections(sig)[0]

    def sca(self, sig):
        return self.cross_sections(sig)[1]

    def abs(self, sig):
        return self.cross_sections(sig)[2]

    extinction, scattering, absorption = ext, sca, abs


class planewaveretmirror:
    """Plane wave for a comparticlemirror (Simulation/retarded/
    @planewaveretmirror): pol in the xy plane, dir along +z or -z.  The
    potential holds two excitations of the symmetry part, pol x and pol y,
    each for dir +z and -z, with their symmetry values (potential.m:16-41);
    ``full(sig)`` combines their solutions on the full particle for the
    requested pol/dir (full.m:22-64), where the cross sections are taken."""

    # potential.m:27-40: symmetry keys of the x, y, z components
    _KEYS = {"x": (("+", "-", "-"), ("-", "+", "+")), "y": (("+", "-", "+"), ("-", "+", "-")),
             "xy": (("++", "--", "-+"), ("--", "++", "+-"))}

    def __init__(self, pol, dir, op=None, medium=1, pinfty=None):
        pol = np.atleast_2d(np.asarray(pol, dtype=np.float64))
        dr = np.atleast_2d(np.asarray(dir, dtype=np.float64))
        if np.any(pol[:, 2] != 0) or np.any(dr[:, :2] != 0) or np.any(np.abs(dr[:, 2]) != 1):
            raise ValueError("planewaveretmirror: pol must lie in the xy plane and dir be +z or -z")
        self.pol, self.dir = pol, dr
        self.exc = planewaveret(pol, dr, op, medium, pinfty)
        self.medium = medium

    def __call__(self, p, enei):
        pf = p.full()
        pot = compstructmirror(p, enei, self.full)
        dr = np.asfortranarray([[0, 0, 1.0], [0, 0, -1.0]])
        for i, keys in enumerate(self._KEYS[p.sym]):
            pol = np.zeros((2, 3))
            pol[:, i] = 1.0
            pol = np.asfortranarray(pol)
            a2 = dev.colmajor((pf.n, 3, 2))
            a2p = dev.colmajor((pf.n, 3, 2))
            _lib.call("mnp_planewave_potential", pf.green.handle, float(enei),
                      _lib.MnpComplex.of(pf.eps[self.medium - 1](enei)[0]), 2, pol.ctypes.data_as(C.c_void_p),
                      dr.ctypes.data_as(C.c_void_p), dev.ptr(dev.base(a2)), dev.ptr(dev.base(a2p)))
            # the symmetry part is the first n faces of the full particle
            part = {k: dev.base(v)[..., :p.n].contiguous().permute(2, 1, 0) for k, v in (("a2", a2), ("a2p", a2p))}
            pot.val.append(compstruct(p, enei, symval=p.symvalue(list(keys)), **part))
        return pot

    potential = __call__

    def full(self, sig):
        """Fields on the full particle for this plane wave's pol and dir."""
        p = sig.p
        pf = p.full()
        J, n, npol = len(p.symtable), p.n, len(self.pol)
        w = np.ascontiguousarray(self.pol[:, :2], dtype=np.float64)
        src = np.ascontiguousarray((self.dir[:, 2] < 0).astype(np.int32))
        out = {}
        v1, v2 = sig.val
        for key in ("sig1", "sig2", "h1", "h2"):
            if v1.get(key) is None:
                continue
            ncomp = 1 if key.startswith("sig") else 3
            rows = [v.symval[-1:] if ncomp == 1 else v.symval for v in (v1, v2)]
            s1, s2 = (np.ascontiguousarray(r.T, dtype=np.int32) for r in rows)  # J x ncomp
            o = dev.colmajor((J * n, npol)) if ncomp == 1 else dev.colmajor((J * n, 3, npol))
            ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
            _lib.call("mnp_mirror_full", n, J, ncomp, 2, npol, ip(s1), ip(s2),
                      w.ctypes.data_as(C.POINTER(C.c_double)), ip(src), dev.ptr(dev.base(v1[key])),
                      dev.ptr(dev.base(v2[key])), dev.ptr(dev.base(o)))
            out[key] = o
        return compstruct(pf, sig.enei, **out)

    def cross_sections(self, sig):
        return self.exc.cross_sections(self.full(sig))

    def ext(self, sig):
        return self.cross_sections(sig)[0]

    def sca(self, sig):
        return self.cross_sections(sig)[1]

    def abs(self, sig):
        return self.cross_sections(sig)[2]

    extinction, scattering, absorption = ext, sca, abs


class planewavestatmirror:
    """Quasistatic plane wave for a comparticlemirror (Simulation/static/
    @planewavestatmirror): pol in the xy plane.  The potential holds two
    excitations of the symmetry part, pol x and pol y, with their symmetry
    values (potential.m:16-28); ``full(sig)`` combines their surface charges
    on the full particle, pol_x sig_x + pol_y sig_y (full.m:23-39), where the
    cross sections of planewavestat are taken.

    The symmetry values are those of planewaveretmirror.  The reference's
    table for sym 'x' gives the pol-y excitation the z-key '-'
    (potential.m:22); its potential -y is even under x -> -x, so the key is
    '+' here, which makes the expanded charge equal the full particle's."""

    _KEYS = planewaveretmirror._KEYS

    def __init__(self, pol, op=None, medium=1):
        pol = np.atleast_2d(np.asarray(pol, dtype=np.float64))
        if np.any(pol[:, 2] != 0):
            raise ValueError("planewavestatmirror: pol must lie in the xy plane")
        self.pol = pol
        self.exc = planewavestat(pol, op, medium)
        self.medium = medium

    def __call__(self, p, enei):
        pf = p.full()
        pot = compstructmirror(p, enei, self.full)
        for i, keys in enumerate(self._KEYS[p.sym]):
            pol = np.zeros((1, 3))
            pol[0, i] = 1.0
            phip = dev.colmajor((pf.n, 1))
            _lib.call("mnp_planewavestat_potential", pf.green.handle, 1,
                      np.asfortranarray(pol).ctypes.data_as(C.c_void_p), dev.ptr(dev.base(phip)))
            # the symmetry part is the first n faces of the full particle
            part = dev.colmajor((p.n, 1))
            part.copy_(phip[:p.n])
            pot.val.append(compstruct(p, enei, symval=p.symvalue(list(keys)), phip=part))
        return pot

    potential = __call__

    def full(self, sig):
        """Surface charges on the full particle for this plane wave's pols."""
        p = sig.p
        pf = p.full()
        J, n, npol = len(p.symtable), p.n, len(self.pol)
        w = np.ascontiguousarray(self.pol[:, :2], dtype=np.float64)
        src = np.zeros(npol, dtype=np.int32)
        v1, v2 = sig.val
        s1, s2 = (np.ascontiguousarray(v.symval[-1:].T, dtype=np.int32) for v in (v1, v2))  # J x 1
        o = dev.colmajor((J * n, npol))
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        _lib.call("mnp_mirror_full", n, J, 1, 1, npol, ip(s1), ip(s2), w.ctypes.data_as(C.POINTER(C.c_double)),
                  ip(src), dev.ptr(dev.base(v1.sig)), dev.ptr(dev.base(v2.sig)), dev.ptr(dev.base(o)))
        return compstruct(pf, sig.enei, sig=o)

    def cross_sections(self, sig):
        return self.exc.cross_sections(self.full(sig))

    def ext(self, sig):
        return self.cross_sections(sig)[0]

    def sca(self, sig):
        return self.cross_sections(sig)[1]

    def abs(self, sig):
        return self.cross_sections(sig)[2]

    extinction, scattering, absorption = ext, sca, abs


def planewave(pol, dir=None, op=None, **kw):
    """Base/planewave.m: planewaveret for op.sim == 'ret', planewavestat for
    'stat', planewaveretmirror when op.sym is set."""
    if op is not None and op.get("sim", "ret") == "stat":
        if op.get("sym"):
            if op.get("layer") is not None:
                raise NotImplementedError("mirror symmetry with layers is not built")
            return planewavestatmirror(pol, op, **kw)
        if op.get("layer") is not None:
            from .layer import planewavestatlayer
            return planewavestatlayer(pol, dir, op, **kw)
        return planewavestat(pol, op, **kw)
    if op is not None and op.get("layer") is not None:
        if op.get("sym"):
            raise NotImplementedError("mirror symmetry with layers is not built")
        from .layer import planewaveretlayer
        return planewaveretlayer(pol, dir, op, **kw)
    if op is not None and op.get("sym"):
        return planewaveretmirror(pol, dir, op, **kw)
    return planewaveret(pol, dir, op, **kw)


def _flat_normals(verts, faces):
    v1, v2, v3 = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    nv = np.cross(v1 - v2, v3 - v2)
    nn = np.sqrt(np.sum(nv * nv, axis=1))
    return nv / nn[:, None], 0.5 * nn


class dipoleret:
    """Dipole excitation (Simulation/retarded/@dipoleret) for dipoles outside the particle.

    ``pt`` (npt x 3) are the positions; ``dip`` is ndip x 3 (the same
    orientations at every position, default eye(3)) or, with ``full=True``,
    npt x 3 x ndip (init.m:11-46).  ``medium`` is the 1-based index of the
    medium holding the dipoles (default: the particle's outside).
    Columns of the potentials and solutions are ipt + npt * idip.
    """

    def __init__(self, pt, dip=None, op=None, medium=None, pinfty=None, full=False):
        self.pt = np.atleast_2d(np.asarray(pt, dtype=np.float64))
        npt = len(self.pt)
        if dip is None:
            dip = np.eye(3)
        dip = np.asarray(dip, dtype=np.float64)
        if full:
            self.dip = dip.reshape(npt, 3, -1)
        else:
            d = np.atleast_2d(dip)
            self.dip = np.repeat(d.T[None, :, :], npt, axis=0)
        self.medium = medium
        if pinfty is None:
            from . import shapes
            v, f = shapes.trisphere256(2.0)
            nvec, area = _flat_normals(v, f)
        else:
            nvec, area = pinfty
        dev.ensure()
        self.pinf_nvec = dev.to_device_colmajor(nvec, torch.float64)
        self.pinf_area = torch.as_tensor(np.asarray(area, np.float64), device="cuda")

    @property
    def npol(self):
        return self.dip.shape[0] * self.dip.shape[2]

    def _medium(self, p):
        med = p.inout[0, 1] if self.medium is None else self.medium
        if med != p.inout[0, 1]:
            raise NotImplementedError("dipoles must lie in the particle's outside medium")
        return med

    def _eps_med(self, p, enei):
        return p.eps[self._medium(p) - 1](enei)[0]

    def _host(self):
        pts = np.asfortranarray(self.pt)
        dip = np.asfortranarray(self.dip)
        return pts, dip

    def __call__(self, p, enei):
        """exc(p, enei): phi2, phi2p, a2, a2p on the faces (potential.m:64-122)."""
        n, npol = p.n, self.npol
        out = dict(phi2=dev.colmajor((n, npol)), phi2p=dev.colmajor((n, npol)),
                   a2=dev.colmajor((n, 3, npol)), a2p=dev.colmajor((n, 3, npol)))
        pts, dip = self._host()
        _lib.call("mnp_dipole_potential", p.green.handle, float(enei),
                  _lib.MnpComplex.of(self._eps_med(p, enei)), len(pts), pts.ctypes.data_as(C.c_void_p),
                  dip.shape[2], dip.ctypes.data_as(C.c_void_p),
                  *[dev.ptr(dev.base(out[k])) for k in ("phi2", "phi2p", "a2", "a2p")])
        return compstruct(p, enei, **out)

    potential = __call__

    def field(self, sig):
        """Induced e, h (npt x 3 x npol) at the dipole positions (compgreenret/field.m)."""
        p, enei = sig.p, sig.enei
        npt, npol = len(self.pt), sig.sig2.shape[1]
        e = dev.colmajor((npt, 3, npol))
        h = dev.colmajor((npt, 3, npol))
        pts, _ = self._host()
        _lib.call("mnp_field_points", p.green.handle, float(enei), _lib.MnpComplex.of(self._eps_med(p, enei)),
                  npt, pts.ctypes.data_as(C.c_void_p), npol, dev.ptr(dev.base(sig.sig2)),
                  dev.ptr(dev.base(sig.h2)), dev.ptr(dev.base(e)), dev.ptr(dev.base(h)))
        return compstruct(p, enei, e=e, h=h)

    def decayrate(self, sig):
        """[tot, rad, rad0] = decayrate(exc, sig), each npt x ndip (decayrate.m:13-62)."""
        p, enei = sig.p, sig.enei
        npt, ndip = self.dip.shape[0], self.dip.shape[2]
        out = [np.zeros((npt, ndip), order="F") for _ in range(3)]
        pts, dip = self._host()
        _lib.call("mnp_dipole_decayrate", p.green.handle, float(enei),
                  _lib.MnpComplex.of(self._eps_med(p, enei)), npt, pts.ctypes.data_as(C.c_void_p), ndip,
                  dip.ctypes.data_as(C.c_void_p), self.pinf_area.numel(), dev.ptr(dev.base(self.pinf_nvec)),
                  dev.ptr(self.pinf_area), dev.ptr(dev.base(sig.sig2)), dev.ptr(dev.base(sig.h2)),
                  *[o.ctypes.data_as(C.c_void_p) for o in out])
        return tuple(out)


class dipolestat:
    """Quasistatic dipole excitation (Simulation/static/@dipolestat) for
    dipoles outside the particle; ``pt``, ``dip``, ``full`` and ``medium`` as
    for dipoleret.  Columns of the potentials and solutions are ipt + npt *
    idip."""

    def __init__(self, pt, dip=None, op=None, medium=None, full=False):
        self.pt = np.atleast_2d(np.asarray(pt, dtype=np.float64))
        npt = len(self.pt)
        dip = np.eye(3) if dip is None else np.asarray(dip, dtype=np.float64)
        self.dip = dip.reshape(npt, 3, -1) if full else np.repeat(np.atleast_2d(dip).T[None, :, :], npt, axis=0)
        self.medium = medium

    @property
    def npol(self):
        return self.dip.shape[0] * self.dip.shape[2]

    _medium = dipoleret._medium
    _eps_med = dipoleret._eps_med
    _host = dipoleret._host

    def _potential(self, p, w, z0=0.0, wimage=None):
        pts, dip = self._host()
        phip = dev.colmajor((p.n, self.npol))
        w = np.ascontiguousarray(w, dtype=np.complex128)
        wi = None if wimage is None else np.ascontiguousarray(wimage, dtype=np.complex128)
        _lib.call("mnp_dipolestat_potential", p.green.handle, len(pts), pts.ctypes.data_as(C.c_void_p),
                  dip.shape[2], dip.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p), float(z0),
                  None if wi is None else wi.ctypes.data_as(C.c_void_p), dev.ptr(dev.base(phip)))
        return phip

    def __call__(self, p, enei):
        """exc(p, enei): phi' = -nvec . E of the dipoles (potential.m:17)."""
        w = np.full(len(self.pt), 1.0 / complex(self._eps_med(p, enei)))
        return compstruct(p, enei, phip=self._potential(p, w))

    potential = __call__

    def field_own(self, green, sig):
        """E = -Gp sig (k = 0) at each dipole from its own column, npol x 3."""
        pts, _ = self._host()
        ncol = sig.shape[1]
        e = np.zeros((ncol, 3), dtype=np.complex128)
        _lib.call("mnp_stat_field_own", green.handle, len(pts), pts.ctypes.data_as(C.c_void_p), ncol,
                  dev.ptr(dev.base(sig)), e.ctypes.data_as(C.c_void_p))
        return e

    @staticmethod
    def induced_dipole(sig):
        """sum area pos sig per column, npol x 3 (FP64 device reduction)."""
        npol = sig.sig.shape[-1]
        dip = np.zeros((npol, 3), dtype=np.complex128)
        _lib.call("mnp_stat_dipole_moment", sig.p.green.handle, npol, dev.ptr(dev.base(sig.sig)),
                  dip.ctypes.data_as(C.c_void_p))
        return dip

    def decayrate(self, sig):
        """[tot, rad, rad0], each npt x ndip (decayrate.m:25-56): tot = 1 +
        Im(e . dip) / (0.5 nb gamma), rad = |nb^2 p_ind + dip|^2."""
        p, enei = sig.p, sig.enei
        npt, ndip = self.dip.shape[0], self.dip.shape[2]
        nb = np.sqrt(complex(self._eps_med(p, enei)))
        gamma = 4 / 3 * (2 * np.pi / enei) ** 3
        e = self.field_own(p.green, sig.sig)
        indip = self.induced_dipole(sig)
        tot, rad = np.zeros((npt, ndip)), np.zeros((npt, ndip))
        for ip in range(npt):
            for i in range(ndip):
                q = ip + npt * i
                d = self.dip[ip, :, i]
                rad[ip, i] = np.linalg.norm(nb ** 2 * indip[q] + d) ** 2
                tot[ip, i] = np.real(1 + np.imag(e[q] @ d) / (0.5 * nb * gamma))
        return tot, rad, np.full((npt, ndip), np.real(nb * gamma))


class dipolestatmirror:
    """Quasistatic dipoles for a comparticlemirror (Simulation/static/
    @dipolestatmirror); ``pt``, ``dip``, ``full`` as for dipolestat, medium
    the particle's outside.  The potential holds J excitations of the
    symmetry part: unit dipoles along x, y, z at the points and at their
    mirror images (moments mirrored), combined with the signs of each
    symmetry value (init.m:11-25, potential.m:19-41); ``full(sig)`` averages
    their expansions to the full particle, which leaves the solution of the
    unit dipoles at the points alone, and contracts it with the moments
    (full.m:14-43).  Decay rates are dipolestat's on the full particle."""

    # potential.m:28-40: (sign of each mirror set, symmetry keys x, y, z)
    _COMB = {"x": [((1, 1), ("-", "+", "+")), ((1, -1), ("+", "-", "-"))],
             "y": [((1, 1), ("+", "-", "+")), ((1, -1), ("-", "+", "-"))],
             "xy": [((1, 1, 1, 1), ("-+", "+-", "++")), ((1, -1, 1, -1), ("++", "--", "-+")),
                    ((1, 1, -1, -1), ("--", "++", "+-")), ((1, -1, -1, 1), ("+-", "-+", "--"))]}
    # mirror sets: (flipped axes, moment signs) (init.m:14-24)
    _SETS = {"x": [((), (1, 1, 1)), ((0,), (-1, 1, 1))], "y": [((), (1, 1, 1)), ((1,), (1, -1, 1))],
             "xy": [((), (1, 1, 1)), ((0,), (-1, 1, 1)), ((1,), (1, -1, 1)), ((0, 1), (-1, -1, 1))]}

    def __init__(self, pt, dip=None, op=None, full=False):
        self.dip = dipolestat(pt, dip, op, full=full)
        self.pt = self.dip.pt

    def __call__(self, p, enei):
        pf = p.full()
        npt = len(self.pt)
        w = np.full(npt, 1.0 / complex(p.eps2(enei)))
        phis = []
        for axes, sgn in self._SETS[p.sym]:
            pts = self.pt.copy()
            for a in axes:
                pts[:, a] = -pts[:, a]
            m = dipolestat(pts, np.diag(np.asarray(sgn, dtype=np.float64)))
            phis.append(m._potential(pf, w)[:p.n])   # the symmetry part: first n faces
        exc = compstructmirror(p, enei, self.full)
        for signs, keys in self._COMB[p.sym]:
            v = dev.colmajor((p.n, 3 * npt))
            v.copy_(sum(s * ph for s, ph in zip(signs, phis)))
            exc.val.append(compstruct(p, enei, symval=p.symvalue(list(keys)), phip=v))
        return exc

    potential = __call__

    def full(self, sig):
        """Surface charge of the full particle for the dipoles' moments (full.m)."""
        import torch
        p = sig.p
        pf = p.full()
        J, npt = len(p.symtable), len(self.pt)
        vmean = None
        for val in sig.val:
            v = torch.cat([float(s) * val.sig for s in val.symval[-1]], dim=0) / J
            vmean = v if vmean is None else vmean + v
        dip = self.dip.dip                                   # npt x 3 x ndip
        out = dev.colmajor((pf.n, npt * dip.shape[2]))
        out.zero_()
        for j in range(dip.shape[2]):
            for c in range(3):
                wgt = torch.as_tensor(dip[:, c, j], dtype=torch.float64, device=out.device)
                out[:, npt * j:npt * (j + 1)] += vmean[:, npt * c:npt * (c + 1)] * wgt[None, :]
        return compstruct(pf, sig.enei, sig=out)

    def decayrate(self, sig):
        """[tot, rad, rad0] of the dipoles on the full particle (decayrate.m)."""
        return self.dip.decayrate(self.full(sig))


class dipoleretmirror:
    """Retarded dipoles for a comparticlemirror (Simulation/retarded/
    @dipoleretmirror): the construction of dipolestatmirror with dipoleret's
    potentials phi2, phi2p, a2, a2p (init.m, potential.m); ``full(sig)``
    expands sig1, sig2 (scalar key) and h1, h2 (per-component keys),
    averages and contracts with the moments (full.m); decay rates are
    dipoleret's on the full particle (decayrate.m)."""

    _COMB = dipolestatmirror._COMB
    _SETS = dipolestatmirror._SETS

    def __init__(self, pt, dip=None, op=None, full=False, pinfty=None):
        self.dip = dipoleret(pt, dip, op, pinfty=pinfty, full=full)
        self.pt = self.dip.pt

    def __call__(self, p, enei):
        pf = p.full()
        npt, n = len(self.pt), p.n
        sets = []
        for axes, sgn in self._SETS[p.sym]:
            pts = self.pt.copy()
            for a in axes:
                pts[:, a] = -pts[:, a]
            m = dipoleret(pts, np.diag(np.asarray(sgn, dtype=np.float64)), pinfty=(np.zeros((1, 3)), np.zeros(1)))
            pot = m(pf, enei)
            sets.append({k: pot[k][:n] for k in ("phi2", "phi2p", "a2", "a2p")})
        exc = compstructmirror(p, enei, self.full)
        for signs, keys in self._COMB[p.sym]:
            fields = {}
            for k in ("phi2", "phi2p", "a2", "a2p"):
                shape = (n, 3 * npt) if k.startswith("phi") else (n, 3, 3 * npt)
                v = dev.colmajor(shape)
                v.copy_(sum(sg * st[k] for sg, st in zip(signs, sets)))
                fields[k] = v
            exc.val.append(compstruct(p, enei, symval=p.symvalue(list(keys)), **fields))
        return exc

    potential = __call__

    def full(self, sig):
        """Charges and currents of the full particle for the dipoles' moments (full.m)."""
        import torch
        p = sig.p
        pf = p.full()
        J, npt = len(p.symtable), len(self.pt)
        dip = self.dip.dip
        out = {}
        for key in ("sig1", "sig2", "h1", "h2"):
            vmean = None
            for val in sig.val:
                sv = val.symval
                if key.startswith("sig"):
                    v = torch.cat([float(sv[-1, j]) * val[key] for j in range(J)], dim=0)
                else:
                    sc = torch.as_tensor(sv, dtype=torch.float64, device=val[key].device)  # 3 x J
                    v = torch.cat([val[key] * sc[:, j][None, :, None] for j in range(J)], dim=0)
                vmean = v / J if vmean is None else vmean + v / J
            shape = (pf.n, npt * dip.shape[2]) if key.startswith("sig") else (pf.n, 3, npt * dip.shape[2])
            o = dev.colmajor(shape)
            o.zero_()
            for j in range(dip.shape[2]):
                for c in range(3):
                    wgt = torch.as_tensor(dip[:, c, j], dtype=torch.float64, device=o.device)
                    o[..., npt * j:npt * (j + 1)] += vmean[..., npt * c:npt * (c + 1)] * wgt
            out[key] = o
        return compstruct(pf, sig.enei, **out)

    def decayrate(self, sig):
        """[tot, rad, rad0] of the dipoles on the full particle (decayrate.m)."""
        return self.dip.decayrate(self.full(sig))


def dipole(pt, dip=None, op=None, **kw):
    """Base/dipole.m: dipoleret for op.sim == 'ret' (dipoleretlayer with
    op.layer), dipolestat for 'stat', dipolestatlayer for 'stat' with
    op.layer."""
    if op is not None and op.get("sim", "ret") == "stat":
        if op.get("sym"):
            if op.get("layer") is not None:
                raise NotImplementedError("mirror symmetry with layers is not built")
            return dipolestatmirror(pt, dip, op, **kw)
        if op.get("layer") is not None:
            from .layer import dipolestatlayer
            return dipolestatlayer(pt, dip, op, **kw)
        return dipolestat(pt, dip, op, **kw)
    if op is not None and op.get("layer") is not None:
        if op.get("sym"):
            raise NotImplementedError("mirror symmetry with layers is not built")
        from .layer import dipoleretlayer
        kw.pop("medium", None)  # the dipoles lie in the upper medium
        return dipoleretlayer(pt, dip, op, **kw)
    if op is not None and op.get("sym"):
        return dipoleretmirror(pt, dip, op, **kw)
    return dipoleret(pt, dip, op, **kw)
