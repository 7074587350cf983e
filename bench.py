"""Benchmark: wavelengths/sec of the 20,480-face retarded BEM extinction spectrum.

Workload (one GPU): prolate ellipsoid with semi-axes 10 x 10 x 40 nm from an
icosphere subdivided 5 times (20,480 flat faces), Drude gold
eps = 10 - 9.065^2 / (w (w + 0.0708 i)) in a medium of eps = 1.77, plane
waves polarised x (propagating z) and z (propagating x), wavelengths from
500 to 1000 nm.  One step = one wavelength: assembly of G1, G2, H1, H2, the
bemret chain, the solve for both polarisations and the FP64 extinction and
scattering cross sections, with every input resident on the GPU.

The JSON line also carries `hmatrix_dipole`: the H-matrix iterative path
with dipole excitation on an 81,920-face sphere (16 wavelengths, decay
rates), with the H-matrix product's achieved bandwidth.

`value` is the mixed-precision path (complex FP32 chain on MFMA used as the
preconditioner of the exact FP64 operator, refined to a relative residual of
1e-12).  The FP64 drop-in is timed on `--fp64-steps` wavelengths for the
speed-up and the cross-section error of the mixed path is reported against
it.  `cpu_baseline` times the NumPy/OpenBLAS FP64 oracle on the host for one
wavelength of the same shape at 5,120 faces, its assembly, chain and solve
apart, and scales the chain by N^3 and the rest by N^2 to 20,480 faces.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "wavelengths/sec, 20k-face retarded BEM, FP32-MFMA+refine vs FP64 @1e-6 rel err"
DRUDE = (10.0, 9.065, 0.0708)
EPS_OUT = 1.77
POL = [[1.0, 0, 0], [0, 0, 1.0]]
DIR = [[0, 0, 1.0], [1.0, 0, 0]]
PEAK = {"fp32": 157.3, "fp64": 78.6, "bf16": 2500.0}  # dense MFMA TFLOP/s, MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(level, steps, warmup, fp64_steps, split_bf16=True):
    import torch

    from mnpbem_amd import _lib, bem as B, device as dev, shapes

    dev.init(0)
    v, f = shapes.ellipsoid(level, (10.0, 10.0, 40.0))
    n = len(f)
    op = B.bemoptions(sim="ret")
    p = B.comparticle([B.epsconst(EPS_OUT), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, op)
    exc = B.planewave(POL, DIR, op)
    t0 = time.time()
    p.green  # geometry, refinement tables, sum rule (once per mesh)
    torch.cuda.synchronize()
    log(f"[bench] n={n} refined pairs={p.green.nnz} setup {time.time() - t0:.2f}s")
    # the timed wavelengths spread over the 64 of the spectrum; the FP64
    # drop-in runs on fp64_steps of them, spread likewise
    spec = np.linspace(500.0, 1000.0, 64)
    lam_t = spec[np.unique(np.round(np.linspace(0, 63, steps)).astype(int))]
    lam_w = spec[1:1 + warmup]
    lam64 = lam_t[np.unique(np.round(np.linspace(0, len(lam_t) - 1, min(fp64_steps, len(lam_t)))).astype(int))]

    def spectrum(bem, lams, timed):
        out = []
        if timed:
            _lib.call("mnp_timers_enable", 1)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for lam in lams:
            sig = bem.solve(exc(p, lam))
            out.append(exc.cross_sections(sig))
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        tm = None
        if timed:
            tm = _lib.Timers()
            _lib.call("mnp_timers_read", C.byref(tm))
            _lib.call("mnp_timers_enable", 0)
        return out, el, tm

    mixed = B.bemsolver(p, op, precision="mixed")
    if warmup:
        spectrum(mixed, lam_w, False)
        log(f"[bench] mixed warmup done ({warmup})")
    mixed.enei = None  # every timed step initialises its wavelength
    res_mx, t_mx, tm = spectrum(mixed, lam_t, True)
    sweeps = mixed.stats()
    log(f"[bench] mixed {steps} wavelengths in {t_mx:.2f}s, refinement {sweeps}")
    del mixed
    torch.cuda.empty_cache()

    res_bf, t_bf, tm_bf, sweeps_bf = [], 0.0, None, None
    if split_bf16:
        sb = B.bemsolver(p, op, precision="mixed_bf16")
        if warmup:
            spectrum(sb, lam_w, False)
        sb.enei = None
        res_bf, t_bf, tm_bf = spectrum(sb, lam_t, True)
        sweeps_bf = sb.stats()
        log(f"[bench] mixed_bf16 {steps} wavelengths in {t_bf:.2f}s, refinement {sweeps_bf}")
        del sb
        torch.cuda.empty_cache()

    res_64, t_64, tm64 = [], 0.0, None
    if fp64_steps:
        fp64 = B.bemsolver(p, op, precision="fp64")
        res_64, t_64, tm64 = spectrum(fp64, lam64, True)
        log(f"[bench] fp64 {len(lam64)} wavelengths in {t_64:.2f}s")
        del fp64
    where = {float(l): i for i, l in enumerate(lam_t)}

    def max_err(res):
        # extinction and scattering, both polarisations, at the FP64 wavelengths
        e = [np.abs(res[where[float(l)]][c] - b[c]) / np.abs(b[c]) for l, b in zip(lam64, res_64) for c in (0, 1)]
        return float(np.max(e)) if e else None
    return dict(n=n, t_mx=t_mx, t_64=t_64, n64=len(lam64), timers=tm, timers64=tm64,
                max_err=max_err(res_mx), sweeps=sweeps, ext=[list(map(float, r[0])) for r in res_mx],
                t_bf=t_bf, timers_bf=tm_bf, sweeps_bf=sweeps_bf,
                max_err_bf=max_err(res_bf) if res_bf else None,
                lam_t=[float(x) for x in lam_t], lam64=[float(x) for x in lam64])


def run_quasistatic(level=5, steps=3):
    """Quasistatic solver (bemstat) on the same 20,480-face ellipsoid and
    wavelengths as the headline: one complex LU of Lambda + F per wavelength,
    mixed (FP32 factors + FP64 refinement) and FP64; extinction from the
    induced dipole."""
    import torch

    from mnpbem_amd import bem as B, shapes

    v, f = shapes.ellipsoid(level, (10.0, 10.0, 40.0))
    op = B.bemoptions(sim="stat")
    p = B.comparticle([B.epsconst(EPS_OUT), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, op)
    exc = B.planewave(POL, op=op)
    lams = np.linspace(500.0, 1000.0, 64)
    res = {}
    for prec in ("mixed", "mixed_bf16", "fp64"):
        bem = B.bemsolver(p, op, precision=prec)
        exc.ext(bem.solve(exc(p, lams[0])))  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        ext = [exc.ext(bem.solve(exc(p, lam))) for lam in lams[1:1 + steps]]
        torch.cuda.synchronize()
        res[prec] = (steps / (time.perf_counter() - t), ext)
        del bem
    err = max(float(np.max(np.abs(a / b - 1))) for a, b in zip(res["mixed"][1], res["fp64"][1]))
    err_bf = max(float(np.max(np.abs(a / b - 1))) for a, b in zip(res["mixed_bf16"][1], res["fp64"][1]))
    log(f"[bench] quasistatic: mixed {res['mixed'][0]:.2f}/s, mixed_bf16 {res['mixed_bf16'][0]:.2f}/s, "
        f"fp64 {res['fp64'][0]:.2f}/s")
    # iterative twin (bemstatiter, bemiter defaults): F filled once per mesh
    it = B.bemstatiter(p, dict(op, iter=dict(cleaf=200, htol=1e-6, tol=1e-6)))
    torch.cuda.synchronize()
    t = time.perf_counter()
    it(lams[0])
    torch.cuda.synchronize()
    t_fill = time.perf_counter() - t
    exc.ext(it.solve(exc(p, lams[0])))
    torch.cuda.synchronize()
    t = time.perf_counter()
    ext_it, iters = [], []
    for lam in lams[1:1 + steps]:
        ext_it.append(exc.ext(it.solve(exc(p, lam))))
        iters.append(it.stats()["iter"])
    torch.cuda.synchronize()
    v_it = steps / (time.perf_counter() - t)
    err_it = max(float(np.max(np.abs(a / b - 1))) for a, b in zip(ext_it, res["fp64"][1]))
    compression = it.stats()["compression"]
    del it
    log(f"[bench] quasistatic iterative: fill {t_fill:.3f} s, {v_it:.1f}/s, iterations {iters}, "
        f"rel err vs dense FP64 {err_it:.2e}")
    return {"workload": f"quasistatic bemstat, same ellipsoid ({len(f)} faces) and wavelengths as the headline",
            "wavelengths_per_s_mixed": res["mixed"][0], "wavelengths_per_s_mixed_bf16": res["mixed_bf16"][0],
            "wavelengths_per_s_fp64": res["fp64"][0],
            "speedup": res["mixed"][0] / res["fp64"][0], "speedup_mixed_bf16": res["mixed_bf16"][0] / res["fp64"][0],
            "max_rel_err_vs_fp64": err, "max_rel_err_mixed_bf16_vs_fp64": err_bf,
            "iterative": {"solver": "bemstatiter, FP64 H-matrix of F (cleaf 200, htol 1e-6), GMRES tol 1e-6, "
                                    "block preconditioner",
                          "fill_s": t_fill, "wavelengths_per_s": v_it, "gmres_iterations": iters,
                          "compression": compression, "max_rel_err_ext_vs_dense_fp64": err_it}}


def run_mirror(steps=3):
    """Mirror symmetry 'xy' (bemretmirror): the quarter x, y >= 0 of the
    prolate 10 x 10 x 40 nm ellipsoid meshed by octants (2 x 50^2 = 5,000
    faces), completed to 20,000 faces; Drude gold in eps = 1.77, pol x and
    pol y along +z, the headline's wavelengths.  FP64, mixed and split-bf16
    mirror solvers, and the mixed solver without symmetry on the same
    20,000-face particle; cross sections compared with the FP64 mirror
    solver."""
    import torch

    from mnpbem_amd import bem as B, shapes

    part = shapes.ellipsoid_octants(50, (10.0, 10.0, 40.0), "xy")
    op = B.bemoptions(sim="ret", sym="xy")
    p = B.comparticlemirror([B.epsconst(EPS_OUT), B.epsdrude(*DRUDE)], [part], [2, 1], 1, op)
    pol, dirs = [[1.0, 0, 0], [0, 1.0, 0]], [[0, 0, 1.0], [0, 0, 1.0]]
    exc = B.planewave(pol, dirs, op)
    lams = np.linspace(500.0, 1000.0, 64)

    def timed(bem, ex, part_, nst):
        ex.ext(bem.solve(ex(part_, lams[0])))  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        cs = [np.array(ex.cross_sections(bem.solve(ex(part_, lam)))) for lam in lams[1:1 + nst]]
        torch.cuda.synchronize()
        return nst / (time.perf_counter() - t), cs

    res, sweeps = {}, {}
    for prec, nst in (("mixed", steps), ("mixed_bf16", steps), ("fp64", 1)):
        bem = B.bemsolver(p, op, precision=prec)
        res[prec] = timed(bem, exc, p, nst)
        sweeps[prec] = bem.val_stats
        del bem
    fop = B.bemoptions(sim="ret")
    pf = B.comparticle(p.eps, [(p.full().verts, p.full().faces)], [2, 1], 1, fop)
    bem = B.bemsolver(pf, fop, precision="mixed")
    res["full_mixed"] = timed(bem, B.planewave(pol, dirs, fop), pf, 1)
    del bem
    err = {k: max(float(np.max(np.abs(a / b - 1))) for a, b in zip(res[k][1], res["fp64"][1]))
           for k in ("mixed", "mixed_bf16", "full_mixed")}
    log(f"[bench] mirror xy: mixed {res['mixed'][0]:.2f}/s, mixed_bf16 {res['mixed_bf16'][0]:.2f}/s, "
        f"fp64 {res['fp64'][0]:.2f}/s, full particle mixed {res['full_mixed'][0]:.3f}/s; rel err vs FP64 {err}")
    # quasistatic twin (bemstatmirror) on the same part, and bemstat on the full particle
    sop = B.bemoptions(sim="stat", sym="xy")
    ps = B.comparticlemirror(p.eps, [part], [2, 1], 1, sop)
    sexc = B.planewave(pol, None, sop)
    sres = {}
    for prec in ("mixed", "fp64"):
        bem = B.bemsolver(ps, sop, precision=prec)
        sres[prec] = timed(bem, sexc, ps, 8)
        del bem
    fsop = B.bemoptions(sim="stat")
    psf = B.comparticle(p.eps, [(p.full().verts, p.full().faces)], [2, 1], 1, fsop)
    bem = B.bemsolver(psf, fsop, precision="mixed")
    sres["full_mixed"] = timed(bem, B.planewave(pol, None, fsop), psf, 8)
    del bem
    serr = {k: max(float(np.max(np.abs(a / b - 1))) for a, b in zip(sres[k][1], sres["fp64"][1]))
            for k in ("mixed", "full_mixed")}
    log(f"[bench] quasistatic mirror xy: mixed {sres['mixed'][0]:.1f}/s, fp64 {sres['fp64'][0]:.1f}/s, "
        f"full particle mixed {sres['full_mixed'][0]:.2f}/s; rel err vs FP64 {serr}")
    return {"workload": f"mirror symmetry xy: quarter of the ellipsoid 10x10x40 nm meshed by octants, {p.n} faces "
                        f"of {p.full().n}, pol x and y along +z, wavelengths from linspace(500,1000,64)",
            "wavelengths_per_s_mixed": res["mixed"][0], "wavelengths_per_s_mixed_bf16": res["mixed_bf16"][0],
            "wavelengths_per_s_fp64": res["fp64"][0], "wavelengths_per_s_full_particle_mixed": res["full_mixed"][0],
            "speedup_mixed_vs_full_particle_mixed": res["mixed"][0] / res["full_mixed"][0],
            "refinement_sweeps_mixed": sweeps["mixed"], "refinement_sweeps_mixed_bf16": sweeps["mixed_bf16"],
            "max_rel_err_vs_fp64": err["mixed"], "max_rel_err_mixed_bf16_vs_fp64": err["mixed_bf16"],
            "max_rel_err_full_particle_vs_mirror_fp64": err["full_mixed"],
            "quasistatic": {"solver": "bemstatmirror, same part and pols, 8 wavelengths",
                            "wavelengths_per_s_mixed": sres["mixed"][0], "wavelengths_per_s_fp64": sres["fp64"][0],
                            "wavelengths_per_s_full_particle_mixed": sres["full_mixed"][0],
                            "max_rel_err_vs_fp64": serr["mixed"],
                            "max_rel_err_full_particle_vs_mirror_fp64": serr["full_mixed"]}}


def run_substrate(level=5, steps=3):
    """Quasistatic solver above a glass substrate (bemstatlayer): the headline
    ellipsoid (20,480 faces) with its lowest point 5 nm above the interface,
    air above and eps = 2.25 below, pol x at normal incidence and a TM wave at
    0.6 rad; mixed, mixed_bf16 and FP64."""
    import torch

    from mnpbem_amd import bem as B, shapes
    from mnpbem_amd.layer import layerstructure

    v, f = shapes.ellipsoid(level, (10.0, 10.0, 40.0))
    v = v + np.array([0, 0, 45.0])
    epstab = [B.epsconst(1.0), B.epsdrude(*DRUDE), B.epsconst(2.25)]
    op = B.bemoptions(sim="stat", layer=layerstructure(epstab, [1, 3], 0.0))
    p = B.comparticle(epstab, [(v, f)], [2, 1], 1, op)
    th = 0.6
    exc = B.planewave([[1.0, 0, 0], [np.cos(th), 0, np.sin(th)]], [[0, 0, -1.0], [np.sin(th), 0, -np.cos(th)]], op)
    lams = np.linspace(500.0, 1000.0, 64)
    res = {}
    for prec in ("mixed", "mixed_bf16", "fp64"):
        bem = B.bemsolver(p, op, precision=prec)
        exc.ext(bem.solve(exc(p, lams[0])))  # warm-up (also fills the image Green function)
        torch.cuda.synchronize()
        t = time.perf_counter()
        cs = [np.array(exc.cross_sections(bem.solve(exc(p, lam)))) for lam in lams[1:1 + steps]]
        torch.cuda.synchronize()
        res[prec] = (steps / (time.perf_counter() - t), cs)
        del bem
    err = {k: max(float(np.max(np.abs(a / b - 1))) for a, b in zip(res[k][1], res["fp64"][1]))
           for k in ("mixed", "mixed_bf16")}
    log(f"[bench] substrate: mixed {res['mixed'][0]:.2f}/s, mixed_bf16 {res['mixed_bf16'][0]:.2f}/s, "
        f"fp64 {res['fp64'][0]:.2f}/s; rel err vs FP64 {err}")
    return {"workload": f"quasistatic bemstatlayer: ellipsoid 10x10x40 nm ({len(f)} faces) 5 nm above "
                        "eps 2.25, air above, pol x normal incidence and TM at 0.6 rad",
            "wavelengths_per_s_mixed": res["mixed"][0], "wavelengths_per_s_mixed_bf16": res["mixed_bf16"][0],
            "wavelengths_per_s_fp64": res["fp64"][0],
            "max_rel_err_vs_fp64": err["mixed"], "max_rel_err_mixed_bf16_vs_fp64": err["mixed_bf16"]}


def run_substrate_ret(level=4, steps=4):
    """Retarded solver above a glass substrate (bemretlayer, complex FP64):
    the 10 x 10 x 40 nm ellipsoid at 5,120 faces 5 nm above eps = 2.25, air
    above, five TM plane waves from above at 0..80 degrees (the excitation of
    the reference's substrate spectra demo), the reflected Green function
    tabulated on the 30 x 30 tabspace grid at linspace(350, 800, 5) and
    interpolated in wavelength; per wavelength the reflected and direct
    matrices, the chain, the solve and the cross sections."""
    import torch

    from mnpbem_amd import bem as B, shapes
    from mnpbem_amd.layer import compgreentablayer, layerstructure

    v, f = shapes.ellipsoid(level, (10.0, 10.0, 40.0))
    v = v + np.array([0, 0, 45.0])
    epstab = [B.epsconst(1.0), B.epsdrude(*DRUDE), B.epsconst(2.25)]
    layer = layerstructure(epstab, [1, 3], 0.0)
    p = B.comparticle(epstab, [(v, f)], [2, 1], 1, B.bemoptions(sim="ret"))
    torch.cuda.synchronize()
    t = time.perf_counter()
    tab = compgreentablayer(layer, layer.tabspace(p)).set(np.linspace(350.0, 800.0, 5))
    torch.cuda.synchronize()
    t_tab = time.perf_counter() - t
    op = B.bemoptions(sim="ret", layer=layer, greentab=tab)
    theta = np.linspace(0, 80, 5) * np.pi / 180
    dirs = np.column_stack([np.sin(theta), 0 * theta, -np.cos(theta)])
    pols = np.column_stack([np.cos(theta), 0 * theta, np.sin(theta)])
    exc = B.planewave(pols, dirs, op)
    lams = np.linspace(450.0, 750.0, 20)
    res = {}
    for prec in ("mixed", "fp64"):
        t = time.perf_counter()
        bem = B.bemsolver(p, op, precision=prec)
        torch.cuda.synchronize()
        t_init = time.perf_counter() - t
        exc.cross_sections(bem.solve(exc(p, lams[0])))  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        cs, sweeps = [], []
        for lam in lams[1:1 + steps]:
            cs.append(exc.cross_sections(bem.solve(exc(p, lam))))
            sweeps.append(bem.stats()["iterations"])
        torch.cuda.synchronize()
        res[prec] = (steps / (time.perf_counter() - t), cs, sweeps, bem.stats()["fallbacks"], t_init)
        del bem
    err = max(float(np.max(np.abs(a[j] - b[j]) / np.max(np.abs(b[j]))))
              for a, b in zip(res["mixed"][1], res["fp64"][1]) for j in (0, 1))
    log(f"[bench] substrate (retarded): mixed {res['mixed'][0]:.3f}/s (sweeps {res['mixed'][2]}), "
        f"fp64 {res['fp64'][0]:.3f}/s, rel err {err:.1e}; tables {t_tab:.2f} s")
    return {"workload": f"retarded bemretlayer: ellipsoid 10x10x40 nm ({len(f)} faces) 5 nm above eps 2.25, "
                        "air above, 5 TM plane waves from above at 0-80 deg, 30x30 table set at 5 wavelengths",
            "wavelengths_per_s_mixed": res["mixed"][0], "wavelengths_per_s_fp64": res["fp64"][0],
            "speedup_mixed_vs_fp64": res["mixed"][0] / res["fp64"][0],
            "max_rel_err_mixed_vs_fp64": err, "refinement_sweeps": res["mixed"][2],
            "table_set_s": t_tab, "refinement_setup_s": res["fp64"][4],
            "pivoting_fallbacks": {"mixed": res["mixed"][3], "fp64": res["fp64"][3]},
            "ext_first": [float(x) for x in res["fp64"][1][0][0]],
            "sca_first": [float(x) for x in res["fp64"][1][0][1]]}


def run_substrate_iter(steps=4):
    """Iterative retarded solver above a glass substrate (bemretlayeriter):
    the workload of run_substrate_ret (10 x 10 x 40 nm ellipsoid 5 nm above
    eps = 2.25, five TM plane waves, 30 x 30 tables set at 5 wavelengths) at
    5,120 faces against the dense FP64 solver, and at 20,480 faces alone;
    cleaf 32, htol 1e-6, FP32 H-matrix storage, GMRES tol 1e-6."""
    import torch

    from mnpbem_amd import bem as B, shapes
    from mnpbem_amd.layer import compgreentablayer, layerstructure

    theta = np.linspace(0, 80, 5) * np.pi / 180
    dirs = np.column_stack([np.sin(theta), 0 * theta, -np.cos(theta)])
    pols = np.column_stack([np.cos(theta), 0 * theta, np.sin(theta)])
    lams = np.linspace(450.0, 750.0, 20)
    out = {}
    for level in (4, 5):
        v, f = shapes.ellipsoid(level, (10.0, 10.0, 40.0))
        v = v + np.array([0, 0, 45.0])
        epstab = [B.epsconst(1.0), B.epsdrude(*DRUDE), B.epsconst(2.25)]
        layer = layerstructure(epstab, [1, 3], 0.0)
        p = B.comparticle(epstab, [(v, f)], [2, 1], 1, B.bemoptions(sim="ret"))
        tab = compgreentablayer(layer, layer.tabspace(p)).set(np.linspace(350.0, 800.0, 5))
        op = B.bemoptions(sim="ret", layer=layer, greentab=tab)
        exc = B.planewave(pols, dirs, op)
        torch.cuda.synchronize()
        t = time.perf_counter()
        it = B.bemsolver(p, dict(op, iter=dict(cleaf=32, htol=1e-6, tol=1e-6, hprecision="fp32")))
        exc.cross_sections(it.solve(exc(p, lams[0])))  # warm-up
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t
        t = time.perf_counter()
        cs, iters = [], []
        for lam in lams[1:1 + steps]:
            cs.append(exc.cross_sections(it.solve(exc(p, lam))))
            iters.append(it.stats()["iter"])
        torch.cuda.synchronize()
        rate = steps / (time.perf_counter() - t)
        r = {"faces": len(f), "wavelengths_per_s": rate, "gmres_iterations": iters, "setup_s": t_setup,
             "ext_first": [float(x) for x in cs[0][0]]}
        if level == 4:
            dense = B.bemsolver(p, op, precision="fp64")
            err = max(float(np.max(np.abs(a[j] - b[j]) / np.max(np.abs(b[j]))))
                      for lam, a in zip(lams[1:1 + steps], cs)
                      for b in [exc.cross_sections(dense.solve(exc(p, lam)))] for j in (0, 1))
            r["max_rel_err_vs_dense_fp64"] = err
            del dense
        del it
        log(f"[bench] substrate (iterative) {len(f)} faces: {rate:.3f}/s, iterations {iters}"
            + (f", rel err vs dense {r['max_rel_err_vs_dense_fp64']:.1e}" if level == 4 else ""))
        out[f"faces_{len(f)}"] = r
    out["workload"] = ("retarded bemretlayeriter: ellipsoid 10x10x40 nm 5 nm above eps 2.25, air above, 5 TM "
                       "plane waves from above at 0-80 deg, cleaf 32, htol 1e-6, FP32 H-matrices, GMRES tol 1e-6")
    return out


def run_substrate_dipole(steps=4):
    """Dipoles above a glass substrate (dipoleretlayer): three orientations
    at one point 5 nm above the tip of the 10 x 10 x 40 nm ellipsoid (5,120
    faces) 5 nm above eps = 2.25, air above; per wavelength the excitation
    (direct plus reflected potentials), the dense FP64 bemretlayer solve
    and the total and radiative decay rates (induced and reflected fields at
    the dipole, far fields over the 256-point sphere); 30 x 30 tables set at
    5 wavelengths."""
    import torch

    from mnpbem_amd import bem as B, shapes
    from mnpbem_amd.layer import compgreentablayer, layerstructure

    v, f = shapes.ellipsoid(4, (10.0, 10.0, 40.0))
    v = v + np.array([0, 0, 45.0])
    pts = np.array([[0.0, 0.0, 90.0]])  # the tip is at z = 85
    epstab = [B.epsconst(1.0), B.epsdrude(*DRUDE), B.epsconst(2.25)]
    layer = layerstructure(epstab, [1, 3], 0.0)
    p = B.comparticle(epstab, [(v, f)], [2, 1], 1, B.bemoptions(sim="ret"))
    tab = compgreentablayer(layer, layer.tabspace(p, dip=pts)).set(np.linspace(350.0, 800.0, 5))
    op = B.bemoptions(sim="ret", layer=layer, greentab=tab)
    exc = B.dipole(pts, np.eye(3), op)
    lams = np.linspace(450.0, 750.0, 20)
    bem = B.bemsolver(p, op)
    exc.decayrate(bem.solve(exc(p, lams[0])))  # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    rates = []
    for lam in lams[1:1 + steps]:
        tot, rad, _ = exc.decayrate(bem.solve(exc(p, lam)))
        rates.append((tot.ravel().tolist(), rad.ravel().tolist()))
    torch.cuda.synchronize()
    rate = steps / (time.perf_counter() - t)
    log(f"[bench] substrate dipole {len(f)} faces: {rate:.3f} wavelengths/s, tot {rates[0][0]}")
    return {"workload": "retarded dipoleretlayer + bemretlayer (FP64): 3 dipole orientations 5 nm above the tip of "
                        "the 10x10x40 nm ellipsoid (5120 faces) 5 nm above eps 2.25, air above",
            "faces": len(f), "wavelengths_per_s": rate, "tot_rad_first": rates[0]}


def run_sensitivity(nsweep=32):
    """Config 5: extinction and its gradient with respect to all vertex
    coordinates and the Drude parameters (eps_inf, wp, gamma) by the adjoint
    (one transposed solve with the forward factors per wavelength), on the
    5,120-face ellipsoid 10 x 10 x 40 nm in eps = 1.77, pol x (dir z) and z
    (dir x), of the summed extinction of both polarisations.  At 650 nm the
    vertex gradient is checked against central differences of the FP64
    solver (step 1e-4 nm per vertex, refined pairs held fixed) along seeded
    directions and the Drude gradient against central differences (relative
    step 1e-6); the sweep linspace(500, 1000, 32) is timed with the mixed and
    FP64 solvers."""
    import torch

    from mnpbem_amd import bem as B, shapes
    from mnpbem_amd.sensitivity import drude_derivatives, extinction_gradient

    v, f = shapes.ellipsoid(4, (10.0, 10.0, 40.0))

    def make(prec):
        op = B.bemoptions(sim="ret")
        p = B.comparticle([B.epsconst(EPS_OUT), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, op)
        return p, B.bemsolver(p, op, precision=prec), B.planewave(POL, DIR, op)

    def gradient(p, bem, exc, lam):
        sig = bem.solve(exc(p, lam))
        ext = float(np.sum(exc.ext(sig)))
        dv, de = extinction_gradient(exc, bem, sig)
        dd = drude_derivatives(p.eps[1], lam)
        return ext, dv, de[0] * dd.real + de[1] * dd.imag

    p, b64, exc = make("fp64")
    ext, dv, dpar = gradient(p, b64, exc, 650.0)
    # Drude parameters: central differences, relative step 1e-6
    fdp = np.zeros(3)
    for j in range(3):
        h = 1e-6 * DRUDE[j]
        vals = []
        for sgn in (1, -1):
            par = list(DRUDE)
            par[j] += sgn * h
            p.eps[1] = B.epsdrude(*par)
            b64.enei = None
            vals.append(float(np.sum(exc.ext(b64.solve(exc(p, 650.0))))))
        fdp[j] = (vals[0] - vals[1]) / (2 * h)
    p.eps[1] = B.epsdrude(*DRUDE)
    b64.enei = None
    err_p = float(np.max(np.abs(dpar - fdp) / np.abs(fdp)))
    # vertices: central differences along seeded directions, step 1e-4 nm per vertex
    rng = np.random.default_rng(2025)
    an, fd = [], []
    for _ in range(4):
        u = rng.normal(size=v.shape)
        u *= np.sqrt(len(v)) / np.linalg.norm(u)
        vals = []
        for sgn in (1, -1):
            q = p.moved(v + sgn * 1e-4 * u)
            vals.append(float(np.sum(exc.ext(B.bemsolver(q, q.op, precision="fp64").solve(exc(q, 650.0))))))
        fd.append((vals[0] - vals[1]) / 2e-4)
        an.append(float(np.sum(dv * u)))
    an, fd = np.array(an), np.array(fd)
    err_v = float(np.max(np.abs(an - fd)) / np.max(np.abs(an)))
    lams = np.linspace(500.0, 1000.0, nsweep)
    res, grads = {}, {}
    for prec, nl in (("mixed", nsweep), ("fp64", 4)):
        p, bem, exc = make(prec)
        gradient(p, bem, exc, lams[-1])  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        grads[prec] = [gradient(p, bem, exc, lam) for lam in lams[:nl]]
        torch.cuda.synchronize()
        res[prec] = nl / (time.perf_counter() - t)
        del bem
    err_mx = max(float(np.max(np.abs(a[1] - b[1])) / np.max(np.abs(b[1])))
                 for a, b in zip(grads["mixed"], grads["fp64"]))
    log(f"[bench] adjoint: vertex gradient vs FD rel err {err_v:.2e}, Drude vs FD {err_p:.2e}, mixed vs FP64 "
        f"{err_mx:.2e}; mixed {res['mixed']:.2f} λ/s, fp64 {res['fp64']:.2f} λ/s (extinction + full gradient)")
    return {"workload": f"d(sum of ext)/d(vertex coordinates, eps_inf, wp, gamma) by the adjoint, ellipsoid "
                        f"10x10x40 nm ({len(f)} faces, {len(v)} vertices) in eps=1.77, pols x(dir z) and z(dir x), "
                        "650 nm FD checks and linspace(500,1000,32) sweep",
            "wavelengths_per_s_mixed": res["mixed"], "wavelengths_per_s_fp64": res["fp64"],
            "vertex_gradient_max_rel_err_vs_central_fd_fp64": err_v,
            "drude_gradient_max_rel_err_vs_central_fd_fp64": err_p,
            "mixed_vs_fp64_vertex_gradient_max_rel_diff": err_mx,
            "ext_650nm": ext, "drude_gradient_650nm": dpar.tolist(),
            "vertex_gradient_norm_650nm": float(np.linalg.norm(dv))}


def run_multi_particle(per_particle=8, fp64_every=4):
    """The seeded many-particle sweep (12 Drude particles, 320-5,120 faces, in
    vacuum, 128 wavelengths 400-900 nm each): all 1,536 particle-wavelength
    solves through the batched path (mnpbem_amd.batch.spectra: per face count
    the batched pipeline, captured as a hipGraph, with the variant tuned
    beforehand on one wavelength per particle at the same batch sizes; the
    tuning time is reported apart) in mixed precision; every fp64_every-th
    wavelength in FP64; and per_particle
    wavelengths per particle through the dense solver, one solve at a time,
    as the reference's loop runs them (the comparison and the parity check)."""
    import torch

    from mnpbem_amd import bem as B
    from mnpbem_amd.batch import plan_capacity, spectra
    from mnpbem_amd.workloads import multi_particle_sweep

    parts, enei = multi_particle_sweep()
    exc = B.planewave([[1.0, 0, 0]], [[0, 0, 1.0]])
    ps = [B.comparticle([B.epsconst(1.0), B.epsdrude(*q["drude"])], [(q["verts"], q["faces"])], [2, 1], 1)
          for q in parts]
    t0 = time.perf_counter()
    for p in ps:
        p.green
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    out = {"workload": f"{len(parts)} seeded Drude ellipsoids (faces {sorted(p.n for p in ps)}), eps 1, pol x dir z, "
                       f"128 wavelengths 400-900 nm each",
           "faces_total": int(sum(p.n for p in ps)), "setup_s_per_particle": setup / len(ps)}
    full = {}
    for prec, lams in (("mixed", enei), ("fp64", enei[::fp64_every])):
        # one wavelength per particle at the timed run's batch sizes: tunes the
        # variant of each face count (the tuner keeps it for the process)
        caps = plan_capacity(ps, len(lams), prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        tune = {}
        spectra(ps, enei[:1], exc, precision=prec, capacity=caps, stats=tune)
        torch.cuda.synchronize()
        t_tune = time.perf_counter() - t
        stats = {}
        t = time.perf_counter()
        ext, sca, _ = spectra(ps, lams, exc, precision=prec, capacity=caps, stats=stats)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        full[prec] = (lams, ext[..., 0])
        out[f"batched_{prec}"] = {
            "particle_wavelengths": len(ps) * len(lams), "seconds": dt, "tuning_seconds": t_tune,
            "particle_wavelengths_per_s": len(ps) * len(lams) / dt,
            "per_face_count": {str(n): st for n, st in stats.items()},
            "tuner_ms_per_variant": {str(n): st.get("variant_ms") for n, st in tune.items()}}
        for st in stats.values():
            st.pop("variant_ms", None)
        log(f"[bench] multi-particle batched {prec}: {len(ps) * len(lams)} solves in {dt:.2f} s "
            f"({len(ps) * len(lams) / dt:.1f} particle-wavelengths/s); "
            + "; ".join(f"{n} faces: {'dense' if st.get('dense') else 'variant %s x%s' % (st.get('variant'), st.get('capacity'))}"
                        for n, st in stats.items()))
    # the dense solver one solve at a time, per_particle wavelengths per particle
    sub = np.arange(0, len(enei), len(enei) // per_particle)[:per_particle]
    seq = {}
    for prec in ("mixed", "fp64"):
        tot, ext, per_n = 0.0, [], {}
        for p in ps:
            bem = B.bemsolver(p, B.bemoptions(sim="ret"), precision=prec)
            torch.cuda.synchronize()
            t = time.perf_counter()
            ext.append([exc.ext(bem.solve(exc(p, lam)))[0] for lam in enei[sub]])
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            tot += dt
            per_n.setdefault(str(p.n), []).append(dt / len(sub))
            del bem
        out[f"s_per_solve_dense_{prec}"] = {n: float(np.mean(v)) for n, v in per_n.items()}
        seq[prec] = np.array(ext)
        out[f"particle_wavelengths_per_s_{prec}_dense_sequential"] = len(ps) * len(sub) / tot
        log(f"[bench] multi-particle dense one at a time {prec}: {len(ps) * len(sub) / tot:.1f} particle-wavelengths/s")
    out["speedup_batched_vs_dense_sequential_mixed"] = (out["batched_mixed"]["particle_wavelengths_per_s"] /
                                                        out["particle_wavelengths_per_s_mixed_dense_sequential"])
    out["speedup_batched_vs_dense_sequential_fp64"] = (out["batched_fp64"]["particle_wavelengths_per_s"] /
                                                       out["particle_wavelengths_per_s_fp64_dense_sequential"])
    # parity on the wavelengths both runs share
    ref = seq["fp64"]
    out["max_rel_err_batched_mixed_vs_dense_fp64"] = float(np.max(np.abs(full["mixed"][1][:, sub] / ref - 1)))
    assert np.all(sub % fp64_every == 0)  # the FP64 sweep holds every compared wavelength
    out["max_rel_err_batched_fp64_vs_dense_fp64"] = float(np.max(np.abs(full["fp64"][1][:, sub // fp64_every] / ref - 1)))
    out["max_rel_err_dense_mixed_vs_dense_fp64"] = float(np.max(np.abs(seq["mixed"] / ref - 1)))
    return out


def run_batch_sweep(precision="mixed"):
    """Only the batched part of run_multi_particle (tuning pass, then the
    1,536 solves of the seeded sweep) in one precision, for profiling."""
    import torch

    from mnpbem_amd import bem as B
    from mnpbem_amd.batch import plan_capacity, spectra
    from mnpbem_amd.workloads import multi_particle_sweep

    parts, enei = multi_particle_sweep()
    exc = B.planewave([[1.0, 0, 0]], [[0, 0, 1.0]])
    ps = [B.comparticle([B.epsconst(1.0), B.epsdrude(*q["drude"])], [(q["verts"], q["faces"])], [2, 1], 1)
          for q in parts]
    for p in ps:
        p.green
    caps = plan_capacity(ps, len(enei), precision)
    spectra(ps, enei[:1], exc, precision=precision, capacity=caps)
    torch.cuda.synchronize()
    stats = {}
    t = time.perf_counter()
    spectra(ps, enei, exc, precision=precision, capacity=caps, stats=stats)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    return {"precision": precision, "particle_wavelengths": len(ps) * len(enei), "seconds": dt,
            "particle_wavelengths_per_s": len(ps) * len(enei) / dt,
            "per_face_count": {str(n): {k: v for k, v in st.items() if k != "variant_ms"} for n, st in stats.items()}}


def run_hmatrix_dipole(nlam=16, seed=2024, coarse=2):
    """H-matrix iterative path with dipole excitation: sphere of diameter 150 nm
    (icosphere level 6, 81,920 faces), Drude gold in eps=1, dipole 5 nm above
    the surface along a seeded random direction with a seeded random moment,
    cleaf 32, eta 2.5, htol 1e-6, kmax [4, 100], GMRES tol 1e-6 with FP32
    H-matrix storage, 16 wavelengths from 500 to 700 nm; total and radiative
    decay rates."""
    import torch

    from mnpbem_amd import _lib, bem as B, shapes

    v, f = shapes.icosphere(6, 75.0)
    rng = np.random.default_rng(seed)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    dip = rng.normal(size=3)
    dip /= np.linalg.norm(dip)
    op = B.bemoptions(sim="ret", iter=dict(cleaf=32, htol=1e-6, kmax=[4, 100], tol=1e-6, hprecision="fp32",
                                           coarse=coarse))
    p = B.comparticle([B.epsconst(1.0), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, op)
    t0 = time.perf_counter()
    p.green
    it = B.bemsolver(p, op)
    exc = B.dipoleret(80.0 * u[None], dip[None])
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    lams = np.linspace(500.0, 700.0, nlam)
    _lib.call("mnp_timers_enable", 1)
    torch.cuda.synchronize()
    t = time.perf_counter()
    tot, rad, iters, flags = [], [], [], []
    for lam in lams:
        sig = it.solve(exc(p, lam))
        a, b, _ = exc.decayrate(sig)
        tot.append(float(a[0, 0]))
        rad.append(float(b[0, 0]))
        st = it.stats()
        iters.append(st["iter"])
        flags.append(st["flag"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    tm = _lib.Timers()
    _lib.call("mnp_timers_read", C.byref(tm))
    _lib.call("mnp_timers_enable", 0)
    hv = 4  # MNP_TIMER_HMATVEC
    mv_ms = tm.ms[hv] / max(tm.launches[hv], 1)
    mv_bytes = tm.work[hv] / max(tm.launches[hv], 1)
    gbs = mv_bytes / (mv_ms * 1e-3) / 1e9 if mv_ms > 0 else 0.0
    log(f"[bench] H-matrix dipole: {nlam} wavelengths in {el:.2f}s (setup {setup:.2f}s), iterations {iters}")
    if any(flags):
        raise RuntimeError(f"H-matrix dipole: GMRES did not converge at wavelengths "
                           f"{[float(l) for l, fl in zip(lams, flags) if fl]}")
    return {
        "workload": "sphere d=150 nm, icosphere level 6 (81920 faces), Drude gold in eps=1, dipole 5 nm "
                    "from the surface (seeded), bemretiter cleaf 32, htol 1e-6, GMRES tol 1e-6, FP32 "
                    "H-matrix storage, 16 wavelengths 500-700 nm, total and radiative decay rates",
        "wavelengths_per_s": nlam / el, "ms_per_wavelength": 1e3 * el / nlam, "setup_s": setup,
        "gmres_iterations": iters, "gmres_flags": flags, "compression": it.stats()["compression"],
        "tot": tot, "rad": rad,
        "hmatvec": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0,
                    "avg_launch_ms": mv_ms, "bytes_per_launch": mv_bytes,
                    "launches": int(tm.launches[hv]), "share_of_time": tm.ms[hv] * 1e-3 / el},
        "aca_share_of_time": tm.ms[5] * 1e-3 / el,
    }


def run_hmatrix_product(nprod=50):
    """The H-matrix product alone, for timing and hardware counters: G of the
    81,920-face sphere (icosphere level 6, d = 150 nm) at 600 nm, cleaf 32,
    htol 1e-6, FP32 storage, 4 right-hand sides, `nprod` products; beside it
    the device's plain streaming rates (torch sum over 2 GiB: the read
    calibration of FETCH_SIZE; fill of 2 GiB: the write calibration of
    WRITE_SIZE; copy of 2 GiB)."""
    import torch

    from mnpbem_amd import bem as B, shapes
    from mnpbem_amd import hmatrix as H

    def timed(fn, reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    x = torch.ones(1 << 28, dtype=torch.float64, device="cuda")  # 2 GiB
    y = torch.empty_like(x)
    gib2 = float(1 << 31)
    read_ms = timed(lambda: x.sum(), 10)
    fill_ms = timed(lambda: y.fill_(2.0), 3)
    copy_ms = timed(lambda: y.copy_(x), 10)
    del x, y
    v, f = shapes.icosphere(6, 75.0)
    p = B.comparticle([B.epsconst(1.0), B.epsdrude(*DRUDE)], [(v, f)], [2, 1], 1, B.bemoptions(sim="ret"))
    g = p.green
    tree = H.clustertree(g.geometry()["pos"], cleaf=32)
    hm = H.hmatrix(g, tree, htol=1e-6, kmax=100, precision="fp32").fill(2 * np.pi / 600.0, "G")
    st = hm.stats()
    nbytes = 8.0 * (st["full_entries"] + st["lowrank_entries"])
    xr = torch.randn(g.n, 4, dtype=torch.complex128, device="cuda")
    mv_ms = timed(lambda: hm.matvec(xr), nprod)
    log(f"[bench] H-matrix product: {mv_ms:.3f} ms for {nbytes / 1e9:.3f} GB")
    return {"workload": "G of the 81920-face sphere at 600 nm, cleaf 32, htol 1e-6, FP32 storage, 4 right-hand sides",
            "ms_per_product": mv_ms, "bytes_per_product": nbytes, "gb_per_s": nbytes / (mv_ms * 1e-3) / 1e9,
            "stream_read_gb_per_s": gib2 / (read_ms * 1e-3) / 1e9,
            "stream_fill_gb_per_s": gib2 / (fill_ms * 1e-3) / 1e9,
            "stream_copy_gb_per_s": 2 * gib2 / (copy_ms * 1e-3) / 1e9}


def cpu_baseline_hmatrix(iterations, compression, nsample=60, seed=7):
    """Oracle port of the H-matrix path on the host for a bounded sample,
    scaled to one wavelength of the 81,920-face workload: partially pivoted
    ACA (NumPy restatement) of `nsample` seeded admissible blocks of F and
    the products L (R^T x) and V x (4 right-hand sides, complex128) of the
    sampled low-rank blocks and as many full blocks, scaled by the totals to
    4 fills plus `iterations` x 4 products."""
    from mnpbem_amd import shapes
    from mnpbem_amd.hmatrix import admissibility, block_ranges, clustertree
    from oracle import hmat_oracle as HO
    from oracle import mnpbem_oracle as O
    v, f = shapes.icosphere(6, 75.0)
    pos, _, _, nvec, area = O.norm_flat(v, f)
    tree = clustertree(pos, 32)
    full, low = admissibility(tree)
    lr, fr = block_ranges(tree, low), block_ranges(tree, full)
    rng = np.random.default_rng(seed)
    k = 2 * np.pi / 600.0
    perm = tree.perm
    t_aca = t_mv = 0.0
    mn_s = ent_s = 0
    x = rng.normal(size=(len(f), 4)) + 1j * rng.normal(size=(len(f), 4))
    for b in rng.choice(len(low), nsample, replace=False):
        r0, r1, c0, c1 = lr[:, b]
        rows, cols = perm[r0:r1], perm[c0:c1]
        t = time.perf_counter()
        L, R = HO.aca(lambda r: HO.kernel_block(pos, nvec, area, rows[[r]], cols, k, "F")[0],
                      lambda c: HO.kernel_block(pos, nvec, area, rows, cols[[c]], k, "F")[:, 0],
                      len(rows), len(cols), tol=1e-6, kmax=100)
        t_aca += time.perf_counter() - t
        mn_s += len(rows) + len(cols)
        xs = x[cols]
        t = time.perf_counter()
        _ = L @ (R.T @ xs)
        t_mv += time.perf_counter() - t
        ent_s += L.size + R.size
    for b in rng.choice(len(full), nsample, replace=False):
        r0, r1, c0, c1 = fr[:, b]
        V = rng.normal(size=(r1 - r0, c1 - c0)) + 1j
        xs = x[perm[c0:c1]]
        t = time.perf_counter()
        _ = V @ xs
        t_mv += time.perf_counter() - t
        ent_s += V.size
    mn_all = float((lr[1] - lr[0] + lr[3] - lr[2]).sum())
    ent_all = compression * len(f) ** 2  # stored entries of one operator (compression of the GPU run)
    aca_all = t_aca / mn_s * mn_all
    mv_all = t_mv / ent_s * ent_all
    per_lam = 4 * aca_all + iterations * 4 * mv_all
    return {"value": 1.0 / per_lam, "unit": "wavelengths/s", "cores": 1, "kind": "port",
            "sample": f"ACA of {nsample} seeded admissible blocks ({t_aca:.2f} s) and NumPy products of "
                      f"{nsample} low-rank + {nsample} full blocks ({t_mv * 1e3:.1f} ms), scaled to 4 fills "
                      f"({4 * aca_all:.0f} s) + {iterations} x 4 products ({iterations * 4 * mv_all:.0f} s) "
                      "per wavelength"}


def cpu_baseline(level=4):
    """Oracle (NumPy/SciPy-OpenBLAS FP64) time for one wavelength at 5,120 faces."""
    from mnpbem_amd import shapes
    from oracle import mnpbem_oracle as O
    try:
        from threadpoolctl import threadpool_info
        cores = max((i.get("num_threads") or 1) for i in threadpool_info()) or os.cpu_count()
    except Exception:  # pragma: no cover
        cores = os.cpu_count()
    v, f = shapes.ellipsoid(level, (10.0, 10.0, 40.0))
    g = O.GreenRet(v, f)
    pv, pf = shapes.trisphere256(2.0)
    _, _, _, pn, pa = O.norm_flat(pv, pf)
    pol, dirs = np.array(POL), np.array(DIR)
    lam = 650.0
    # the oracle's initmat is assembly + chain; time the assembly alone first
    k0 = 2 * np.pi / lam
    t = time.perf_counter()
    g.eval(k0 * np.sqrt(complex(O.eps_drude(lam, *DRUDE))), ("G", "H1"))
    g.eval(k0 * np.sqrt(EPS_OUT), ("G", "H2"))
    t_asm = time.perf_counter() - t
    t = time.perf_counter()
    bem = O.BemRet(g).initmat(lam, O.eps_drude(lam, *DRUDE), EPS_OUT)
    t_init = time.perf_counter() - t
    t = time.perf_counter()
    a2, a2p = O.planewave_potential(g.pos, g.nvec, pol, dirs, lam, EPS_OUT)
    sig = bem.solve({"a2": a2, "a2p": a2p})
    O.extinction(g.pos, g.area, sig, pol, dirs, lam, EPS_OUT)
    O.scattering(g.pos, g.area, sig, lam, EPS_OUT, pn, pa)
    t_sol = time.perf_counter() - t
    t_chain = max(t_init - t_asm, 0.0)
    r = 20480 / len(f)
    # O(N^2) parts (assembly, solve, cross sections) scale by r^2, the chain by r^3
    est = (t_asm + t_sol) * r ** 2 + t_chain * r ** 3
    return {"value": 1.0 / est, "unit": "wavelengths/s", "cores": int(cores),
            "kind": "port",
            "sample": f"one wavelength (650 nm, 2 pols) of the same ellipsoid at {len(f)} faces: assembly "
                      f"{t_asm:.2f} s, chain (4 inverses + 2 products) {t_chain:.2f} s, solve + cross "
                      f"sections {t_sol:.2f} s; scaled to 20480 faces by {r:.0f}^2 (assembly, solve) and "
                      f"{r:.0f}^3 (chain): {est:.0f} s per wavelength"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=4)
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-hmatrix", action="store_true", help="skip the H-matrix dipole measurement")
    ap.add_argument("--no-split-bf16", action="store_true", help="skip the split-bf16 chain measurement")
    ap.add_argument("--no-quasistatic", action="store_true", help="skip the quasistatic measurement")
    ap.add_argument("--no-multi-particle", action="store_true", help="skip the many-particle sweep measurement")
    ap.add_argument("--no-mirror", action="store_true", help="skip the mirror-symmetry measurement")
    ap.add_argument("--no-substrate", action="store_true", help="skip the substrate measurement")
    ap.add_argument("--no-sensitivity", action="store_true", help="skip the Drude-gradient measurement")
    ap.add_argument("--hmatrix-only", action="store_true",
                    help="only the H-matrix dipole measurement (for profiling)")
    ap.add_argument("--hmatrix-coarse", type=int, default=2,
                    help="row tiles per coarse aggregate of the H-matrix preconditioner")
    ap.add_argument("--only", default=None,
                    help="comma list of side measurements to run alone (for profiling): "
                         "quasistatic, mirror, substrate, substrate_ret, multi_particle, hmatrix_dipole, sensitivity, batch_sweep")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy three times (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; wavelengths shard across GPUs without "
              "communication, but multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        import torch
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copy calibration: 3 copies of 1 GiB")
        del src, dst
    if a.hmatrix_only:
        print(json.dumps({"hmatrix_dipole": run_hmatrix_dipole(coarse=a.hmatrix_coarse)}), flush=True)
        return
    if a.only:
        runs = {"quasistatic": lambda: run_quasistatic(a.level), "mirror": run_mirror,
                "substrate": lambda: run_substrate(a.level), "substrate_ret": run_substrate_ret,
                "substrate_iter": run_substrate_iter,
                "substrate_dipole": run_substrate_dipole,
                "multi_particle": run_multi_particle,
                "sensitivity": run_sensitivity, "batch_sweep": run_batch_sweep,
                "hmatrix_dipole": run_hmatrix_dipole, "hmatrix_product": run_hmatrix_product}
        names = [x.strip() for x in a.only.split(",") if x.strip()]
        bad = [x for x in names if x not in runs]
        if bad:
            ap.error(f"--only: unknown measurement(s) {bad}")
        print(json.dumps({x: runs[x]() for x in names}), flush=True)
        return
    r = run_gpu(a.level, a.steps, a.warmup, a.fp64_steps, not a.no_split_bf16)
    tm = r["timers"]
    gi = 1  # MNP_TIMER_GEMM_FP32
    gemm_ms = tm.ms[gi] / max(tm.launches[gi], 1)
    gemm_flops = tm.work[gi] / max(tm.launches[gi], 1)
    achieved = gemm_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    n = r["n"]
    chain_flops = 48.0 * n ** 3
    step_s = r["t_mx"] / a.steps
    out = {
        "metric": METRIC,
        "value": a.steps / r["t_mx"],
        "unit": "wavelengths/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * step_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+c128",
        "data": "synthetic",
        "config": {
            "workload": f"prolate ellipsoid 10x10x40 nm, icosphere level {a.level} "
                        f"({n} faces), Drude gold in eps=1.77, pols x(dir z) and z(dir x), "
                        "wavelengths from linspace(500,1000,64)",
            "faces": n, "npol": 2, "precision": "complex FP32 chain + FP64 refinement",
        },
        "roofline": {
            "bound": "mfma", "kernel": "k_cgemm_f32lds<4,3> (3M complex GEMM on FP32 MFMA, direct-to-LDS operands)",
            "achieved": achieved, "peak": PEAK["fp32"], "unit": "TFLOP/s",
            "frac": achieved / PEAK["fp32"], "traffic": None,
            "launches_per_step": tm.launches[gi] / a.steps,
            "avg_launch_ms": gemm_ms, "avg_flops_per_launch": gemm_flops,
            "avg_algorithmic_bytes_per_launch": tm.bytes[gi] / max(tm.launches[gi], 1),
        },
        "fp64_wavelengths_per_s": r["n64"] / r["t_64"] if r["n64"] else None,
        "speedup_vs_fp64": (a.steps / r["t_mx"]) / (r["n64"] / r["t_64"]) if r["n64"] else None,
        "max_rel_err_vs_fp64": r["max_err"],
        "fp64_wavelengths": r["lam64"],
        "timed_wavelengths": r["lam_t"],
        "refinement": {"sweeps": r["sweeps"][0], "relres": r["sweeps"][1]},
        "chain_effective_tflops": chain_flops / step_s / 1e12,
        "gemm_share_of_step": tm.ms[gi] * 1e-3 / r["t_mx"],
        "panel_share_of_step": tm.ms[2] * 1e-3 / r["t_mx"],
        "assembly_share_of_step": tm.ms[3] * 1e-3 / r["t_mx"],
        "fp64_residual_gemm_share_of_step": tm.ms[0] * 1e-3 / r["t_mx"],
    }
    t64 = r["timers64"]
    if t64 is not None and t64.launches[0] and t64.ms[0] > 0:
        a64 = t64.work[0] / (t64.ms[0] * 1e-3) / 1e12
        out["fp64_gemm"] = {"achieved": a64, "peak": PEAK["fp64"], "frac": a64 / PEAK["fp64"]}
    tb = r["timers_bf"]
    if tb is not None:
        si = 6  # MNP_TIMER_GEMM_SBF16
        sb_ms = tb.ms[si] / max(tb.launches[si], 1)
        sb_flops = tb.work[si] / max(tb.launches[si], 1)
        sb_ach = sb_flops / (sb_ms * 1e-3) / 1e12 if sb_ms > 0 else 0.0
        v_bf = a.steps / r["t_bf"]
        out["split_bf16"] = {
            "precision": "complex FP32 chain, large products from hi/lo bf16 splits on bf16 MFMA, "
                         "+ FP64 refinement",
            "wavelengths_per_s": v_bf, "ms_per_step": 1e3 * r["t_bf"] / a.steps,
            "speedup_vs_fp64": v_bf / (r["n64"] / r["t_64"]) if r["n64"] else None,
            "speedup_vs_fp32_mfma": v_bf / out["value"],
            "max_rel_err_vs_fp64": r["max_err_bf"],
            "refinement": {"sweeps": r["sweeps_bf"][0], "relres": r["sweeps_bf"][1]},
            "roofline": {
                "bound": "mfma", "kernel": "k_cgemm_sbf16<4,4,2> (4M complex GEMM, 3 bf16 products per real one, 128 x 128 tiles)",
                "achieved": sb_ach, "peak": PEAK["bf16"], "unit": "TFLOP/s", "frac": sb_ach / PEAK["bf16"],
                "executed_bf16_tflops": 3.0 * sb_ach, "executed_frac": 3.0 * sb_ach / PEAK["bf16"],
                "launches_per_step": tb.launches[si] / a.steps, "avg_launch_ms": sb_ms,
                "avg_flops_per_launch": sb_flops,
            },
            "gemm_share_of_step": tb.ms[si] * 1e-3 / r["t_bf"],
        }
    if not a.no_quasistatic:
        out["quasistatic"] = run_quasistatic(a.level)
    if not a.no_mirror:
        out["mirror"] = run_mirror()
    if not a.no_substrate:
        out["substrate"] = run_substrate(a.level)
        out["substrate_ret"] = run_substrate_ret()
    if not a.no_sensitivity:
        out["sensitivity"] = run_sensitivity()
    if not a.no_multi_particle:
        out["multi_particle"] = run_multi_particle()
    if not a.no_hmatrix:
        try:
            out["hmatrix_dipole"] = run_hmatrix_dipole()
        except RuntimeError as e:  # a wavelength that did not converge voids the measurement
            out["hmatrix_dipole"] = {"error": str(e)}
        if not a.no_cpu_baseline and "error" not in out["hmatrix_dipole"]:
            it = out["hmatrix_dipole"]["gmres_iterations"]
            out["hmatrix_dipole"]["cpu_baseline"] = cpu_baseline_hmatrix(int(round(sum(it) / len(it))),
                                                                         out["hmatrix_dipole"]["compression"])
    if not a.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
