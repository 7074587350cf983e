"""Benchmark: wavelengths/sec of the 20,480-face retarded BEM extinction spectrum.

Workload (one GPU): prolate ellipsoid with semi-axes 10 x 10 x 40 nm from an
icosphere subdivided 5 times (20,480 flat faces), Drude gold
eps = 10 - 9.065^2 / (w (w + 0.0708 i)) in a medium of eps = 1.77, plane
waves polarised x (propagating z) and z (propagating x), wavelengths from
500 to 1000 nm.  One step = one wavelength: assembly of G1, G2, H1, H2, the
bemret chain, the solve for both polarisations and the FP64 extinction and
scattering cross sections, with every input resident on the GPU.

`value` is the mixed-precision path (complex FP32 chain on MFMA used as the
preconditioner of the exact FP64 operator, refined to a relative residual of
1e-12).  The FP64 drop-in is timed on `--fp64-steps` wavelengths for the
speed-up and the cross-section error of the mixed path is reported against
it.  `cpu_baseline` times the NumPy/OpenBLAS FP64 oracle on the host for one
wavelength of the same shape at 5,120 faces and scales it by N^3 to 20,480.

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
PEAK = {"fp32": 157.3, "fp64": 78.6}  # dense MFMA TFLOP/s, MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(level, steps, warmup, fp64_steps):
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
    nlam = warmup + steps
    enei = np.linspace(500.0, 1000.0, 64)[:max(nlam, fp64_steps)]

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
        spectrum(mixed, enei[:warmup], False)
        log(f"[bench] mixed warmup done ({warmup})")
    res_mx, t_mx, tm = spectrum(mixed, enei[warmup:warmup + steps], True)
    sweeps = mixed.stats()
    log(f"[bench] mixed {steps} wavelengths in {t_mx:.2f}s, refinement {sweeps}")
    del mixed
    torch.cuda.empty_cache()

    fp64 = B.bemsolver(p, op, precision="fp64")
    lam64 = enei[warmup:warmup + fp64_steps]
    res_64, t_64, tm64 = spectrum(fp64, lam64, True)
    log(f"[bench] fp64 {len(lam64)} wavelengths in {t_64:.2f}s")
    errs = []
    for a, b in zip(res_mx, res_64):
        for q in range(2):
            errs.append(abs(a[q] - b[q]) / np.abs(b[q]))
    max_err = float(np.max(errs)) if errs else None
    return dict(n=n, t_mx=t_mx, t_64=t_64, n64=len(lam64), timers=tm, timers64=tm64,
                max_err=max_err, sweeps=sweeps, ext=[list(map(float, r[0])) for r in res_mx])


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
    t = time.perf_counter()
    bem = O.BemRet(g).initmat(lam, O.eps_drude(lam, *DRUDE), EPS_OUT)
    a2, a2p = O.planewave_potential(g.pos, g.nvec, pol, dirs, lam, EPS_OUT)
    sig = bem.solve({"a2": a2, "a2p": a2p})
    O.extinction(g.pos, g.area, sig, pol, dirs, lam, EPS_OUT)
    O.scattering(g.pos, g.area, sig, lam, EPS_OUT, pn, pa)
    el = time.perf_counter() - t
    scale = (20480 / len(f)) ** 3
    return {"value": 1.0 / (el * scale), "unit": "wavelengths/s", "cores": int(cores),
            "kind": "port",
            "sample": f"one wavelength (650 nm, 2 pols) of the same ellipsoid at {len(f)} faces "
                      f"took {el:.2f} s; scaled by (20480/{len(f)})^3 = {scale:.0f}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; wavelengths shard across GPUs without "
              "communication, but multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)
    r = run_gpu(a.level, a.steps, a.warmup, a.fp64_steps)
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
            "bound": "mfma", "kernel": "k_cgemm<float> (complex GEMM, 4 real MFMA products)",
            "achieved": achieved, "peak": PEAK["fp32"], "unit": "TFLOP/s",
            "frac": achieved / PEAK["fp32"], "traffic": None,
            "launches_per_step": tm.launches[gi] / a.steps,
            "avg_launch_ms": gemm_ms, "avg_flops_per_launch": gemm_flops,
        },
        "fp64_wavelengths_per_s": r["n64"] / r["t_64"] if r["t_64"] else None,
        "speedup_vs_fp64": (a.steps / r["t_mx"]) / (r["n64"] / r["t_64"]) if r["t_64"] else None,
        "max_rel_err_vs_fp64": r["max_err"],
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
    if not a.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
