"""Builds libmnpbem_hip.so (all HIP sources in csrc/) for gfx950 with hipcc.

    python -m mnpbem_amd.build        # or mnpbem_amd.build.build()

Each source compiles to its own object (in parallel, MAX_JOBS at a time,
default the CPU count) under mnpbem_amd/_build/, then one link writes the
library next to this file so that it travels with the source tree; the
Python layer loads it from there.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libmnpbem_hip.so")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("MNP_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    f"--offload-arch={ARCH}",
    # no packed-FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # keep MFMA sources and accumulators in ArchVGPRs: with the default
    # heuristic the GEMM tiles shuttle accumulators between AGPRs and VGPRs
    # around each MFMA (v_accvgpr_read/write), serialising the MFMA pipe
    "-mllvm", "--amdgpu-mfma-vgpr-form",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    return res.returncode, err


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    inc = ["-I", os.path.join(HERE, "..", "include")]
    srcs = sources()
    objs = [os.path.join(OBJ, os.path.splitext(os.path.basename(s))[0] + ".o") for s in srcs]
    cmds = [[hipcc, *FLAGS, *inc, "-c", s, "-o", o] for s, o in zip(srcs, objs)]
    jobs = max(1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=min(jobs, len(cmds))) as ex:
        results = list(ex.map(_run, cmds))
    for cmd, (rc, err) in zip(cmds, results):
        if verbose:
            print(" ".join(cmd))
        if rc != 0:
            raise RuntimeError(f"hipcc failed ({rc}) on {cmd[-3]}:\n{err}")
        if verbose and err.strip():
            print(err)
    tmp = LIB + ".tmp"
    rc, err = _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    if rc != 0:
        raise RuntimeError(f"hipcc link failed ({rc}):\n{err}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
