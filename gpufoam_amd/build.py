"""Builds libgpufoam_amd.so in-tree with hipcc for gfx950.

Every call compiles from source (no JIT, nothing under ~/.cache): the built
library sits next to this file and travels with the source tree.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgpufoam_amd.so")
OBJ = os.path.join(HERE, "_obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]
# Device code: gfx950 only; packed-FP32 instructions disabled; a*b+c is fused
# only inside one source expression, so inlined copies of the same device
# function round identically in every kernel that uses it.
DEVICE = [f"--offload-arch={ARCH}", "-ffp-contract=on",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


# The debug bounds-check variant (GFC_BOUNDS_CHECK): same sources, its own
# objects and library.
OUT_BOUNDS = os.path.join(HERE, "libgpufoam_amd_bounds.so")
OBJ_BOUNDS = os.path.join(HERE, "_obj_bounds")


def _compile(job) -> str:
    src, bounds = job
    odir = OBJ_BOUNDS if bounds else OBJ
    os.makedirs(odir, exist_ok=True)
    obj = os.path.join(odir, os.path.basename(src) + ".o")
    extra = ["-DGFC_BOUNDS_CHECK"] if bounds else []
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, *DEVICE, *extra, "-c", src, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, *extra, "-x", "c++", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def _link(objs, out: str) -> None:
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)


def build(verbose: bool = False, bounds_check: bool = True) -> str:
    """Builds libgpufoam_amd.so and, with bounds_check, the bounds-check
    variant libgpufoam_amd_bounds.so (both compiled in one job pool)."""
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.cpp")) +
                  glob.glob(os.path.join(CSRC, "*.hip")))
    # the slowest (largest) sources first, so the pool stays busy
    srcs.sort(key=lambda f: -os.path.getsize(f))
    jobs = [(f, False) for f in srcs] + ([(f, True) for f in srcs] if bounds_check else [])
    with ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(_compile, jobs))
    _link(objs[:len(srcs)], OUT)
    if verbose:
        print(f"built {OUT}", file=sys.stderr)
    if bounds_check:
        _link(objs[len(srcs):], OUT_BOUNDS)
        if verbose:
            print(f"built {OUT_BOUNDS}", file=sys.stderr)
    return OUT


if __name__ == "__main__":
    build(verbose=True, bounds_check="--no-bounds-check" not in sys.argv)
