"""Build libastrea_hip.so in-tree with hipcc for gfx950.

Run as ``python -m astrea_amd._build`` or through __graft_entry__.build().
Translation units compile in parallel; FP64 units use -ffp-contract=off so
their arithmetic rounds exactly as the numpy reference does, FP32 units (the
split-precision path) allow contraction.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libastrea_hip.so")
OBJ = os.path.join(HERE, "_obj")
ARCH = os.environ.get("AH_OFFLOAD_ARCH", "gfx950")
BASE = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
UNITS = {
    "ah_api.hip": ["-ffp-contract=off"],
    "ah_fp64_1d.hip": ["-ffp-contract=off"],
    "ah_fp64_2d.hip": ["-ffp-contract=off"],
    "ah_fp64_2d_fused.hip": ["-ffp-contract=off"],
    "ah_fp64_2d_author.hip": ["-ffp-contract=off"],
    "ah_fp32_1d.hip": [],
    "ah_fp32_2d.hip": [],
    "ah_fp32_2d_fused.hip": [],
}


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


def _compile(unit):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(OBJ, unit.replace(".hip", ".o"))
    cmd = [_hipcc(), *BASE, *UNITS[unit], "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {unit}:\n{r.stderr[-4000:]}")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(UNITS), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, sorted(UNITS)))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
