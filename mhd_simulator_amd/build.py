"""Build libmhdp.so (HIP kernels for gfx950 + the C ABI) in-tree with hipcc.

    python -m mhd_simulator_amd.build

Every source is compiled with packed-FP32 instructions disabled; the setup kernels
(face mobilities, RHS, Galerkin graphs, velocity correction, display fields) and the
momentum step (predict, viscous) are compiled without floating-point contraction so they
reproduce the reference's FP32 formulas exactly.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libmhdp.so")
ARCH = os.environ.get("MHDP_ARCH", "gfx950")

SOURCES = {
    "setup.hip": ["-ffp-contract=off"],
    "pcg.hip": [],
    "mg.hip": [],
    "fused.hip": [],
    "brick.hip": [],
    "march.hip": [],
    "rbgs.hip": [],
    "cc.hip": [],
    "step.hip": ["-ffp-contract=off"],
    "em.hip": ["-ffp-contract=off"],
    "api.cpp": ["-x", "hip", "-pthread"],
    "em_api.cpp": ["-x", "hip"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain (ROCm) is required to build libmhdp.so")


def common_flags() -> list[str]:
    return ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            f"-I{INCLUDE}", f"-I{CSRC}"]


def _compile(src: str, extra: list[str]) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *common_flags(), *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
