"""Build libsophus_gpu.so (gfx950 HIP kernels + C ABI) in-tree with hipcc.

Usage: python sophus-rs_amd/build.py   (or call build_library()).
The shared library is written next to this file so that it travels with the tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libsophus_gpu.so"

HIP_SOURCES = ["kernels.hip", "pcg.hip", "learned.hip", "batch.hip", "noalloc.hip", "symbolic.hip"]
HOST_SOURCES = ["problem.cpp", "optimizer.cpp", "stage_api.cpp", "generator.cpp", "train.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
# gfx950 device code; packed FP32 instructions stay disabled (hipcc emits them from plain
# float code otherwise).
DEVICE = [
    "--offload-arch=gfx950",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:4])} ...")


def build_library(verbose: bool = False) -> Path:
    """Compile every source and link libsophus_gpu.so; returns its path."""
    hipcc = _hipcc()
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        obj = objdir / (src + ".o")
        jobs.append([hipcc, "-x", "hip", *COMMON, *DEVICE, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = objdir / (src + ".o")
        jobs.append([hipcc, *COMMON, "--offload-arch=gfx950", "-D__HIP_PLATFORM_AMD__", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    procs = [(cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
             for cmd in jobs]
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(out)
        elif verbose and out:
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("compilation failed")
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "--offload-arch=gfx950", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
