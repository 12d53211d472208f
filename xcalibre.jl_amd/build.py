"""Build the in-tree native libraries of the package.

* ``libxcal_mesh.so`` -- host mesh builders and VTK result writers (g++).
* ``libxcalhip.so``   -- the gfx950 HIP library behind include/xcal.h (hipcc).

All HIP code is compiled for gfx950 with FMA contraction off (the assembly
kernels reproduce the reference's rounding order) and without packed FP32.
Objects go to ``_build/`` at the repository root; the libraries land next to
this file.  Run ``python xcalibre.jl_amd/build.py`` or call :func:`build`.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INC = ROOT / "include"
BUILD = ROOT / "_build"
ARCH = os.environ.get("XCAL_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{INC}", f"-I{CSRC}",
]
HIP_SOURCES = ["core.hip", "assemble.hip", "calculate.hip", "krylov.hip", "solver.hip", "batch.hip",
               "turbulence.hip", "simple.hip", "serve.hip", "noalloc.hip", "ingest.hip"]
HEADERS = ["xcal_internal.h", "device_common.h"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_mesh() -> Path:
    out = PKG / "libxcal_mesh.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c++17", f"-I{INC}",
          str(CSRC / "mesh" / "meshgen.cpp"), str(CSRC / "io" / "writers.cpp"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_hip(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []

    def compile_one(src):
        obj = BUILD / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    out = PKG / "libxcalhip.so"
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp),
          "-lpthread"])
    os.replace(tmp, out)
    return out


def build(jobs: int = 8) -> list:
    return [build_mesh(), build_hip(jobs)]


if __name__ == "__main__":
    for p in build(int(sys.argv[1]) if len(sys.argv) > 1 else 8):
        print(p)
