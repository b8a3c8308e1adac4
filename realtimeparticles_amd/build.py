"""Builds librtp.so (HIP kernels for gfx950 + the C ABI) in-tree with hipcc.

    python -m realtimeparticles_amd.build

Floating-point flags matter for parity with the CPU oracle: no contraction into FMA,
correctly rounded fp32 division and square root, IEEE denormals.  Packed-fp32 instruction
selection is disabled.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OUT_DIR = os.path.join(PKG, "lib")
LIB = os.path.join(OUT_DIR, "librtp.so")
OBJ_DIR = os.path.join(OUT_DIR, "obj")

HIP_SOURCES = ["grid_sort.hip", "pbf.hip", "boids.hip", "mlp.hip", "capi.hip", "batch.hip",
               "selfcheck.hip", "render.hip", "clouds.hip"]
HOST_SOURCES = ["host/Model.cpp", "host/json_params.cpp"]

ARCH = os.environ.get("RTP_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def build(verbose: bool = False, jobs: int = 4, extra_flags=()) -> str:
    """Compile every source and link librtp.so; extra_flags go to every hipcc compile
    (e.g. -DRTP_QUEUE_DEPTH=16 to size the neighbour passes' survivor queue)."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    cc = hipcc()
    inc = ["-I", INCLUDE, "-I", CSRC]
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
        cmds.append([cc, *HIPCC_FLAGS, *extra_flags, *inc, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
        cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", *inc,
                     "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs,
          "-lrocprofiler-sdk-roctx"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, extra_flags=[a for a in sys.argv[1:] if a.startswith("-D")]))
