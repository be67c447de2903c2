"""Build recipe of libtomo_hip.so (gfx950) and the C++ driver executables.

    python tomographic-image-reconstruction-algorithms_amd/build.py [-v]

Outputs go to ``_build/`` next to this file (in-tree, so they travel with the
source tree).  Always recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
DRIVERS = os.path.join(PKG_DIR, "drivers")
OUT = os.path.join(PKG_DIR, "_build")
INCLUDE = os.path.join(ROOT, "include")

ARCH = os.environ.get("TOMO_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 (v_pk_*_f32)
# instructions in device code.  -ffp-contract=off: no automatic FMA
# contraction; the kernels write their fused multiply-adds explicitly, so the
# rounding of a cell does not depend on how a kernel variant was scheduled.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-I", INCLUDE, "-I", CSRC]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-I", INCLUDE, "-I", CSRC]

HIP_SOURCES = ["tomo_kernels.hip", "tomo_engine.hip", "tomo_cnn.hip", "tomo_batch.hip", "tomo_noalloc.hip"]
HOST_SOURCES = ["tomo_host.cpp"]
LIB = os.path.join(OUT, "libtomo_hip.so")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_library(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OUT, src.replace(".hip", ".o"))
        jobs.append(([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], obj))
    for src in HOST_SOURCES:
        obj = os.path.join(OUT, src.replace(".cpp", ".o"))
        jobs.append(([CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], obj))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda j: _run(j[0], verbose), jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *[o for _, o in jobs]], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_drivers(verbose: bool = False) -> list:
    """C++ command-line drivers mirroring the reference executables."""
    if not os.path.isdir(DRIVERS):
        return []
    outs = []
    srcs = sorted(s for s in os.listdir(DRIVERS) if s.endswith(".cpp") and s != "driver_common.cpp")
    common = os.path.join(DRIVERS, "driver_common.cpp")
    extra = [common] if os.path.exists(common) else []

    def one(src):
        exe = os.path.join(OUT, src[:-4].replace("_", "-"))
        _run([CXX, "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-I", DRIVERS, os.path.join(DRIVERS, src), *extra,
              "-o", exe, "-L", OUT, "-ltomo_hip", f"-Wl,-rpath,$ORIGIN"], verbose)
        return exe

    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        outs = list(ex.map(one, srcs))
    return outs


def build(verbose: bool = False) -> str:
    lib = build_library(verbose)
    build_drivers(verbose)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
