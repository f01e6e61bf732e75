"""Build libvct.so (HIP, gfx950) and the mvct drop-in drivers in-tree.

    python -m virtualct_amd._build          # or __graft_entry__.build()

Every call recompiles.  Device code is built with contraction off (the tracer
and backprojector reproduce the reference's FP64 operation sequence) and with
packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
OBJDIR = os.path.join(PKG, "_build")

ARCH = os.environ.get("VCT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

LIB_SOURCES = ["vct_api.cpp", "vct_trace.hip", "vct_fbp.hip", "vct_batch.hip", "vct_npf.hip", "vct_measured.hip"]
DRIVERS = {"mvct_mbfpa": "mvct_mbfpa.cpp", "mvct_fbp": "mvct_fbp.cpp", "mvct_bhc": "mvct_bhc.cpp"}

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False, jobs: int = 8) -> str:
    """Compile libvct.so and the drivers; returns the library path."""
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd)
        if verbose:
            print("compiled", src, flush=True)
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, srcs))
    lib = os.path.join(LIBDIR, "libvct.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    for name, src in DRIVERS.items():
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        exe = os.path.join(BINDIR, name)
        _run([HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", path, "-o", exe + ".tmp",
              f"-L{LIBDIR}", "-lvct", "-Wl,-rpath,$ORIGIN/../lib"])
        os.replace(exe + ".tmp", exe)
        if verbose:
            print("linked", name, flush=True)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
