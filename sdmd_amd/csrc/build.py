"""Build sdmd_amd/libsdmd.so: the C host (gcc) + the gfx950 HIP engine (hipcc).

Both halves are compiled without floating-point contraction
(-ffp-contract=off) so every product and sum rounds separately, as in the
x86-64 build of the reference; packed FP32 instructions are disabled for the
device code.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(PKG, "libsdmd.so")
ARCH = os.environ.get("SDMD_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["host/sdmd_host.c", "host/sdmd_remd.c", "host/sdmd_io.c"]
HIP_SOURCES = ["sdmd_gpu.hip", "sdmd_fast.hip"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        _run(["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
              "-Wall", "-Wno-format-truncation", "-I" + INC, "-c", os.path.join(HERE, src), "-o", obj])
        objs.append(obj)
    hipcc = _hipcc()
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        _run([hipcc, "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
              "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-I" + INC, "-c", os.path.join(HERE, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, LIB)
    # the C caller of the fine-grained entry points (tests/c_abi_loop.c),
    # linked against the library with only include/*.h
    drv = os.path.join(ROOT, "tests", "c_abi_loop.c")
    if os.path.exists(drv):
        _run(["gcc", "-O2", "-std=c99", "-Wall", "-I" + INC, drv, "-o", os.path.join(ROOT, "tests", "c_abi_loop.bin"),
              "-L" + PKG, "-lsdmd", "-Wl,-rpath," + PKG])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
