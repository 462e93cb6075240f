"""Build recipe of the native libraries (run from __graft_entry__.build()).

  libpbig_hip.so  HIP kernels for gfx950 + the C-ABI of include/pbig_hip.h
  libpbig.so      the reference-compatible C++ API (include/pbig.h,
                  report.h, rect.h) on top of the C-ABI, float 2-D records
  libpbig_u32.so  the same API for the reference's shipped build: unsigned
                  3-D records (rect.h with PBIG_COMPAT_U32)

Both are written in-tree next to this file so that they travel with the
source tree.  hipcc cross-compiles for gfx950 without a GPU.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Packed FP32 instructions are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # IEEE float32 semantics: no denormal flushing, no contraction.
    "-fno-gpu-flush-denormals-to-zero",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wall",
]

HIP_SOURCES = ["pbig_kernels.hip", "pbig_binned.hip", "pbig_u32.hip", "pbig_capi.hip"]
HOST_SOURCES = ["pbig_rice.cpp"]  # host-only parts of libpbig_hip.so
LIB_HIP = os.path.join(HERE, "libpbig_hip.so")
LIB_API = os.path.join(HERE, "libpbig.so")
LIB_API_U32 = os.path.join(HERE, "libpbig_u32.so")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> None:
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(CSRC, src.replace(".cpp", ".o"))
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB_HIP + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB_HIP)
    tmp = LIB_API + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{INCLUDE}",
          os.path.join(CSRC, "pbig_api.cpp"), f"-L{HERE}", "-lpbig_hip",
          "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, LIB_API)
    tmp = LIB_API_U32 + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-DPBIG_COMPAT_U32", "-DDIMS=3",
          f"-I{INCLUDE}", os.path.join(CSRC, "pbig_api.cpp"), f"-L{HERE}", "-lpbig_hip",
          "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, LIB_API_U32)
    for o in objs:
        os.remove(o)


if __name__ == "__main__":
    build()
    sys.exit(0)
