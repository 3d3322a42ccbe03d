"""Build libmolmodel_amd.so (HIP kernels for gfx950 + the C ABI) in-tree.

    python -m molmodel_amd.build        # or molmodel_amd.build.build()

Every HIP translation unit is compiled with packed FP32 instructions
disabled (-packed-fp32-ops); the neighbor-list unit additionally with
-ffp-contract=off so its FP64 cell assignment and distance tests round like
the reference's x86-64 build.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmolmodel_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["engine.cpp", "capi.cpp", "refapi.cpp", "nbrlist.hip", "rspace.hip", "kspace.hip",
           "pppm.hip", "finish.hip", "remd.hip", "tune.cpp", "adjoint.cpp", "adjoint.hip",
           "noalloc.hip", "sites.cpp", "msimio.cpp", "shape.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-x", "hip",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument", "-I", os.path.join(HERE, "..", "include")]
EXTRA = {"nbrlist.hip": ["-ffp-contract=off"]}


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + COMMON + EXTRA.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile of {src} failed:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
