"""Build libbie_amd.so (HIP for gfx950) in-tree with hipcc.

The library is plain HIP/C++ with a C ABI (include/bie_amd.h); it links no PyTorch.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
CSRC = os.path.join(PKG, "csrc")
LIB_PATH = os.path.join(HERE, "libbie_amd.so")

SOURCES = ["rules.cpp", "schedule.cpp", "assemble.hip", "sweep.hip", "adjoint.hip", "apply.hip",
           "solve.hip", "sparse.hip", "near_field.cpp", "api.hip"]
# rocSOLVER (LU) and rocBLAS (GEMV) from the ROCm install, found at run time by rpath
ROCM_LIB = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
LINK_LIBS = ["-L" + ROCM_LIB, "-Wl,-rpath," + ROCM_LIB, "-lrocsolver", "-lrocblas"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every source for gfx950 and link libbie_amd.so beside this file."""
    cc = hipcc()
    objs = []
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    cmds = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmds.append([cc] + HIP_FLAGS + ["-x", "hip", "-c", path, "-o", obj])
        objs.append(obj)
    # the translation units are independent: compile them side by side
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), file=sys.stderr)
        for f in [pool.submit(subprocess.run, cmd, check=True) for cmd in cmds]:
            f.result()
    tmp = LIB_PATH + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + LINK_LIBS
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
