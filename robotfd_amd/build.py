"""Builds librfd.so (gfx950 HIP kernels + C ABI + C++ drop-in API) in-tree.

    python -m robotfd_amd.build

hipcc compiles for gfx950 only.  Every HIP translation unit is built with
-ffp-contract=off (only the explicit FMAs of the per-cell formulas fuse) and
with packed FP32 ops disabled.
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
LIB = os.path.join(PKG, "librfd.so")

HIP_SOURCES = ["rfd_capi.hip", "rfd_batch.hip", "rfd_bounds.hip", "rfd_dec.hip", "rfd_mlp.hip"]
# Per-file additions: the fused bound sweep's unit has no NaN on its path and
# is compiled without NaN semantics (see the head of rfd_bounds.hip).
EXTRA_FLAGS = {"rfd_bounds.hip": ["-fno-honor-nans"]}
CXX_SOURCES = ["wrapper.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _jobs() -> int:
    for var in ("MAX_JOBS", "RFD_BUILD_JOBS"):
        if os.environ.get(var, "").isdigit():
            return max(1, int(os.environ[var]))
    return max(1, min(8, os.cpu_count() or 1))


def build(verbose: bool = False) -> str:
    obj_dir = os.path.join(PKG, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    cmds = []
    objs = []
    for src in HIP_SOURCES + CXX_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *HIPCC_FLAGS, *EXTRA_FLAGS.get(src, []), "-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [hipcc(), "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
                   f"-I{INCLUDE}", f"-I{CSRC}", "-c", path, "-o", obj]
        cmds.append(cmd)
        objs.append(obj)
    # The translation units are independent: compile them side by side.
    with concurrent.futures.ThreadPoolExecutor(max_workers=_jobs()) as pool:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
        for fut in [pool.submit(subprocess.run, cmd, check=True) for cmd in cmds]:
            fut.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # The reference CLI (fd <c1> <c2> <dim> [1]) linked against the library.
    main = os.path.join(CSRC, "fd_main.cpp")
    if os.path.exists(main):
        cmd = [hipcc(), "-O2", "-std=c++17", f"-I{INCLUDE}", main, "-o", os.path.join(PKG, "fd"),
               f"-L{PKG}", "-lrfd", f"-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
