"""Builds libbh_potential.so (HIP kernels + C ABI) in-tree for gfx950.

    python -m bipoly_hydro_amd.build

hipcc cross-compiles without a GPU.  The library lands next to this file so
that it travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbh_potential.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["bh_kernels.hip", "bh_gemm.hip", "bh_mixed.hip", "bh_ops.cpp", "bh_resources.cpp", "bh_plan.cpp",
           "bh_exec.cpp", "bh_batch.cpp", "bh_density.hip", "bh_serve.cpp",
           "bh_dft.hip", "bh_spectral.hip", "bh_noalloc.hip", "bh_cont.cpp", "bh_hydro.hip", "bh_spec64.hip",
           "bh_hostcopy.cpp"]

# no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]


def build(verbose: bool = False, jobs: int = 0) -> str:
    from concurrent.futures import ThreadPoolExecutor
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(run, cmds))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
