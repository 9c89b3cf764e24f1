"""Builds the HIP engine library in-tree with hipcc for gfx950.

``python -m mrilab_amd.build`` (or ``build()``) compiles mrilab_amd/csrc/*
into mrilab_amd/libmrilab_amd.so. Packed fp32 instructions are disabled
(``-packed-fp32-ops``) for every HIP translation unit.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmrilab_amd.so")
ARCH = os.environ.get("MRS_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["engine.hip", "multipool.hip", "batch.hip", "adjoint.hip", "recon.hip", "sar.hip", "sar_average.hip"]
CXX_SOURCES = ["decode.cpp", "seqgen.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the engine needs ROCm's hipcc to build")


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compiles every source (in parallel, ``jobs`` at a time: default the
    host's cores, at most 16) and links the library."""
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    objs, cmds = [], []
    common = ["-O3", "-fPIC", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
              "-Wall", "-Wno-unused-result"]
    for src in HIP_SOURCES + CXX_SOURCES:
        obj = os.path.join(HERE, "_build", os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", *common, "-c", os.path.join(CSRC, src), "-o", obj]
        if src in CXX_SOURCES:
            cmd = [hipcc(), *common, "-x", "c++", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(obj)
    n = jobs or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=n) as pool:
        for r in list(pool.map(lambda c: subprocess.run(c), cmds)):
            if r.returncode != 0:
                raise subprocess.CalledProcessError(r.returncode, r.args)
    tmp = LIB + ".tmp"
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(hipcc()))), "lib")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, f"-L{rocm_lib}", "-lhipfft",
           f"-Wl,-rpath,{rocm_lib}", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
