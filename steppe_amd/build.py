"""Build libsteppe_hip.so (the HIP engine + C ABI) in-tree for gfx950.

    python -m steppe_amd.build          # compiles steppe_amd/csrc/* -> steppe_amd/lib/libsteppe_hip.so

Each translation unit is compiled separately (in parallel) with hipcc and linked
into one shared library.  All device code is built with packed-FP32 instruction
selection disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "lib", "obj")
LIB = os.path.join(LIBDIR, "libsteppe_hip.so")

ARCH = os.environ.get("STEPPE_OFFLOAD_ARCH", "gfx950")
SOURCES = ["decode.hip", "f2_fp64.hip", "f2_split.hip", "f2_fixup.hip", "synth.hip", "qpadm.hip", "qpadm_large.hip", "qpgraph.hip", "fstats.hip",
           "frontend.hip", "engine.cpp", "f2_graph.cpp", "autotune.cpp", "qpadm_host.cpp", "qpgraph_host.cpp", "qpgraph_enum.cpp",
           "qpgraph_search.cpp", "fstats_host.cpp", "geno_io.cpp", "frontend_host.cpp", "tiered.cpp", "capi.cpp"]
# Built without FP contraction: each product is rounded before it is added, as in the
# reference's fit chain (plain double sums in a fixed order).
NO_CONTRACT = {"qpadm.hip", "qpadm_large.hip", "qpgraph.hip", "fstats.hip", "frontend.hip", "f2_fixup.hip"}


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP engine)")


def common_flags() -> list:
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-Wall",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _newest_header() -> float:
    heads = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    inc = os.path.join(ROOT, "include")
    heads += [os.path.join(inc, f) for f in os.listdir(inc) if f.endswith(".h")]
    return max(os.path.getmtime(h) for h in heads)


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    if incremental and os.path.exists(obj):
        t = os.path.getmtime(obj)
        if t > os.path.getmtime(os.path.join(CSRC, src)) and t > _newest_header():
            return obj
    cmd = [hipcc(), *common_flags(), "-c", os.path.join(CSRC, src), "-o", obj]
    if src in NO_CONTRACT:
        cmd.insert(-4, "-ffp-contract=off")
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compile every source (incremental=True: only those older than their object
    file or any header) and link the library."""
    os.makedirs(OBJDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda f: _compile(f, incremental), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
    sys.exit(0)
