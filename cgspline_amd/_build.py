"""Compiles libcgs.so (HIP kernels + C ABI) for gfx950, in-tree.

Numerics flags: no FMA contraction, IEEE f32 division/sqrt and denormals, so
the kernels reproduce the reference host path's FP32 results bit for bit.
Packed FP32 instructions are disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcgs.so")
SOURCES = ["cgs_lib.cpp", "cgs_trace.hip", "cgs_render_sampled.hip", "cgs_phf.hip", "cgs_models.cpp", "cgs_dl.cpp",
           "cgs_cull.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, jobs: int = 3, defines=(), lib_path: str = LIB) -> str:
    """Compiles every source to an object in csrc/_build, then links LIB."""
    out_dir = os.path.join(CSRC, "_build" + ("_" + "_".join(d.replace("=", "") for d in defines) if defines else ""))
    os.makedirs(out_dir, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(out_dir, src.rsplit(".", 1)[0] + ".o")
        objs.append(obj)
        cmd = [cc, *HIPCC_FLAGS, *["-D" + d for d in defines], "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = lib_path + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, lib_path)
    return lib_path


COUNTING_LIB = os.path.join(HERE, "libcgs_counts.so")


def build_counting(verbose: bool = False) -> str:
    """The work-counting variant (CGS_WORK_COUNTS=1): the same kernels, each
    counting its work events (cgs_work_counts), for the work model's counts
    of a launch.  Never the one timed: counting slows the kernels."""
    return build(verbose, defines=("CGS_WORK_COUNTS=1",), lib_path=COUNTING_LIB)


def build_facade_check(verbose: bool = False) -> str:
    """Compiles tests/cpp/facade_check (the C++ facade driven like the
    reference's callers) with g++ against the in-tree libcgs.so."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "facade_check.cpp")
    exe = os.path.join(root, "tests", "cpp", "facade_check")
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-I", os.path.join(root, "include"), src, "-o", exe + ".tmp",
           "-L", HERE, "-lcgs", "-Wl,-rpath,$ORIGIN/../../cgspline_amd", "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("facade_check build failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(exe + ".tmp", exe)
    return exe


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    text = out.decode(errors="replace")
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{text}")
    warn = [l for l in text.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    if warn:
        sys.stderr.write("\n".join(warn[:20]) + "\n")


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_facade_check(verbose=True))
