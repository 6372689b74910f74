"""Builds libgpw_amd/_lib/libgpw_amd.so for gfx950 with hipcc (in-tree).

Kernels are compiled with FP contraction on (FMA) and packed FP32 disabled;
the host planner with -ffp-contract=off so cube centres and sphere bounds are
reproducible bit for bit.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libgpw_amd.so")

ARCH = os.environ.get("GPW_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    # no packed FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

SOURCES = [
    ("gpw_tiles_colloc_ortho.hip", []),
    ("gpw_tiles_colloc_general.hip", []),
    ("gpw_tiles_integ_ortho.hip", []),
    ("gpw_tiles_integ_general.hip", []),
    ("gpw_kernels.hip", []),
    ("gpw_plan.cpp", ["-ffp-contract=off"]),
    ("gpw_single.cpp", ["-ffp-contract=off"]),
    ("gpw_folded.cpp", []),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs, cmds = [], []
    for src, extra in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmds.append([HIPCC, "-x", "hip", *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the translation units compile independently: run them side by side
    from concurrent.futures import ThreadPoolExecutor
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd))
    with ThreadPoolExecutor(max_workers=min(len(cmds), max(1, os.cpu_count() or 1))) as ex:
        for f in [ex.submit(_run, cmd) for cmd in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    # task-file replay driver: plain C++ against the drop-in header and the
    # C ABI (the batch mode allocates its device buffers with the HIP runtime)
    exe = os.path.join(LIBDIR, "gpw_replay")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    _run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
          "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(rocm, "include"),
          os.path.join(CSRC, "gpw_replay_main.cpp"), "-o", exe + ".tmp", "-L" + LIBDIR, "-lgpw_amd",
          "-L" + os.path.join(rocm, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath," + os.path.join(rocm, "lib")])
    os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
