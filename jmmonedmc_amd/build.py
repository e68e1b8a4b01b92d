"""Builds the engine's shared library (gfx950) and the host driver in-tree.

    python -m jmmonedmc_amd.build            # library + driver
The library lands in jmmonedmc_amd/build/libjmm_engine.so; the driver in
jmmonedmc_amd/build/jmm_onedmc.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OUT = os.path.join(PKG, "build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(OUT, "libjmm_engine.so")
DRIVER = os.path.join(OUT, "jmm_onedmc")

ARCH = os.environ.get("JMM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the FP64 kernels round once per operation in a fixed
# order (bit-reproducible positions); packed FP32 instructions are disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}",
]
SOURCES = ["kernels.hip", "engine.hip"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build_engine(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_driver(verbose: bool = False) -> str:
    """C++ host driver (INPUT parser, jmmMCState-style facade, main loop)."""
    os.makedirs(OUT, exist_ok=True)
    srcs = [os.path.join(HOST, f) for f in ("readInput.cpp", "mcstate.cpp", "main.cpp")]
    srcs = [s for s in srcs if os.path.exists(s)]
    if not srcs:
        return ""
    tmp = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-fopenmp", f"-I{INCLUDE}", f"-I{HOST}", "-o", tmp, *srcs,
          f"-L{OUT}", "-ljmm_engine", f"-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"], verbose)
    os.replace(tmp, DRIVER)
    return DRIVER


def build(verbose: bool = False) -> None:
    build_engine(verbose)
    build_driver(verbose)


if __name__ == "__main__":
    build(verbose=True)
