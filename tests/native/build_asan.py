"""Builds the AddressSanitizer variant of the engine's host code and the C ABI
lifecycle program that tests/test_abi_leaks.py runs.  Only host code is
instrumented: the library links the regular build's kernel object, and the
sanitizer runtime comes with the (host-only) test program.

    python tests/native/build_asan.py        # after the regular build
"""
import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from jmmonedmc_amd.build import ARCH, CSRC, HIP_FLAGS, HIPCC, INCLUDE, OUT  # noqa: E402

LIB = os.path.join(OUT, "libjmm_engine_asan.so")
PROG = os.path.join(OUT, "abi_lifecycle_asan")
CLANG = "/opt/rocm/llvm/bin/clang++"
HOST_SAN = shlex.split("-g -Xarch_host -fsanitize=address -Xarch_host -fno-omit-frame-pointer")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"asan build step failed: {' '.join(cmd[:4])} ...")


def build() -> str:
    kobj = os.path.join(OUT, "kernels.o")
    if not os.path.exists(kobj):
        raise RuntimeError(f"{kobj} missing: run the regular build first")
    eobj = os.path.join(OUT, "engine_asan.o")
    _run([HIPCC, *HIP_FLAGS, *HOST_SAN, "-c", os.path.join(CSRC, "engine.hip"), "-o", eobj])
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB, eobj, kobj])
    _run([CLANG, "-O1", "-std=c++17", *HOST_SAN, "-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}",
          "-I/opt/rocm/include", os.path.join(HERE, "abi_lifecycle.cpp"), "-o", PROG,
          f"-L{OUT}", "-ljmm_engine_asan", "-L/opt/rocm/lib", "-lamdhip64",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    return PROG


if __name__ == "__main__":
    print(build())
