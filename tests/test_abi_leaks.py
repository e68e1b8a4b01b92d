"""Memory errors and leaks at the C ABI: the library's host code built with
AddressSanitizer (tests/native/build_asan.py, run by the build) and a
lifecycle program that calls every entry point, the error paths, repeated
create/destroy cycles and a creation that fails after allocating
(tests/native/abi_lifecycle.cpp).  The reference's setupMCS/freeMCS leak
files, the RNG and the Compute objects (src/jmmMCState.cpp:792-842); the
engine must not.

A run fails on any AddressSanitizer error, on any leak whose allocation
stack passes through the engine library or the program, on any failed
check, and (GPU) when device memory is not returned after destroy.  Leaks
wholly inside the ROCm runtime libraries are suppressed
(tests/native/lsan.supp)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "jmmonedmc_amd", "build", "abi_lifecycle_asan")
SUPP = os.path.join(ROOT, "tests", "native", "lsan.supp")
OURS = ("libjmm_engine_asan", "engine.hip", "kernels.hip", "abi_lifecycle")


def run(mode, timeout):
    assert os.path.exists(PROG), f"{PROG} missing: build with __graft_entry__.build()"
    env = dict(os.environ)
    # exitcode=23: the status a leak report ends the program with
    env["ASAN_OPTIONS"] = "detect_leaks=1:verify_asan_link_order=0:exitcode=23"
    env["LSAN_OPTIONS"] = f"suppressions={SUPP}:print_suppressions=0"
    sym = "/opt/rocm/llvm/bin/llvm-symbolizer"
    if os.path.exists(sym):
        env["ASAN_SYMBOLIZER_PATH"] = sym
    return subprocess.run([PROG, mode], capture_output=True, text=True, timeout=timeout, env=env)


def our_leaks(text):
    blocks = re.split(r"\n(?=(?:Direct|Indirect) leak of)", text)
    return [b for b in blocks if re.match(r"(Direct|Indirect) leak of", b)
            and any(o in b for o in OURS)]


def check(r, mode):
    out = r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in out, out[-6000:]
    ours = our_leaks(out)
    assert not ours, "leaks in the engine or the program:\n" + "\n".join(ours)[:6000]
    assert f"{mode}: 0 failed checks" in r.stdout, out[-6000:]
    # 23: LeakSanitizer found leaks, all of them outside the engine (above)
    assert r.returncode in (0, 23), (r.returncode, out[-6000:])
    assert "CHECK failed" not in out, out[-6000:]


def test_error_paths_leak_free():
    """Argument and specification errors (no device needed)."""
    check(run("errors", 60), "errors")


@pytest.mark.gpu
def test_lifecycle_leak_free_on_device(gpu):
    print("[gpu] abi lifecycle under AddressSanitizer", flush=True)
    check(run("full", 300), "full")
