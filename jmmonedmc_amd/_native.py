"""ctypes declarations of the C ABI in include/jmm_engine.h.

The library is loaded from the package's build directory only; there is no
fallback implementation, so a missing or stale build fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

from .build import LIB

POT_LJ, POT_HARMONIC, POT_HRSW, POT_YUKAWA = 0, 1, 2, 3
WALLS, PERIODIC = 0, 1
NLT, NPT = 0, 1
VOL_LINEAR, VOL_LOG = 0, 1
FORM_PLAIN_FP64, FORM_BATCHED_FP64, FORM_BATCHED_FP32 = 0, 1, 2
FORMS = {"plain_fp64": FORM_PLAIN_FP64, "batched_fp64": FORM_BATCHED_FP64,
         "batched_fp32": FORM_BATCHED_FP32}

EXPORTED = [
    "jmm_abi_version", "jmm_last_error", "jmm_schedule", "jmm_create", "jmm_destroy",
    "jmm_run_sweeps", "jmm_synchronize", "jmm_full_energy", "jmm_read_samples", "jmm_read_samples_virial", "jmm_read_samples_noalloc",
    "jmm_get_state", "jmm_get_states", "jmm_set_state", "jmm_set_states", "jmm_perturb_energy", "jmm_get_counters",
    "jmm_read_accept_log", "jmm_kernel_time", "jmm_observables", "jmm_observables_device", "jmm_pair_observables",
    "jmm_hist_configure", "jmm_hist_sample", "jmm_hist_read", "jmm_set_steps",
    "jmm_scaled_energy", "jmm_set_lengths", "jmm_chain_layout",
]


class ChainSpec(C.Structure):
    _fields_ = [
        ("potential", C.c_int32), ("boundary", C.c_int32), ("ensemble", C.c_int32),
        ("vol_move", C.c_int32), ("nbn", C.c_int32), ("shifted", C.c_int32),
        ("N", C.c_uint64), ("T", C.c_double), ("P", C.c_double), ("L0", C.c_double),
        ("max_step", C.c_double), ("max_dv", C.c_double), ("cutoff", C.c_double),
        ("pot_p", C.c_double * 4), ("seed", C.c_uint64), ("chain_id", C.c_uint32),
        ("eci", C.c_uint32), ("bcap", C.c_uint32), ("reserved", C.c_uint32),
    ]


class EngineOpts(C.Structure):
    _fields_ = [
        ("form", C.c_int32), ("device", C.c_int32), ("sample_capacity", C.c_uint32),
        ("block_sweeps", C.c_uint32), ("wg_threads", C.c_uint32), ("use_graph", C.c_uint32),
        ("log_accepts", C.c_uint32), ("lds_budget", C.c_uint32), ("seg_lds", C.c_uint32),
        ("seg_lanes", C.c_uint32), ("sample_virial", C.c_uint32), ("reserved", C.c_uint32),
    ]


class Sample(C.Structure):
    _fields_ = [("U", C.c_double), ("L", C.c_double), ("acc_d", C.c_uint32), ("acc_v", C.c_uint32)]


class HistSpec(C.Structure):
    _fields_ = [("rhonb", C.c_uint32), ("gns", C.c_uint32), ("gnb", C.c_uint32),
                ("reserved", C.c_uint32), ("rbw", C.c_double), ("gsw", C.c_double),
                ("gbw", C.c_double)]


MODE_LDS, MODE_HBM, MODE_SPLIT = 0, 1, 2
MODES = {MODE_LDS: "lds", MODE_HBM: "hbm", MODE_SPLIT: "split"}


class Layout(C.Structure):
    _fields_ = [("B", C.c_uint32), ("w", C.c_uint32), ("mode", C.c_uint32),
                ("segments", C.c_uint32), ("lds_bytes", C.c_uint64), ("lanes", C.c_uint32),
                ("reserved", C.c_uint32)]


class Counters(C.Structure):
    _fields_ = [
        ("sweeps", C.c_uint64), ("trials", C.c_uint64), ("acc_d", C.c_uint64),
        ("wall_rejects", C.c_uint64), ("core_rejects", C.c_uint64), ("pair_terms", C.c_uint64),
        ("vol_trials", C.c_uint64), ("acc_v", C.c_uint64), ("recomputes", C.c_uint64),
        ("max_drift", C.c_double), ("U", C.c_double), ("L", C.c_double),
    ]


class JmmError(RuntimeError):
    pass


_lib = None


def lib(path: str = LIB):
    """Load libjmm_engine.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise JmmError(f"{path} is missing: build it with `python -m jmmonedmc_amd.build` "
                       "(there is no non-GPU implementation)")
    L = C.CDLL(path)
    dp, u64p, u32p = C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)
    vp = C.c_void_p
    L.jmm_abi_version.restype = C.c_int
    L.jmm_last_error.restype = C.c_char_p
    L.jmm_schedule.argtypes = [C.POINTER(ChainSpec), u64p, u64p, u64p]
    L.jmm_create.argtypes = [C.POINTER(ChainSpec), C.c_int, C.POINTER(EngineOpts), C.POINTER(vp)]
    L.jmm_destroy.argtypes = [vp]
    L.jmm_run_sweeps.argtypes = [vp, C.c_uint32]
    L.jmm_synchronize.argtypes = [vp]
    L.jmm_full_energy.argtypes = [vp, dp]
    L.jmm_read_samples.argtypes = [vp, C.POINTER(Sample), C.c_uint32, u32p]
    L.jmm_read_samples_virial.argtypes = [vp, C.POINTER(Sample), dp, C.c_uint32, u32p]
    L.jmm_read_samples_noalloc.argtypes = [vp, C.POINTER(Sample), C.c_uint32, u32p]
    L.jmm_get_state.argtypes = [vp, C.c_int, dp, dp, dp, u64p]
    L.jmm_get_states.argtypes = [vp, dp, dp, dp]
    L.jmm_set_state.argtypes = [vp, C.c_int, dp, C.c_double, C.c_uint64]
    L.jmm_set_states.argtypes = [vp, dp, dp, dp, u64p]
    L.jmm_perturb_energy.argtypes = [vp, C.c_int, C.c_double]
    L.jmm_get_counters.argtypes = [vp, C.c_int, C.POINTER(Counters)]
    L.jmm_read_accept_log.argtypes = [vp, C.POINTER(C.c_uint8), C.c_uint32, u32p]
    L.jmm_kernel_time.argtypes = [vp, u64p, dp, u64p, C.c_int]
    L.jmm_observables.argtypes = [vp, dp, dp]
    L.jmm_observables_device.argtypes = [vp, vp, vp]
    L.jmm_pair_observables.argtypes = [C.POINTER(ChainSpec), C.c_int, vp, vp, vp, vp, C.c_int, vp]
    L.jmm_hist_configure.argtypes = [vp, C.POINTER(HistSpec), C.c_uint32]
    L.jmm_hist_read.argtypes = [vp, u64p, u64p, u64p, C.c_int]
    L.jmm_hist_sample.argtypes = [vp]
    L.jmm_set_steps.argtypes = [vp, dp, dp]
    L.jmm_scaled_energy.argtypes = [vp, dp, dp]
    L.jmm_set_lengths.argtypes = [vp, dp]
    for name in EXPORTED:
        getattr(L, name).restype = C.c_char_p if name == "jmm_last_error" else C.c_int
    _lib = L
    return L


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().jmm_last_error()
        raise JmmError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
