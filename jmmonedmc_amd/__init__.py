"""jmmonedmc_amd -- MI355X-native Metropolis engine for 1D classical particles.

The hot path (displacement trials, box-length trials, many chains) runs in
hand-written gfx950 HIP kernels behind the C ABI of include/jmm_engine.h;
this package is its Python view: the Engine class, the workload builders,
statistics and the autograd binding of the backward kernel.
"""
from . import _native
from ._native import (FORM_BATCHED_FP32, FORM_BATCHED_FP64, FORM_PLAIN_FP64, NLT, NPT, PERIODIC,
                      POT_HARMONIC, POT_HRSW, POT_LJ, POT_YUKAWA, VOL_LINEAR, VOL_LOG, WALLS,
                      JmmError)
from .engine import Chain, Engine, schedule

__all__ = [
    "Chain", "Engine", "schedule", "JmmError", "FORM_PLAIN_FP64", "FORM_BATCHED_FP64",
    "FORM_BATCHED_FP32", "POT_LJ", "POT_HARMONIC", "POT_HRSW", "POT_YUKAWA", "WALLS", "PERIODIC",
    "NLT", "NPT", "VOL_LINEAR", "VOL_LOG",
]
