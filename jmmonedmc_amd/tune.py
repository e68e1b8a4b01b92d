"""Autotuner for the batched forms' launch parameters.

The parameters tuned here change how chains are packed into workgroups and
launched, never what a chain computes: a chain's trials, its random numbers
and its reductions are the same under every candidate (the tests check that
the samples are bit-identical).  For a given chain mix the tuner times a
short run of every candidate with the engine's own HIP-event kernel timer
and keeps the fastest.

Candidates (candidates_for): the LDS budget per workgroup (how many chains
share one), the workgroup width (256, 512 or 1024 lanes: wider workgroups
admit chains of up to that many blocks and put more blocks in one; 1024
lanes cap the kernel at 128 VGPRs, so they spill),
whether sweep blocks replay a captured hipGraph, and, for chains too large
for one workgroup's LDS, the split threshold (how many workgroups share a
chain) and the lanes per block of a split chain (1 or 2).  Results can be
cached in a JSON file keyed by a signature of the chain mix and the device.
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from dataclasses import dataclass, field

import numpy as np

from ._native import JmmError
from .engine import Chain, Engine

DEFAULT_CANDIDATES = [
    dict(lds_budget=32768, wg_threads=256, use_graph=True),
    dict(lds_budget=49152, wg_threads=256, use_graph=True),
    dict(lds_budget=65536, wg_threads=256, use_graph=True),
    dict(lds_budget=98304, wg_threads=256, use_graph=True),
    dict(lds_budget=65536, wg_threads=256, use_graph=False),
]


# a chain whose LDS image (8 bytes per particle and its ghosts) exceeds
# this many bytes may be split over several workgroups
SPLIT_CANDIDATE_BYTES = 96 * 1024


def candidates_for(chains) -> list:
    """DEFAULT_CANDIDATES, the 512- and 1024-lane workgroups at two LDS
    budgets each, and,
    when a chain may be split, split thresholds of 64 and 32 KiB per
    workgroup and two lanes per block of a split chain."""
    cands = [dict(c) for c in DEFAULT_CANDIDATES]
    cands += [dict(lds_budget=65536, wg_threads=512, use_graph=True),
              dict(lds_budget=98304, wg_threads=512, use_graph=True),
              dict(lds_budget=65536, wg_threads=1024, use_graph=True),
              dict(lds_budget=98304, wg_threads=1024, use_graph=True)]
    if any(8 * c.N > SPLIT_CANDIDATE_BYTES for c in chains):
        base = dict(lds_budget=65536, wg_threads=256, use_graph=True)
        cands += [dict(base, seg_lanes=2), dict(base, seg_lds=65536), dict(base, seg_lds=32768)]
    return cands


@dataclass
class TuneResult:
    best: dict
    ms_per_sweep: dict = field(default_factory=dict)   # candidate key -> ms per sweep (None: refused)
    signature: str = ""


def signature(chains, form: str) -> str:
    """Stable key of a chain mix: potentials, ensembles, boundaries and a
    histogram of N (the packing depends on these, not on seeds)."""
    rows = sorted((c.potential, c.ensemble, c.boundary, c.nbn, round(c.cutoff, 6),
                   int(np.log2(max(c.N, 1)) * 4)) for c in chains)
    h = hashlib.sha1(json.dumps([form, len(chains), rows]).encode()).hexdigest()[:16]
    return h


def _key(c: dict) -> str:
    return ",".join(f"{k}={c[k]}" for k in sorted(c))


def tune(chains, form: str = "batched_fp32", candidates=None, block_sweeps: int = 100,
         blocks: int = 2, device: int = 0, cache: str | None = None) -> TuneResult:
    """Time every candidate on `blocks` sweep blocks (after one warm-up block)
    and return the fastest.  With `cache` (a JSON path) a stored result for
    the same signature is returned without timing, and new results are
    stored."""
    chains = list(chains)
    sig = signature(chains, form) + f"/dev{device}"
    if cache and os.path.exists(cache):
        with open(cache) as f:
            stored = json.load(f)
        if sig in stored:
            r = stored[sig]
            return TuneResult(best=r["best"], ms_per_sweep=r["ms_per_sweep"], signature=sig)
    cands = candidates or candidates_for(chains)
    times = {}
    for c in cands:
        opts = dict(c)
        cap = (blocks + 1) * block_sweeps
        try:
            e = Engine(chains, form=form, device=device, sample_capacity=cap, block_sweeps=block_sweeps, **opts)
        except JmmError:
            # a candidate the device cannot hold (split chains whose
            # workgroups would not all be resident) is refused at creation
            times[_key(c)] = None
            continue
        with e:
            e.run(block_sweeps)
            e.sync()
            e.kernel_time(reset=True)
            e.run(blocks * block_sweeps)
            e.sync()
            n, ms, sw = e.kernel_time(reset=True)
        times[_key(c)] = ms / max(sw, 1)
    ok = [c for c in cands if times[_key(c)] is not None]
    if not ok:
        raise JmmError("tune: no candidate could be created for these chains")
    best = min(ok, key=lambda c: times[_key(c)])
    res = TuneResult(best=dict(best), ms_per_sweep=times, signature=sig)
    if cache:
        stored = {}
        if os.path.exists(cache):
            with open(cache) as f:
                stored = json.load(f)
        stored[sig] = {"best": res.best, "ms_per_sweep": times, "time": time.time()}
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(stored, f, indent=1)
    return res
