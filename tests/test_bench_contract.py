"""bench.py's output contract, on the CPU: the refusal of more than one GPU,
the algorithmic FLOP and transcendental counts against hand-computed
values, and the CPU baselines on a small sample."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from jmmonedmc_amd import _native as nat  # noqa: E402
from jmmonedmc_amd.engine import Chain  # noqa: E402

from helpers import hrsw, lj, yukawa  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


class _Counters:
    """Stands in for an engine: all_counters() before and after a run."""

    def __init__(self, terms, trials):
        self.terms, self.trials = np.asarray(terms, np.uint64), np.asarray(trials, np.uint64)

    def all_counters(self):
        return {"pair_terms": self.terms, "trials": self.trials}


def test_algorithmic_flops_and_transcendentals_by_hand():
    chains = [Chain(N=10, L0=20.0, **lj()), Chain(N=10, L0=20.0, **hrsw()),
              Chain(N=10, L0=20.0, **yukawa()),
              Chain(N=10, L0=20.0, potential=nat.POT_HARMONIC, cutoff=float("inf"),
                    pot_p=(0, 0, 0, 0), shifted=0)]
    before = {"pair_terms": np.array([5, 0, 7, 1], np.uint64), "trials": np.array([1, 2, 3, 4], np.uint64)}
    eng = _Counters([105, 40, 207, 21], [51, 22, 53, 14])
    flops, terms, trials, trans = bench.algorithmic_flops(eng, chains, before)
    # pair evaluations 100, 40, 200, 20 and trials 50, 20, 50, 10:
    # F = f_pair n + n + 8 trials with f_pair = 10 (LJ), 4 (rod + well),
    # 10 (Yukawa), 4 (spring)
    assert terms == 360 and trials == 130
    assert flops == (10 * 100 + 100 + 400) + (4 * 40 + 40 + 160) + (10 * 200 + 200 + 400) + (4 * 20 + 20 + 80)
    # one reciprocal per LJ pair, exponential + reciprocal per Yukawa pair,
    # one Metropolis exponential per trial
    assert trans == 100 + 0 + 400 + 0 + 130


def test_cpu_baselines_on_a_small_sample():
    chains = [Chain(N=64, L0=128.0, seed=1, **lj())]
    v, dt, tr = bench.cpu_baseline(chains, 50, 1)
    assert tr == 50 * 64 and v > 0 and dt > 0
    va, _, tra = bench.cpu_baseline(chains, 50, 1, allpairs=True)
    assert tra == tr and va > 0
