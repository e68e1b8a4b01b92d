"""Ensemble averages against exact 1D NPT results (tests/exact1d.py): the
oracle on the CPU, and the engine's FP64 and FP32 forms on the GPU, within
3 standard errors.  Each case names its boundary and volume move.
"""
import math

import numpy as np
import pytest

from exact1d import (block_mean_sem, gap_moments, harmonic_phi, hrsw_phi, lj_phi, periodic_npt,
                     walls_npt)
from helpers import hrsw, lj, oracle_spec

from jmmonedmc_amd import _native as nat
from jmmonedmc_amd.engine import Chain


def case_hrsw_periodic(seed=0, chain_id=0):
    # periodic ring, ln L moves, hard rod 1 + well depth 1 to 1.5 (C3 potential)
    ch = Chain(N=40, rho=0.5, T=1.0, P=1.0, ensemble=nat.NPT, vol_move=nat.VOL_LOG, max_dv=0.03,
               seed=seed, chain_id=chain_id, **hrsw())
    mom = gap_moments(hrsw_phi(), 1.0, 1.0, 1.0, breaks=(1.5,))
    return ch, periodic_npt(40, mom)


def case_tonks_periodic(seed=0, chain_id=0):
    # hard rods only (well depth 0): Tonks gas
    ch = Chain(N=30, rho=0.5, T=1.3, P=0.6, ensemble=nat.NPT, vol_move=nat.VOL_LOG, max_dv=0.04,
               seed=seed, chain_id=chain_id, **hrsw(pot_p=(1.0, 0.0, 1.5, 0.0)))
    b, P = 1 / 1.3, 0.6
    mom = dict(m=1 + 1 / (b * P), v=1 / (b * P) ** 2, ephi=0.0, ephig=0.0)
    return ch, periodic_npt(30, mom)


def case_harmonic_walls(seed=125, chain_id=0):
    # reference test/INPUTstd: harmonic, NBN 1, walls, linear dL, N ln s
    ch = Chain(N=10, L0=10.0, T=0.4, P=0.7, ensemble=nat.NPT, boundary=nat.WALLS,
               vol_move=nat.VOL_LINEAR, max_dv=1.0, max_step=0.1, nbn=1, seed=seed,
               chain_id=chain_id, potential=nat.POT_HARMONIC, cutoff=math.inf,
               pot_p=(0, 0, 0, 0), shifted=0)
    mom = gap_moments(harmonic_phi(), 1 / 0.4, 0.7, 0.0)
    return ch, walls_npt(10, mom, 1 / 0.4, 0.7)


def case_lj_nbn1_walls(seed=3, chain_id=0):
    # reference LJ (no cutoff) with NBN 1, walls, linear dL
    ch = Chain(N=20, L0=24.0, T=1.0, P=0.5, ensemble=nat.NPT, boundary=nat.WALLS,
               vol_move=nat.VOL_LINEAR, max_dv=3.0, max_step=0.3, nbn=1, seed=seed,
               chain_id=chain_id, **lj(cutoff=math.inf, shifted=0))
    mom = gap_moments(lj_phi(), 1.0, 0.5, 0.5, breaks=(1.0, 1.5, 3.0))
    return ch, walls_npt(20, mom, 1.0, 0.5)


def case_lj_nbn1_periodic_log(seed=4, chain_id=0):
    # shifted LJ rc=3 with NBN 1 on a ring, ln L moves
    shift = 4 * (3.0 ** -12 - 3.0 ** -6)
    ch = Chain(N=24, rho=0.31, T=0.8, P=0.3, ensemble=nat.NPT, vol_move=nat.VOL_LOG, max_dv=0.03,
               max_step=0.4, nbn=1, seed=seed, chain_id=chain_id, **lj())
    mom = gap_moments(lj_phi(rc=3.0, shift=shift), 1 / 0.8, 0.3, 0.5, breaks=(1.0, 1.5, 3.0))
    return ch, periodic_npt(24, mom)


CASES = {"hrsw_periodic": case_hrsw_periodic, "tonks_periodic": case_tonks_periodic,
         "harmonic_walls": case_harmonic_walls, "lj_nbn1_walls": case_lj_nbn1_walls,
         "lj_nbn1_periodic_log": case_lj_nbn1_periodic_log}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_exact_npt(oracle, name):
    """12 independent chains (distinct seeds) of 40000 sweeps; the spread of
    the per-chain means gives the standard error."""
    import pyoracle
    Ls, Us = [], []
    for k in range(12):
        ch, (L_ex, U_ex) = CASES[name](seed=700 + k)
        c = oracle.Chain(oracle_spec(pyoracle, ch))
        c.sweeps(2000)                      # equilibrate
        smp = c.sweeps(40000)
        Ls.append(smp["L"].mean())
        Us.append(smp["U"].mean())
    for vals, ex, what in ((np.array(Ls), L_ex, "L"), (np.array(Us), U_ex, "U")):
        m, se = vals.mean(), vals.std(ddof=1) / math.sqrt(len(vals))
        assert abs(m - ex) <= 3 * se, (what, m, ex, se)


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["batched_fp64", "batched_fp32", "plain_fp64"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_engine_matches_exact_npt(gpu, name, form):
    """128 independent chains (distinct seeds); the spread of per-chain means
    gives the standard error."""
    from jmmonedmc_amd.engine import Engine
    chains, exact = [], None
    for k in range(128):
        ch, exact = CASES[name](seed=1000 + k, chain_id=k)
        chains.append(ch)
    print(f"[gpu] exact {name} {form}", flush=True)
    eng = Engine(chains, form=form, sample_capacity=4000, block_sweeps=500)
    eng.run(3000)                       # equilibrate from the lattice start
    eng.read_samples(4000)
    eng.run(4000)
    smp = eng.read_samples(4000)
    eng.close()
    L_ex, U_ex = exact
    Lc = smp["L"].mean(axis=1)
    Uc = smp["U"].mean(axis=1)
    for vals, ex, what in ((Lc, L_ex, "L"), (Uc, U_ex, "U")):
        m, se = vals.mean(), vals.std(ddof=1) / math.sqrt(len(vals))
        assert abs(m - ex) <= 3 * se, (what, m, ex, se)
