"""Ensemble averages against exact 1D NPT results (tests/exact1d.py): the
oracle on the CPU, and the engine's FP64 and FP32 forms on the GPU, within
3 standard errors.  Each case names its boundary and volume move.
"""
import math

import numpy as np
import pytest

from exact1d import (block_mean_sem, gap_moments, harmonic_phi, hrsw_phi, lj_phi, periodic_npt,
                     walls_npt, walls_nvt_small, yukawa_phi)
from helpers import hrsw, lj, oracle_spec, yukawa

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


# Two and three particles between walls at fixed length: every pair
# interacts, so with three particles the outer pair (two places apart) is
# in the window walk; exact averages by direct integration.
SHIFT3 = 4 * (3.0 ** -12 - 3.0 ** -6)
SMALL = {
    "lj2_walls": lambda seed, cid: (
        Chain(N=2, L0=5.0, T=1.0, boundary=nat.WALLS, seed=seed, chain_id=cid, **lj()),
        walls_nvt_small(2, 5.0, 1.0, lj_phi(rc=3.0, shift=SHIFT3), 0.5, breaks=(1.0, 1.5, 3.0))),
    "lj3_walls": lambda seed, cid: (
        Chain(N=3, L0=4.0, T=1.0, boundary=nat.WALLS, max_step=0.3, seed=seed, chain_id=cid, **lj()),
        walls_nvt_small(3, 4.0, 1.0, lj_phi(rc=3.0, shift=SHIFT3), 0.5, breaks=(0.9, 1.12, 1.5, 3.0))),
    "yukawa3_walls": lambda seed, cid: (
        Chain(N=3, L0=8.0, T=1.5, boundary=nat.WALLS, seed=seed, chain_id=cid, **yukawa()),
        walls_nvt_small(3, 8.0, 1 / 1.5, yukawa_phi(), 0.5, breaks=(1.0, 2.0))),
    "yukawa3_nbn1_walls": lambda seed, cid: (
        Chain(N=3, L0=8.0, T=1.5, boundary=nat.WALLS, nbn=1, seed=seed, chain_id=cid, **yukawa()),
        walls_nvt_small(3, 8.0, 1 / 1.5, yukawa_phi(), 0.5, nbn=1, breaks=(1.0, 2.0))),
}


def test_small_walls_integrals_against_closed_form_and_grid():
    # two hard rods (length 1) with a well of depth eps out to 1.5 in L = 4:
    # the gap d has weight (L - d) e^{beta eps} on [1, 1.5), (L - d) beyond
    L, beta, eps = 4.0, 1.0 / 0.7, 1.0
    w_in = math.exp(beta * eps) * (L * 0.5 - (1.5 ** 2 - 1.0) / 2)
    w_out = (L - 1.5) ** 2 / 2
    exact = -eps * w_in / (w_in + w_out)
    got = walls_nvt_small(2, L, beta, hrsw_phi(1.0, eps, 1.5), 1.0, breaks=(1.5,))
    assert abs(got - exact) <= 1e-9 * abs(exact)
    # three Yukawa particles: midpoint sum over the gap triangle
    L, beta = 8.0, 1 / 1.5
    phi = yukawa_phi()
    n = 2400
    h = (L - 1.0) / n
    g = 0.5 + h * (np.arange(n) + 0.5)
    g1, g2 = np.meshgrid(g, g, indexing="ij")
    ok = g1 + g2 <= L
    f = lambda d: np.exp(-0.2 * d) / d
    u = f(g1) + f(g2) + f(g1 + g2)
    wgt = np.where(ok, (L - g1 - g2) * np.exp(-beta * u), 0.0)
    grid = float((wgt * u).sum() / wgt.sum())
    got = walls_nvt_small(3, L, beta, phi, 0.5, breaks=(1.0, 2.0))
    assert abs(got - grid) <= 2e-4 * abs(grid)
    # the outer pair matters: dropping it (NBN 1) changes <U> by far more
    # than the tests' standard errors
    assert got - walls_nvt_small(3, L, beta, phi, 0.5, nbn=1, breaks=(1.0, 2.0)) > 0.05


@pytest.mark.parametrize("name", sorted(SMALL))
def test_oracle_matches_exact_small_walls(oracle, name):
    """12 independent chains of 40000 sweeps; the spread of the per-chain
    means gives the standard error."""
    import pyoracle
    Us, U_ex = [], None
    for k in range(12):
        ch, U_ex = SMALL[name](800 + k, k)
        c = oracle.Chain(oracle_spec(pyoracle, ch))
        c.sweeps(2000)
        Us.append(c.sweeps(40000)["U"].mean())
    vals = np.array(Us)
    m, se = vals.mean(), vals.std(ddof=1) / math.sqrt(len(vals))
    assert abs(m - U_ex) <= 3 * se, (m, U_ex, se)


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["batched_fp64", "batched_fp32", "plain_fp64"])
@pytest.mark.parametrize("name", sorted(SMALL))
def test_engine_matches_exact_small_walls(gpu, name, form):
    """256 independent chains in one engine; the spread of the per-chain
    means gives the standard error."""
    from jmmonedmc_amd.engine import Engine
    chains, U_ex = [], None
    for k in range(256):
        ch, U_ex = SMALL[name](2000 + k, k)
        chains.append(ch)
    print(f"[gpu] exact small {name} {form}", flush=True)
    eng = Engine(chains, form=form, sample_capacity=8000, block_sweeps=1000)
    eng.run(2000)
    eng.read_samples(8000)
    eng.run(8000)
    smp = eng.read_samples(8000)
    eng.close()
    Uc = smp["U"].mean(axis=1)
    m, se = Uc.mean(), Uc.std(ddof=1) / math.sqrt(len(Uc))
    assert abs(m - U_ex) <= 3 * se, (m, U_ex, se)
