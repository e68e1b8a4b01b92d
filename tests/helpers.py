"""Shared helpers: chain -> oracle spec, small workload builders."""
import math

import numpy as np

from jmmonedmc_amd import _native as nat
from jmmonedmc_amd.engine import Chain


def oracle_spec(oracle, ch: Chain):
    return oracle.make_spec(**ch.as_dict())


def lj(**kw):
    d = dict(potential=nat.POT_LJ, cutoff=3.0, pot_p=(1.0, 1.0, 0.0, 0.0), shifted=1)
    d.update(kw)
    return d


def hrsw(**kw):
    d = dict(potential=nat.POT_HRSW, cutoff=1.5, pot_p=(1.0, 1.0, 1.5, 0.0), shifted=0)
    d.update(kw)
    return d


def yukawa(**kw):
    d = dict(potential=nat.POT_YUKAWA, cutoff=20.0, pot_p=(1.0, 0.2, 0.5, 0.0), shifted=0)
    d.update(kw)
    return d


# Small cases covering every mode the reference and the configurations use.
# Each returns a list of chains; every test states its boundary, cutoff form
# and volume move through these builders.
CASES = {
    # C1 shape, periodic, shifted LJ (config mode)
    "c1_periodic_nvt": lambda: [Chain(N=64, L0=128.0, T=1.0, seed=1, **lj())],
    # C1 shape with the reference's walls and truncated LJ (parity mode)
    "c1_walls_truncated": lambda: [Chain(N=64, L0=128.0, T=1.0, seed=1, boundary=nat.WALLS,
                                         **lj(shifted=0))],
    # NPT, periodic, ln L moves, shifted LJ (C2 shape, small)
    "npt_periodic_log": lambda: [Chain(N=300, rho=0.5, T=1.0, P=0.5, ensemble=nat.NPT,
                                       vol_move=nat.VOL_LOG, max_dv=0.01, seed=7, chain_id=3,
                                       **lj())],
    # NPT, walls, linear dL moves with N ln s Jacobian (reference fav)
    "npt_walls_linear": lambda: [Chain(N=100, L0=200.0, T=1.0, P=0.5, ensemble=nat.NPT,
                                       boundary=nat.WALLS, vol_move=nat.VOL_LINEAR, max_dv=1.0,
                                       seed=11, **lj(shifted=0))],
    # NPT periodic LJ with the exclusion distance at 0.75 sigma and up to
    # 1024 blocks: 512 blocks of 6, so a 512-lane workgroup (the C2 variant
    # with twice the blocks, smaller)
    "lj_exclusion_npt": lambda: [Chain(N=3100, rho=0.5, T=1.0, P=0.5, ensemble=nat.NPT,
                                       vol_move=nat.VOL_LOG, max_dv=0.01, seed=8, chain_id=4, bcap=1024,
                                       **lj(pot_p=(1.0, 1.0, 0.75, 0.0)))],
    # block counts in whole waves (bcap 448, not a power of two): LJ NPT
    # chains of 4096 (448 blocks), 1400 (192) and 300 (32, below a wave)
    "blocks64_npt": lambda: [Chain(N=n, rho=0.5, T=1.0, P=0.5, ensemble=nat.NPT, vol_move=nat.VOL_LOG,
                                   max_dv=0.01, seed=60 + k, chain_id=60 + k, bcap=448, **lj())
                             for k, n in enumerate([4096, 1400, 300])],
    # hard rod + square well, NPT periodic (C3 second half)
    "hrsw_npt": lambda: [Chain(N=200, rho=0.5, T=1.0, P=1.0, ensemble=nat.NPT, max_dv=0.01,
                               seed=5, **hrsw())],
    # reference test/INPUTstd shape: harmonic, NBN 1, walls, linear volume
    "harmonic_nbn1_walls": lambda: [Chain(N=10, L0=10.0, T=0.4, P=0.7, ensemble=nat.NPT,
                                          boundary=nat.WALLS, vol_move=nat.VOL_LINEAR, max_dv=1.0,
                                          max_step=0.1, nbn=1, seed=125, potential=nat.POT_HARMONIC,
                                          cutoff=math.inf, pot_p=(0, 0, 0, 0), shifted=0)],
    # reference test/INPUT_smalltest shape: LJ with no cutoff, walls, NPT
    "lj_inf_walls_npt": lambda: [Chain(N=10, L0=10.0, T=0.9, P=1.0, ensemble=nat.NPT,
                                       boundary=nat.WALLS, vol_move=nat.VOL_LINEAR, max_dv=0.1,
                                       max_step=0.1, seed=92847, **lj(cutoff=math.inf, shifted=0))],
    # reference scripts/INPUT shape: LJ no cutoff with NBN 2, walls
    "lj_inf_nbn2_walls": lambda: [Chain(N=200, L0=250.0, T=0.8, P=0.7, ensemble=nat.NPT,
                                        boundary=nat.WALLS, vol_move=nat.VOL_LINEAR, max_dv=5.0,
                                        max_step=0.1, nbn=2, seed=92847,
                                        **lj(cutoff=math.inf, shifted=0))],
    # Yukawa + hard core, NPT periodic (C5 shape, small)
    "yukawa_npt": lambda: [Chain(N=200, rho=0.8, T=1.5, P=1.0, ensemble=nat.NPT, max_dv=0.01,
                                 seed=3, **yukawa())],
    # the smallest chains, one bundle: a lone particle (no partners) with
    # walls and periodic, pairs whose two images lie one on each side of a
    # ring, w = N - 1 windows, every potential
    "tiny_chains": lambda: [
        Chain(N=1, L0=4.0, T=1.0, P=0.5, ensemble=nat.NPT, boundary=nat.WALLS,
              vol_move=nat.VOL_LINEAR, max_dv=0.5, seed=21, chain_id=21, **lj(shifted=0)),
        Chain(N=1, L0=10.0, T=1.0, seed=22, chain_id=22, **lj()),
        Chain(N=2, L0=3.0, T=0.7, boundary=nat.WALLS, max_step=0.3, seed=23, chain_id=23, **hrsw()),
        Chain(N=2, L0=9.0, T=1.0, P=0.3, ensemble=nat.NPT, vol_move=nat.VOL_LOG, max_dv=0.05,
              seed=24, chain_id=24, **lj()),
        Chain(N=3, L0=60.0, T=1.5, seed=25, chain_id=25, **yukawa()),
        Chain(N=5, L0=6.0, T=0.4, P=0.7, ensemble=nat.NPT, boundary=nat.WALLS,
              vol_move=nat.VOL_LINEAR, max_dv=0.5, max_step=0.1, seed=26, chain_id=26,
              potential=nat.POT_HARMONIC, cutoff=math.inf, pot_p=(0, 0, 0, 0), shifted=0),
        Chain(N=5, L0=8.0, T=0.9, P=1.0, ensemble=nat.NPT, boundary=nat.WALLS,
              vol_move=nat.VOL_LINEAR, max_dv=0.1, max_step=0.1, seed=27, chain_id=27,
              **lj(cutoff=math.inf, shifted=0)),
    ],
    # a mixed batch: several chains per workgroup, several bundles, recompute
    "mixed_batch": lambda: (
        [Chain(N=n, rho=0.5, T=0.5 + 0.1 * k, P=0.2 + 0.1 * k, ensemble=nat.NPT, seed=k,
               chain_id=k, eci=7, **lj()) for k, n in enumerate([32, 40, 57, 90, 130, 200, 333, 700])]
        + [Chain(N=n, rho=0.5, T=1.0, P=0.5, ensemble=nat.NPT, seed=100 + k, chain_id=100 + k,
                 eci=7, **hrsw()) for k, n in enumerate([32, 64, 100, 1000])]
        + [Chain(N=n, rho=0.5, T=1.0, seed=200 + k, chain_id=200 + k, **lj())
           for k, n in enumerate([50, 77])]),
}


def rel(a, b):
    return abs(a - b) / max(1.0, abs(b))
