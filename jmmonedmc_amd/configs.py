"""The five workloads the engine is measured and checked on, as chain lists.

C1  1 chain, N=64, L=128, NVT, shifted LJ (rc 3), T=1, step 0.5, seed 1
C2  256 chains (seeds 0..255), N=4096, NPT T=1 P=0.5, ln L step 0.01,
    density 0.5 start
C3  4096 chains, N log-uniform in [32, 8192], T in [0.5, 2], P in [0.1, 2],
    chains 0..2047 shifted LJ, 2048..4095 hard rod (1) + square well
    (depth 1, edge 1.5), density 0.5 start, FP64 recompute every 100 sweeps
    (parameters drawn with numpy's PCG64 at seed C3_PARAM_SEED)
C4  1024 chains, N=1024, density 0.6, NVT T=1, shifted LJ theta=(1, 1)
C5  64 chains, N=65536, density 0.8, NPT T=1.5 P=1, Yukawa A=1 kappa=0.2
    rc 20 + hard core 0.5
"""
from __future__ import annotations

import numpy as np

from . import _native as nat
from .engine import Chain

C3_PARAM_SEED = 20241202

LJ = dict(potential=nat.POT_LJ, cutoff=3.0, pot_p=(1.0, 1.0, 0.0, 0.0), shifted=1)
HRSW = dict(potential=nat.POT_HRSW, cutoff=1.5, pot_p=(1.0, 1.0, 1.5, 0.0), shifted=0)
YUKAWA = dict(potential=nat.POT_YUKAWA, cutoff=20.0, pot_p=(1.0, 0.2, 0.5, 0.0), shifted=0)


def c1(boundary=nat.PERIODIC, seed=1):
    return [Chain(N=64, L0=128.0, T=1.0, ensemble=nat.NLT, boundary=boundary, seed=seed,
                  max_step=0.5, **LJ)]


def c2(n_chains=256, N=4096, exclusion=0.0, bcap=1024):
    """256 NPT LJ chains of 4096.  bcap: the largest block count (1024:
    4096 / M_min = 4096 / 7 = 585 rounds down to 512 blocks per chain, the
    512-lane kernel; 0 caps at 256).  exclusion > 0: the LJ exclusion
    distance (a trial closer than it is rejected; the default is 0.5 sigma);
    at 0.75 sigma the window is 4 partners instead of 6, and trials landing
    closer than 0.75 sigma, whose pair energy is at least phi(0.75) =
    103.8 eps (Boltzmann factor 8e-46 at T = 1), are rejected without
    evaluation."""
    lj = dict(LJ, pot_p=(LJ["pot_p"][0], LJ["pot_p"][1], exclusion, 0.0))
    return [Chain(N=N, rho=0.5, T=1.0, P=0.5, ensemble=nat.NPT, vol_move=nat.VOL_LOG, max_dv=0.01,
                  max_step=0.5, seed=s, chain_id=s, bcap=bcap, **lj) for s in range(n_chains)]


def c3_params(n_chains=4096, seed=C3_PARAM_SEED):
    rng = np.random.default_rng(seed)
    N = np.rint(np.exp(rng.uniform(np.log(32), np.log(8192), n_chains))).astype(np.int64)
    T = rng.uniform(0.5, 2.0, n_chains)
    P = rng.uniform(0.1, 2.0, n_chains)
    return N, T, P


def c3(n_chains=4096, eci=100, scale_N=1.0):
    N, T, P = c3_params(n_chains)
    out = []
    for c in range(n_chains):
        pot = LJ if c < n_chains // 2 else HRSW
        n = max(8, int(round(N[c] * scale_N)))
        out.append(Chain(N=n, rho=0.5, T=float(T[c]), P=float(P[c]), ensemble=nat.NPT,
                         vol_move=nat.VOL_LOG, max_dv=0.01, max_step=0.5, seed=c, chain_id=c,
                         eci=eci, **pot))
    return out


def c4(n_chains=1024, N=1024, eps=1.0, sigma=1.0):
    return [Chain(N=N, rho=0.6, T=1.0, ensemble=nat.NLT, max_step=0.5, seed=c, chain_id=c,
                  potential=nat.POT_LJ, cutoff=3.0, pot_p=(eps, sigma, 0.0, 0.0), shifted=1)
            for c in range(n_chains)]


def c5(n_chains=64, N=65536, bcap=1024):
    """64 Yukawa chains of 65536 particles.  The window (41 partners per side)
    needs blocks of at least 43 particles; bcap = 1024 gives 1024 blocks of 64
    (64 phases per sweep) where the default 256 would give 256 phases."""
    return [Chain(N=N, rho=0.8, T=1.5, P=1.0, ensemble=nat.NPT, vol_move=nat.VOL_LOG, max_dv=0.01,
                  max_step=0.5, seed=c, chain_id=c, eci=100, bcap=bcap, **YUKAWA)
            for c in range(n_chains)]


def flop_per_pair(potential: int) -> float:
    """Algorithmic FLOPs per pair evaluation (add/mul/compare = 1, FMA = 2,
    rcp/div/sqrt/exp = 1): LJ 10, hard rod + well 4, Yukawa 10, harmonic 4."""
    return {nat.POT_LJ: 10.0, nat.POT_HRSW: 4.0, nat.POT_YUKAWA: 10.0, nat.POT_HARMONIC: 4.0}[potential]


def trans_per_pair(potential: int) -> float:
    """Quarter-rate transcendental instructions per FP32 pair evaluation:
    LJ one reciprocal, Yukawa an exponential and a reciprocal, none for the
    hard rod + well and the harmonic spring."""
    return {nat.POT_LJ: 1.0, nat.POT_HRSW: 0.0, nat.POT_YUKAWA: 2.0, nat.POT_HARMONIC: 0.0}[potential]
