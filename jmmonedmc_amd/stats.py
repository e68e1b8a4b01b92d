"""Statistics over per-sweep samples (U, L per chain per sweep).

Python 3 restatements of the reference's post-processing:
  block means and their standard error   scripts/Analyze_A_SD.py:155-209
  fluctuation properties                 scripts/Analyze_Mean.py:162-176
  block-adjacent correlation             scripts/Plot_AutoCorrelation.py:15-29
  lag correlation                        scripts/Plot_AutoCorrelation.py:36-45
plus the integrated autocorrelation time (Sokal's self-consistent window),
which the reference does not compute.  All functions take numpy arrays;
none touches the GPU.
"""
from __future__ import annotations

import math

import numpy as np


def block_means(x, block: int):
    """Means of consecutive blocks of `block` samples; a trailing partial
    block is dropped (floor(n / block) blocks, Analyze_A_SD.py:157-166)."""
    x = np.asarray(x, dtype=np.float64)
    if block < 1:
        raise ValueError("block size must be >= 1")
    nb = x.shape[0] // block
    return x[: nb * block].reshape((nb, block) + x.shape[1:]).mean(axis=1)


def block_sem(x, block: int):
    """Standard error of the mean from block means: population standard
    deviation of the block means / sqrt(number of blocks)
    (Analyze_A_SD.py:186-190)."""
    m = block_means(x, block)
    if m.shape[0] < 2:
        return float("nan")
    return m.std(axis=0) / math.sqrt(m.shape[0])


def block_sem_scan(x, block_min: int, block_start: int | None = None):
    """The reference's blocking table (Analyze_A_SD.py:147-206): rows of
    (block size, number of blocks, SEM) for block sizes block_start,
    block_min, 2 block_min, ... up to n/2.  With block_start None the scan
    starts at block_min."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    rows = []
    sizes = [block_start] if block_start else []
    b = block_min
    while b <= n // 2:
        sizes.append(b)
        b += block_min
    for b in sizes:
        if b > n // 2:
            continue
        rows.append((b, n // b, float(block_sem(x, b))))
    return np.array(rows, dtype=np.float64).reshape(-1, 3)


def block_adjacent_correlation(x, max_block: int | None = None):
    """c[b] = correlation coefficient of adjacent block means for block size
    b = 1 .. n/2, c[0] = 1 (Plot_AutoCorrelation.py:15-29).  Entries with
    fewer than three blocks are NaN."""
    x = np.asarray(x, dtype=np.float64).ravel()
    top = len(x) // 2 if max_block is None else min(max_block, len(x) // 2)
    c = np.full(top + 1, np.nan)
    c[0] = 1.0
    for b in range(1, top + 1):
        m = block_means(x, b)
        if len(m) >= 3:
            c[b] = np.corrcoef(m[:-1], m[1:])[0, 1]
    return c


def lag_correlation(x, max_lag: int):
    """c[d] = correlation coefficient of x[t] and x[t+d] over the sample
    (Plot_AutoCorrelation.py:36-45), d = 0 .. max_lag."""
    x = np.asarray(x, dtype=np.float64).ravel()
    c = np.empty(max_lag + 1)
    c[0] = 1.0
    for d in range(1, max_lag + 1):
        c[d] = np.corrcoef(x[:-d], x[d:])[0, 1]
    return c


def autocorrelation(x, max_lag: int | None = None):
    """Normalised autocorrelation function rho(t) = C(t)/C(0) with the
    biased (1/n) estimator, computed by FFT."""
    x = np.asarray(x, dtype=np.float64).ravel()
    n = len(x)
    y = x - x.mean()
    f = np.fft.rfft(y, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n] / n
    if acf[0] == 0:
        return np.ones(1 if max_lag is None else max_lag + 1)
    acf /= acf[0]
    return acf if max_lag is None else acf[: max_lag + 1]


def tau_int(x, c: float = 6.0):
    """Integrated autocorrelation time tau = 1/2 + sum_{t>=1} rho(t), summed
    up to the smallest window W with W >= c tau(W) (Sokal).  The variance of
    the mean is 2 tau Var(x)/n."""
    rho = autocorrelation(x)
    tau = 0.5
    for W in range(1, len(rho)):
        tau += rho[W]
        if W >= c * tau:
            return float(tau)
    return float(tau)


def fluctuation_properties(E, E2, L, L2, LE, P: float, T: float):
    """The reference's fluctuation formulas (Analyze_Mean.py:164-176), applied
    elementwise to averages <E>, <E^2>, <L>, <L^2>, <L E> (per block or
    overall), k_B = 1:
        cP     = (var E + 2 P cov(L,E) + P^2 var L) / T^2
        betaT  = var L / (<L> T)
        alphaP = (cov(L,E) + P var L) / (T^2 <L>)
        gammaV = alphaP / betaT
        betaS  = betaT - alphaP^2 T <L> / cP
        muJT   = <L> (alphaP T - 1) / cP
    E is the configurational energy, so cP is the configurational heat
    capacity at constant pressure."""
    E, E2, L, L2, LE = (np.asarray(a, dtype=np.float64) for a in (E, E2, L, L2, LE))
    varE = E2 - E * E
    varL = L2 - L * L
    cov = LE - L * E
    cp = (varE + 2.0 * P * cov + P * P * varL) / (T * T)
    betaT = varL / (L * T)
    alphaP = (cov + P * varL) / (T * T * L)
    gammaV = alphaP / betaT
    betaS = betaT - alphaP * alphaP * T * L / cp
    muJT = L / cp * (alphaP * T - 1.0)
    return dict(cP=cp, betaT=betaT, alphaP=alphaP, gammaV=gammaV, betaS=betaS, muJT=muJT)


def fluctuations_from_samples(U, L, P: float, T: float, nblocks: int = 20):
    """Per-sweep samples of one chain -> fluctuation properties computed per
    block (as the reference does) and summarised as mean and block SEM.
    Returns {name: (mean, sem)}."""
    U = np.asarray(U, dtype=np.float64)
    L = np.asarray(L, dtype=np.float64)
    b = len(U) // nblocks
    if b < 2:
        raise ValueError("need at least 2 samples per block")
    props = fluctuation_properties(block_means(U, b), block_means(U * U, b), block_means(L, b),
                                   block_means(L * L, b), block_means(L * U, b), P, T)
    out = {}
    for k, v in props.items():
        out[k] = (float(v.mean()), float(v.std() / math.sqrt(len(v))))
    return out


def mean_sem(x, nblocks: int = 20):
    """Mean and block SEM with a fixed number of blocks."""
    x = np.asarray(x, dtype=np.float64)
    b = max(len(x) // nblocks, 1)
    return float(x[: b * nblocks].mean()), float(block_sem(x, b))
