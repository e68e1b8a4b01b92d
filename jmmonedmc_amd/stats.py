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


def tau_int_window(x, c: float = 6.0):
    """(tau, W): the integrated autocorrelation time tau = 1/2 +
    sum_{t=1..W} rho(t) and the smallest window W with W >= c tau(W)
    (Sokal's self-consistent window; the whole series if none)."""
    rho = autocorrelation(x)
    if len(rho) < 2:
        return 0.5, 0
    tau = 0.5 + np.cumsum(rho[1:])
    W = np.arange(1, len(rho))
    ok = np.flatnonzero(W >= c * tau)
    k = ok[0] if ok.size else len(tau) - 1
    return float(tau[k]), int(W[k])


def tau_int(x, c: float = 6.0):
    """Integrated autocorrelation time tau = 1/2 + sum_{t>=1} rho(t), summed
    up to the smallest window W with W >= c tau(W) (Sokal).  The variance of
    the mean is 2 tau Var(x)/n."""
    return tau_int_window(x, c)[0]


def mean_se_tau(x, c: float = 6.0):
    """Mean of a correlated series with its standard error
    sqrt(2 tau Var(x) / n), the integrated autocorrelation time tau, the
    number of independent samples n / (2 tau), and the standard error of
    tau itself, tau sqrt(2 (2W + 1) / n) (Sokal)."""
    x = np.asarray(x, dtype=np.float64).ravel()
    n = len(x)
    tau, W = tau_int_window(x, c)
    tau = max(tau, 0.5)
    se = math.sqrt(2.0 * tau * x.var() / n) if n > 1 else float("nan")
    return dict(mean=float(x.mean()), se=se, tau=tau, n_eff=n / (2.0 * tau),
                tau_se=tau * math.sqrt(2.0 * (2 * W + 1) / n) if n else float("nan"))


OBSERVABLES = ("U", "L", "rho", "cP", "betaT", "tau_U")


def chain_observables(U, L, N: int, P: float, T: float, nblocks: int = 20):
    """Per-chain estimates with standard errors from one run's per-sweep
    samples: <U>, <L>, <rho> = <N/L> (tau-corrected SEs), the fluctuation
    properties cP and betaT (block formulas, block SEM), and tau_int of U
    (Sokal SE).  Each entry is (value, se, n_eff); n_eff is the number of
    independent samples of that estimate (for cP and betaT, the smaller of
    the blocks' length over 2 tau of U and of L, times the block count)."""
    U = np.asarray(U, dtype=np.float64)
    L = np.asarray(L, dtype=np.float64)
    out = {}
    mu = mean_se_tau(U)
    ml = mean_se_tau(L)
    mr = mean_se_tau(N / L)
    out["U"] = (mu["mean"], mu["se"], mu["n_eff"])
    out["L"] = (ml["mean"], ml["se"], ml["n_eff"])
    out["rho"] = (mr["mean"], mr["se"], mr["n_eff"])
    b = len(U) // nblocks
    if b >= 2:
        f = fluctuations_from_samples(U, L, P, T, nblocks=nblocks)
        blk = min(b / (2 * mu["tau"]), b / (2 * ml["tau"])) * nblocks
        out["cP"] = (f["cP"][0], f["cP"][1], blk)
        out["betaT"] = (f["betaT"][0], f["betaT"][1], blk)
    out["tau_U"] = (mu["tau"], mu["tau_se"], mu["n_eff"])
    return out


def _acf_rows(X):
    """Normalised autocorrelation of every row (biased estimator, FFT)."""
    n = X.shape[1]
    Y = X - X.mean(axis=1, keepdims=True)
    F = np.fft.rfft(Y, 2 * n, axis=1)
    A = np.fft.irfft(F * np.conj(F), axis=1)[:, :n] / n
    c0 = A[:, :1].copy()
    c0[c0 == 0] = 1.0
    return A / c0


def _tau_rows(X, c: float = 6.0):
    """tau_int and Sokal's window per row (as tau_int_window)."""
    rho = _acf_rows(X)
    n = X.shape[1]
    tau = 0.5 + np.cumsum(rho[:, 1:], axis=1)
    W = np.arange(1, n)
    ok = W[None, :] >= c * tau
    k = np.where(ok.any(axis=1), ok.argmax(axis=1), n - 2)
    rows = np.arange(X.shape[0])
    return np.maximum(tau[rows, k], 0.5), W[k]


def chain_observables_batch(U, L, N, P, T, nblocks: int = 20, chunk: int = 256):
    """chain_observables for many chains at once: U, L [chain, sweep];
    N, P, T per chain.  Returns {observable: (value[c], se[c], n_eff[c])}
    with the same definitions."""
    U = np.asarray(U, dtype=np.float64)
    L = np.asarray(L, dtype=np.float64)
    nc, n = U.shape
    N = np.broadcast_to(np.asarray(N, dtype=np.float64), (nc,))
    P = np.broadcast_to(np.asarray(P, dtype=np.float64), (nc,))
    T = np.broadcast_to(np.asarray(T, dtype=np.float64), (nc,))
    out = {k: (np.full(nc, np.nan), np.full(nc, np.nan), np.zeros(nc)) for k in OBSERVABLES}
    b = n // nblocks
    for lo in range(0, nc, chunk):
        sl = slice(lo, min(nc, lo + chunk))
        u, l = U[sl], L[sl]
        r = N[sl, None] / l
        taus = {}
        for k, X in (("U", u), ("L", l), ("rho", r)):
            tau, W = _tau_rows(X)
            taus[k] = (tau, W)
            out[k][0][sl] = X.mean(axis=1)
            out[k][1][sl] = np.sqrt(2.0 * tau * X.var(axis=1) / n)
            out[k][2][sl] = n / (2.0 * tau)
        tu, wu = taus["U"]
        out["tau_U"][0][sl] = tu
        out["tau_U"][1][sl] = tu * np.sqrt(2.0 * (2 * wu + 1) / n)
        out["tau_U"][2][sl] = n / (2.0 * tu)
        if b >= 2:
            def bm(x):
                return x[:, : b * nblocks].reshape(x.shape[0], nblocks, b).mean(axis=2)
            f = fluctuation_properties(bm(u), bm(u * u), bm(l), bm(l * l), bm(l * u),
                                       P[sl, None], T[sl, None])
            blk = np.minimum(b / (2 * tu), b / (2 * taus["L"][0])) * nblocks
            for k in ("cP", "betaT"):
                out[k][0][sl] = f[k].mean(axis=1)
                out[k][1][sl] = f[k].std(axis=1) / math.sqrt(nblocks)
                out[k][2][sl] = blk
    return out


def _as_batch(r):
    if isinstance(r, dict):
        return r
    nc = len(r)
    out = {}
    for k in OBSERVABLES:
        v = np.full(nc, np.nan)
        se = np.full(nc, np.nan)
        ne = np.zeros(nc)
        for c, d in enumerate(r):
            if k in d:
                v[c], se[c], ne[c] = d[k]
        out[k] = (v, se, ne)
    return out


def compare_runs(a, b, n_eff_min: float = 30.0):
    """z-scores of run b against run a per observable and chain:
    z = (b - a) / sqrt(se_a^2 + se_b^2), over the chains where both
    estimates rest on at least n_eff_min independent samples.  a, b: one
    entry per chain of chain_observables() dicts, or chain_observables_batch
    results.  Returns {observable: summary} with the compared and excluded
    chain counts, the fraction within 3 standard errors, the mean, standard
    deviation and largest |z|, and the median and root-mean-square relative
    difference |b - a| / |a|."""
    A, B = _as_batch(a), _as_batch(b)
    res = {}
    for k in OBSERVABLES:
        va, sa, na = A[k]
        vb, sb, nb = B[k]
        se = np.hypot(sa, sb)
        ok = (np.minimum(na, nb) >= n_eff_min) & (se > 0) & np.isfinite(se) & np.isfinite(vb - va)
        z = (vb[ok] - va[ok]) / se[ok]
        nz = ok & (va != 0)
        rd = np.abs(vb[nz] - va[nz]) / np.abs(va[nz]) if nz.any() else np.zeros(1)
        res[k] = dict(chains_compared=int(z.size), chains_excluded=int((~ok).sum()),
                      frac_within_3sigma=float((np.abs(z) <= 3).mean()) if z.size else None,
                      mean_z=float(z.mean()) if z.size else None,
                      std_z=float(z.std()) if z.size else None,
                      max_abs_z=float(np.abs(z).max()) if z.size else None,
                      median_rel_diff=float(np.median(rd)), rms_rel_diff=float(np.sqrt((rd ** 2).mean())))
    return res


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


def _tau_rows_fast(X, c: float = 6.0, chunk: int = 512, workers: int = 8, lags: int = 0):
    """_tau_rows's tau for many rows, FFTs on `workers` threads.  lags > 0:
    also each row's window (the column k of its running sum, lag k + 1) and
    the running sums 0.5 + sum_{l <= j + 1} rho(l) for the first `lags`
    columns (float32), from which tau_at_windows reads tau at any window."""
    import scipy.fft as sfft
    X = np.asarray(X, dtype=np.float64)
    nc, n = X.shape
    out = np.empty(nc)
    m = sfft.next_fast_len(2 * n, real=True)
    W = np.arange(1, n)
    ncol = max(1, min(lags, n - 1))
    wins = np.empty(nc, dtype=np.int64)
    cums = np.empty((nc, ncol), dtype=np.float32) if lags else None
    for lo in range(0, nc, chunk):
        Y = X[lo:lo + chunk]
        Y = Y - Y.mean(axis=1, keepdims=True)
        F = sfft.rfft(Y, m, axis=1, workers=workers)
        A = sfft.irfft(F * np.conj(F), m, axis=1, workers=workers)[:, :n]
        c0 = A[:, :1].copy()
        c0[c0 == 0] = 1.0
        cum = 0.5 + np.cumsum((A / c0)[:, 1:], axis=1)
        ok = W[None, :] >= c * cum
        k = np.where(ok.any(axis=1), ok.argmax(axis=1), n - 2)
        out[lo:lo + chunk] = np.maximum(cum[np.arange(Y.shape[0]), k], 0.5)
        wins[lo:lo + chunk] = k
        if lags:
            cums[lo:lo + chunk] = cum[:, :ncol]
    return (out, wins, cums) if lags else out


def tau_at_windows(cums, windows):
    """tau_int of every row at its given window column (clamped to the
    stored columns), at least 1/2: the running sum _tau_rows_fast keeps."""
    cums = np.asarray(cums)
    k = np.minimum(np.asarray(windows, dtype=np.int64), cums.shape[1] - 1)
    return np.maximum(cums[np.arange(cums.shape[0]), k].astype(np.float64), 0.5)


def common_window_tau(ref_runs, runs):
    """tau_int of U at one window per chain for every run: the median over
    the reference runs of their own Sokal windows (rounded down).  A run's
    own window moves with its noise (a slow fluctuation lengthens it and
    the estimate with it, which gives the per-run estimates a heavy upper
    tail); at a common window every run's estimate is the same sum of
    autocorrelations, and the window does not depend on the runs compared.
    ref_runs, runs: run_values() results made with acf_lags > 0; each run's
    "tau_U" is replaced.  Returns the windows (as lags, column + 1)."""
    Kw = np.stack([np.asarray(r["_tau_window"]) for r in ref_runs])
    w = np.floor(np.median(Kw, axis=0)).astype(np.int64)
    for r in runs:
        r["tau_U"] = tau_at_windows(r["_tau_cum"], w)
    return w + 1


def run_values(U, L, N, P, T, nblocks: int = 20, workers: int = 8, acf_lags: int = 0):
    """One run's per-chain values of every observable in OBSERVABLES, from
    per-sweep samples U, L [chain, sweep]: <U>, <L>, <rho> = <N/L>; cP and
    betaT as chain_observables defines them (the formulas of
    scripts/Analyze_Mean.py:164-176 per block of n/nblocks sweeps, averaged
    over the blocks); tau_int of U (Sokal's window, c = 6).  Returns
    {observable: array[chain]}; with acf_lags > 0 also "_tau_window" and
    "_tau_cum" (the window and running sums, for common_window_tau)."""
    U = np.asarray(U, dtype=np.float64)
    L = np.asarray(L, dtype=np.float64)
    nc, n = U.shape
    N = np.broadcast_to(np.asarray(N, dtype=np.float64), (nc,))
    P = np.broadcast_to(np.asarray(P, dtype=np.float64), (nc,))
    T = np.broadcast_to(np.asarray(T, dtype=np.float64), (nc,))
    out = {"U": U.mean(axis=1), "L": L.mean(axis=1), "rho": (N[:, None] / L).mean(axis=1)}
    b = n // nblocks
    if b >= 2:
        def bm(x):
            return x[:, : b * nblocks].reshape(nc, nblocks, b).mean(axis=2)
        f = fluctuation_properties(bm(U), bm(U * U), bm(L), bm(L * L), bm(L * U), P[:, None], T[:, None])
        out["cP"] = f["cP"].mean(axis=1)
        out["betaT"] = f["betaT"].mean(axis=1)
    else:
        out["cP"] = np.full(nc, np.nan)
        out["betaT"] = np.full(nc, np.nan)
    if acf_lags:
        out["tau_U"], out["_tau_window"], out["_tau_cum"] = _tau_rows_fast(U, workers=workers, lags=acf_lags)
    else:
        out["tau_U"] = _tau_rows_fast(U, workers=workers)
    return out


def replica_band(K: int, q: float = 0.9973) -> float:
    """Half-width of the two-sided probability-q band of Student's t with
    K - 1 degrees of freedom (q = 0.9973: the probability of a normal
    variable's 3 sigma band)."""
    from scipy import stats as st
    return float(st.t.ppf(0.5 + q / 2.0, K - 1))


def band_allowance(n: int, q: float = 0.9973, level: float = 0.99) -> int:
    """How many of n chains may lie outside a probability-q band before a
    comparison fails: 1 % of them, or, where that is less, the `level`
    quantile of the binomial count outside (n, 1 - q) (with 64 chains one
    outside is likely under the hypothesis, 1 % of them is none)."""
    from scipy import stats as st
    return int(max(math.floor(0.01 * n), st.binom.ppf(level, n, 1.0 - q)))


# compared on a log scale: an estimate of tau_int from one run is skewed
# (a run whose window takes in a slow fluctuation reads several times the
# typical value), its logarithm is close to normal
LOG_COMPARED = ("tau_U",)


def replica_compare(ref, test, q: float = 0.9973, groups=None):
    """One run's per-chain values against K independent reference runs of
    the same chains from the same starting configuration.

    ref: K >= 3 run_values() results; test: one.  Per chain and observable
    t = (x - m) / (s sqrt(1 + 1/K)), m and s the mean and standard deviation
    over the reference runs.  If the test run is one more draw from the
    distribution the reference runs sample, and a run's value is normal, t
    follows Student's t with K - 1 degrees of freedom: the standard error
    comes from the runs themselves, not from a model of the
    autocorrelation.  Observables in LOG_COMPARED are compared through
    their logarithms (then the relative differences below are of the
    logarithms).  Returns {observable: summary}: the chains compared
    (finite values, s > 0; chains whose reference runs agree to 1e-12
    relative are counted as constant and not compared), the fraction with
    |t| within the probability-q
    band (`band`), the number outside it and how many band_allowance
    permits, mean t with its standard error (a bias of the test run
    shows as a mean t away from 0), the standard deviation of t (expected
    sqrt((K-1)/(K-3)) under the hypothesis), the largest |t|, the median
    |x - m| / s (the difference in units of one run's spread) and the median
    and RMS of |x - m| / |m|.  groups: optional {name: bool mask[chain]},
    the fraction and mean t per group as well."""
    K = len(ref)
    if K < 3:
        raise ValueError("need at least 3 reference runs")
    band = replica_band(K, q)
    res = {}
    for k in OBSERVABLES:
        R = np.stack([np.asarray(r[k], dtype=np.float64) for r in ref])
        x = np.asarray(test[k], dtype=np.float64)
        if k in LOG_COMPARED:
            with np.errstate(divide="ignore", invalid="ignore"):
                R, x = np.log(R), np.log(x)
        m = R.mean(axis=0)
        sd = R.std(axis=0, ddof=1)
        # an observable that does not fluctuate in the ensemble (L, rho and
        # betaT at constant length) differs between runs only by the
        # rounding of its mean: constant, not compared
        const = np.isfinite(sd) & (sd <= 1e-12 * np.abs(m))
        ok = np.isfinite(x) & np.isfinite(m) & np.isfinite(sd) & (sd > 0) & ~const
        t = (x[ok] - m[ok]) / (sd[ok] * math.sqrt(1.0 + 1.0 / K))
        nz = ok & (m != 0)
        rd = np.abs(x[nz] - m[nz]) / np.abs(m[nz]) if nz.any() else np.zeros(1)
        r = dict(chains_compared=int(t.size), chains_excluded=int((~ok).sum()),
                 chains_constant=int(const.sum()), band=band,
                 frac_within_band=float((np.abs(t) <= band).mean()) if t.size else None,
                 outside=int((np.abs(t) > band).sum()), allowed_outside=band_allowance(t.size, q),
                 mean_t=float(t.mean()) if t.size else None,
                 mean_t_se=float(t.std() / math.sqrt(t.size)) if t.size > 1 else None,
                 std_t=float(t.std()) if t.size else None,
                 max_abs_t=float(np.abs(t).max()) if t.size else None,
                 median_diff_over_sd=float(np.median(np.abs(x[ok] - m[ok]) / sd[ok])) if t.size else None,
                 median_rel_diff=float(np.median(rd)), rms_rel_diff=float(np.sqrt((rd ** 2).mean())))
        if groups:
            tt = np.full(len(x), np.nan)
            tt[ok] = t
            for g, mask in groups.items():
                sel = np.asarray(mask, bool) & ok
                if sel.any():
                    r[g] = dict(chains=int(sel.sum()), frac_within_band=float((np.abs(tt[sel]) <= band).mean()),
                                mean_t=float(tt[sel].mean()))
        res[k] = r
    return res


def mean_sem(x, nblocks: int = 20):
    """Mean and block SEM with a fixed number of blocks."""
    x = np.asarray(x, dtype=np.float64)
    b = max(len(x) // nblocks, 1)
    return float(x[: b * nblocks].mean()), float(block_sem(x, b))
