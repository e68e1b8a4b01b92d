"""Statistics layer (jmmonedmc_amd/stats.py), pinned without retyping its
formulas: hand-computed block tables, AR(1) closed forms, and the
fluctuation formulas against thermodynamic derivatives of the exact 1D
NPT partition function (finite differences of quadratures)."""
import math

import numpy as np
import pytest

from jmmonedmc_amd import stats
import exact1d


def test_block_sem_hand_table():
    x = np.arange(1.0, 10.0)                  # 9 samples; block 2 drops the 9th
    assert np.allclose(stats.block_means(x, 2), [1.5, 3.5, 5.5, 7.5])
    # means 1.5..7.5: population std sqrt(5), 4 blocks -> sqrt(5)/2
    assert stats.block_sem(x, 2) == pytest.approx(math.sqrt(5.0) / 2.0, rel=1e-15)
    tab = stats.block_sem_scan(x, 2)
    assert tab[:, 0].tolist() == [2.0, 4.0] and tab[:, 1].tolist() == [4.0, 2.0]
    # block 4: means 2.5, 6.5 -> std 2, sem 2/sqrt(2)
    assert tab[1, 2] == pytest.approx(math.sqrt(2.0), rel=1e-15)
    assert math.isnan(stats.block_sem(x, 9))


def _ar1(a, n, seed):
    rng = np.random.default_rng(seed)
    e = rng.standard_normal(n)
    x = np.empty(n)
    x[0] = e[0] / math.sqrt(1 - a * a)
    for t in range(1, n):
        x[t] = a * x[t - 1] + e[t]
    return x


def test_ar1_closed_forms():
    """AR(1) x_t = a x_{t-1} + e_t: rho(t) = a^t, tau_int = (1+a)/(2(1-a)),
    adjacent-sample correlation = a, and the large-block SEM approaches
    sqrt(2 tau Var x / n)."""
    a, n = 0.8, 400_000
    x = _ar1(a, n, 7)
    rho = stats.autocorrelation(x, 5)
    assert np.allclose(rho, a ** np.arange(6), atol=0.01)
    assert stats.lag_correlation(x, 3) == pytest.approx(a ** np.arange(4), abs=0.01)
    assert stats.block_adjacent_correlation(x, 2)[1] == pytest.approx(a, abs=0.01)
    tau = stats.tau_int(x)
    assert tau == pytest.approx((1 + a) / (2 * (1 - a)), rel=0.05)
    want = math.sqrt(2 * tau * x.var() / n)
    assert stats.block_sem(x, 2000) == pytest.approx(want, rel=0.2)
    iid = np.random.default_rng(3).standard_normal(n)
    assert abs(stats.tau_int(iid) - 0.5) < 0.05
    assert abs(stats.block_adjacent_correlation(iid, 4)[4]) < 0.02


def _walls_moments(phi, T, P, N, gmin):
    """Exact NPT moments of the walled 1D nearest-neighbour chain: two wall
    gaps of law exp(-P g/T), N-1 interior gaps of law exp(-(phi+P g)/T)."""
    mom = exact1d.gap_moments(phi, 1.0 / T, P, gmin)
    a = T / P                                  # wall gap mean; its variance is a^2
    L = 2 * a + (N - 1) * mom["m"]
    E = (N - 1) * mom["ephi"]
    varL = 2 * a * a + (N - 1) * mom["v"]
    cov = (N - 1) * (mom["ephig"] - mom["m"] * mom["ephi"])
    varE = (N - 1) * (mom["ephi2"] - mom["ephi"] ** 2)
    lnD = 2 * math.log(a) + (N - 1) * mom["lnZ"]
    return dict(L=L, E=E, L2=varL + L * L, E2=varE + E * E, LE=cov + L * E, lnD=lnD)


@pytest.mark.parametrize("name,phi,gmin", [
    ("harmonic", exact1d.harmonic_phi(), 1e-12),
    ("lj", exact1d.lj_phi(), 0.5),
])
def test_fluctuation_formulas_match_thermodynamic_derivatives(name, phi, gmin):
    """cP = d<H>/dT|P, betaT = -(1/L) dL/dP|T, alphaP = (1/L) dL/dT|P,
    gammaV = -(dL/dT)/(dL/dP), muJT = -(dH/dP|T)/(dH/dT|P) and
    betaS = -(1/L)(dL/dP|T - dL/dT|P dS/dP|T / dS/dT|P), with H = E + P L and
    S = (<H> - G)/T, G = -T ln Delta(T, P), all by central differences."""
    N, T, P = 7, 0.9, 0.6
    m = _walls_moments(phi, T, P, N, gmin)
    f = stats.fluctuation_properties(m["E"], m["E2"], m["L"], m["L2"], m["LE"], P, T)

    def H(T_, P_):
        q = _walls_moments(phi, T_, P_, N, gmin)
        return q["E"] + P_ * q["L"]

    def S(T_, P_):
        q = _walls_moments(phi, T_, P_, N, gmin)
        return (q["E"] + P_ * q["L"] + T_ * q["lnD"]) / T_

    def Lof(T_, P_):
        return _walls_moments(phi, T_, P_, N, gmin)["L"]

    h = 1e-4
    dLdT = (Lof(T + h, P) - Lof(T - h, P)) / (2 * h)
    dLdP = (Lof(T, P + h) - Lof(T, P - h)) / (2 * h)
    dHdT = (H(T + h, P) - H(T - h, P)) / (2 * h)
    dHdP = (H(T, P + h) - H(T, P - h)) / (2 * h)
    dSdT = (S(T + h, P) - S(T - h, P)) / (2 * h)
    dSdP = (S(T, P + h) - S(T, P - h)) / (2 * h)
    L = m["L"]
    want = dict(cP=dHdT, betaT=-dLdP / L, alphaP=dLdT / L, gammaV=-dLdT / dLdP,
                muJT=-dHdP / dHdT, betaS=-(dLdP - dLdT * dSdP / dSdT) / L)
    for k, v in want.items():
        assert float(f[k]) == pytest.approx(v, rel=1e-6), (name, k, float(f[k]), v)
    # Maxwell relation dS/dP|T = -dL/dT|P holds for the quadrature itself
    assert dSdP == pytest.approx(-dLdT, rel=1e-6)


def test_fluctuations_from_samples_ideal_gas():
    """Walled ideal gas (no interactions) at NPT: L is Gamma(N+1, T/P), E = 0.
    Then cP = N+1, betaT = 1/P, alphaP = 1/T (exact), from per-sweep samples."""
    N, T, P = 20, 1.3, 0.7
    rng = np.random.default_rng(11)
    L = rng.gamma(N + 1, T / P, size=200_000)
    U = np.zeros_like(L)
    out = stats.fluctuations_from_samples(U, L, P, T, nblocks=20)
    for k, v in dict(cP=N + 1.0, betaT=1 / P, alphaP=1 / T).items():
        mean, sem = out[k]
        assert abs(mean - v) < 4 * sem + 1e-3 * abs(v), (k, mean, sem, v)


def test_fluctuation_gradient_equals_reweighted_derivative(oracle):
    """d<A>/dtheta from the fluctuation formula equals the theta-derivative of
    the reweighted average sum A_theta' e^{-beta (U_theta' - U_theta)} / sum
    e^{...} on the same samples (central differences), for A = W."""
    import dataclasses
    from jmmonedmc_amd.autograd import fluctuation_gradient
    from jmmonedmc_amd.engine import Chain
    from helpers import lj, oracle_spec
    N, L0, beta = 32, 48.0, 0.8
    rng = np.random.default_rng(2)
    X = []
    for _ in range(40):
        sp = np.diff(np.concatenate([[0.0], np.sort(rng.random(N - 1)), [1.0]]))
        g = 0.85 + (L0 - 0.85 * N) * sp
        X.append(np.cumsum(g) - g[0] - L0 / 2 + 0.5 * g[0])
    base = Chain(N=N, L0=L0, T=1 / beta, **lj())

    def obs(theta):
        ch = dataclasses.replace(base, pot_p=(theta[0], theta[1], 0.0, 0.0))
        s = oracle_spec(oracle, ch)
        return np.array([oracle.observables_x(s, x, L0, want_grad=False)[0] for x in X])

    o0 = obs([1.0, 1.0])
    fg = fluctuation_gradient(o0[:, 1], o0[:, 4:6], o0[:, 2:4], beta)
    h = 1e-5
    for k in range(2):
        v = []
        for sg in (1, -1):
            th = [1.0, 1.0]
            th[k] += sg * h
            o = obs(th)
            wgt = np.exp(-beta * (o[:, 0] - o0[:, 0]))
            v.append((o[:, 1] * wgt).sum() / wgt.sum())
        fd = (v[0] - v[1]) / (2 * h)
        assert fg[k] == pytest.approx(fd, rel=1e-6, abs=1e-8), (k, fg[k], fd)


def test_mean_se_tau_ar1_coverage():
    """AR(1) runs of known mean: the tau-corrected standard error covers
    the true mean at the Gaussian rate (about 99.7 % within 3 SE over many
    runs), where the naive iid SE would not, and the Sokal SE of tau_int
    matches the run-to-run spread of tau."""
    a, n, runs = 0.9, 4000, 400
    rng = np.random.default_rng(21)
    z, taus, tse = [], [], []
    for r in range(runs):
        e = rng.standard_normal(n)
        x = np.empty(n)
        x[0] = e[0] / math.sqrt(1 - a * a)
        for t in range(1, n):
            x[t] = a * x[t - 1] + e[t]
        m = stats.mean_se_tau(x)
        z.append(m["mean"] / m["se"])
        taus.append(m["tau"])
        tse.append(m["tau_se"])
    z = np.array(z)
    assert (np.abs(z) <= 3).mean() >= 0.98
    assert 0.85 < z.std() < 1.15
    assert np.std(taus) == pytest.approx(np.mean(tse), rel=0.35)
    assert np.mean(taus) == pytest.approx((1 + a) / (2 * (1 - a)), rel=0.1)


def test_compare_runs_null_and_planted_bias():
    """compare_runs on synthetic chains: two independent runs of the same
    process pass (>= 99 % within 3 sigma); a run with its mean shifted by
    5 standard errors fails; chains with too few independent samples are
    excluded and counted."""
    rng = np.random.default_rng(22)
    T, P, N = 1.0, 0.5, 100

    def run(shift=0.0, a=0.5, n=2000):
        out = []
        for c in range(300):
            e = rng.standard_normal((2, n))
            U = np.empty(n)
            L = np.empty(n)
            U[0], L[0] = e[0, 0], e[1, 0]
            for t in range(1, n):
                U[t] = a * U[t - 1] + e[0, t]
                L[t] = a * L[t - 1] + e[1, t]
            out.append(stats.chain_observables(U - 50.0 + shift, L + 200.0, N, P, T))
        return out

    r0, r1 = run(), run()
    cmp = stats.compare_runs(r0, r1)
    for k in ("U", "L", "rho"):
        assert cmp[k]["frac_within_3sigma"] >= 0.99, (k, cmp[k])
        assert cmp[k]["chains_compared"] == 300
    se = r0[0]["U"][1]
    bad = stats.compare_runs(r0, run(shift=5 * se * math.sqrt(2)))
    assert bad["U"]["frac_within_3sigma"] < 0.5
    assert bad["L"]["frac_within_3sigma"] >= 0.99
    few = stats.compare_runs(r0, r1, n_eff_min=1e9)
    assert few["U"]["chains_compared"] == 0 and few["U"]["chains_excluded"] == 300


def test_chain_observables_batch_matches_single():
    """The vectorised per-chain statistics equal the one-chain function."""
    rng = np.random.default_rng(23)
    nc, n = 7, 3000
    U = np.cumsum(rng.standard_normal((nc, n)), axis=1) * 0.01 + rng.standard_normal((nc, n)) - 40
    L = 200 + rng.standard_normal((nc, n)) + np.sin(np.arange(n) / 50.0)
    N, P, T = np.arange(50, 57), np.linspace(0.2, 1.0, nc), np.linspace(0.6, 1.8, nc)
    bt = stats.chain_observables_batch(U, L, N, P, T, chunk=3)
    for c in range(nc):
        one = stats.chain_observables(U[c], L[c], int(N[c]), P[c], T[c])
        for k in stats.OBSERVABLES:
            for j in range(3):
                assert bt[k][j][c] == pytest.approx(one[k][j], rel=1e-9, abs=1e-12), (c, k, j)


def test_replica_band_closed_forms():
    """The t band against closed forms of Student's t quantile: one degree
    of freedom (Cauchy), tan(pi (p - 1/2)); two, (2p - 1) sqrt(2 / (4p(1-p)));
    and the normal 3 sigma limit for many degrees of freedom."""
    p = 0.5 + 0.9973 / 2
    assert stats.replica_band(2) == pytest.approx(math.tan(math.pi * (p - 0.5)), rel=1e-9)
    assert stats.replica_band(3) == pytest.approx((2 * p - 1) * math.sqrt(2.0 / (4 * p * (1 - p))), rel=1e-9)
    assert stats.replica_band(100000) == pytest.approx(3.0, abs=2e-3)


def test_replica_compare_null_and_planted_bias():
    """replica_compare on synthetic per-chain values: a test run drawn like
    the K reference runs lies within the band at the band's probability and
    its t has the standard deviation sqrt((K-1)/(K-3)) of Student's t; a
    test run shifted by two of its own spreads fails, with mean t that of
    the noncentral t of 2 / sqrt(1 + 1/K); chains whose reference runs do not vary are excluded
    and counted; the per-group summaries split the chains."""
    rng = np.random.default_rng(24)
    K, nc = 16, 20000
    mu = rng.uniform(-100, 100, nc)
    sd = rng.uniform(0.1, 10, nc)

    def draw(shift=0.0):
        z = rng.standard_normal(nc)
        v = mu + shift * sd + sd * z
        out = {k: v for k in stats.OBSERVABLES}
        # log-compared observables: normal on the log scale
        for k in stats.LOG_COMPARED:
            out[k] = np.exp(mu / 50.0 + (shift + z) * sd / 20.0)
        return out

    ref = [draw() for _ in range(K)]
    null = stats.replica_compare(ref, draw(), groups={"first": np.arange(nc) < 100})
    for k in stats.OBSERVABLES:
        r = null[k]
        assert r["chains_compared"] == nc and r["chains_excluded"] == 0
        assert r["band"] == pytest.approx(stats.replica_band(K))
        assert abs(r["frac_within_band"] - 0.9973) < 5 * math.sqrt(0.0027 * 0.9973 / nc)
        assert r["std_t"] == pytest.approx(math.sqrt((K - 1) / (K - 3)), rel=0.03)
        assert abs(r["mean_t"]) < 4 * r["mean_t_se"]
        assert r["first"]["chains"] == 100
    bad = stats.replica_compare(ref, draw(shift=2.0))
    # noncentral t: mean delta sqrt(nu/2) Gamma((nu-1)/2) / Gamma(nu/2)
    nu, delta = K - 1, 2.0 / math.sqrt(1 + 1 / K)
    want = delta * math.sqrt(nu / 2) * math.exp(math.lgamma((nu - 1) / 2) - math.lgamma(nu / 2))
    assert bad["U"]["mean_t"] == pytest.approx(want, abs=0.03)
    assert bad["U"]["frac_within_band"] < 0.95
    flat = [dict(r) for r in ref]
    for r in flat:
        r["L"] = np.ones(nc)
    ex = stats.replica_compare(flat, draw())
    assert ex["L"]["chains_compared"] == 0 and ex["L"]["chains_excluded"] == nc
    # a constant whose mean over the runs carries rounding (sd ~ 1e-13)
    # is constant, not compared
    L0 = np.full(nc, 1024 / 0.6)
    noisy = [dict(r, L=L0.copy()) for r in ref]
    assert np.stack([r["L"] for r in noisy]).std(axis=0, ddof=1).max() > 0
    cst = stats.replica_compare(noisy, dict(draw(), L=L0.copy()))
    assert cst["L"]["chains_compared"] == 0 and cst["L"]["chains_constant"] == nc
    with pytest.raises(ValueError):
        stats.replica_compare(ref[:2], draw())


def test_run_values_match_chain_observables():
    """The per-run values of many chains equal chain_observables' values
    chain by chain (means, block fluctuation formulas, tau_int of U)."""
    rng = np.random.default_rng(25)
    nc, n = 5, 3000
    U = np.cumsum(rng.standard_normal((nc, n)), axis=1) * 0.01 + rng.standard_normal((nc, n)) - 40
    L = 200 + rng.standard_normal((nc, n)) + np.sin(np.arange(n) / 50.0)
    N, P, T = np.arange(50, 55), np.linspace(0.2, 1.0, nc), np.linspace(0.6, 1.8, nc)
    v = stats.run_values(U, L, N, P, T, workers=2)
    for c in range(nc):
        one = stats.chain_observables(U[c], L[c], int(N[c]), P[c], T[c])
        for k in stats.OBSERVABLES:
            assert v[k][c] == pytest.approx(one[k][0], rel=1e-9, abs=1e-12), (c, k)


def test_band_allowance():
    """1 % of the chains for many chains; for few, the 99 % quantile of the
    binomial count outside a 99.73 % band (64 chains: P(more than 2
    outside) < 1 %, P(more than 1) > 1 %)."""
    assert stats.band_allowance(4096) == 40
    q = 0.0027
    p_le = lambda k, n: sum(math.comb(n, j) * q ** j * (1 - q) ** (n - j) for j in range(k + 1))
    assert p_le(1, 64) < 0.99 <= p_le(2, 64)
    assert stats.band_allowance(64) == 2


def _ar1_rows(a, nc, n, rng):
    """nc independent stationary AR(1) rows (x_t = a x_{t-1} + e_t)."""
    from scipy.signal import lfilter
    e = rng.standard_normal((nc, n))
    e[:, 0] /= math.sqrt(1 - a * a)
    return lfilter([1.0], [1.0, -a], e, axis=1)


def test_common_window_tau():
    """tau_int at one window per chain: the window is the median of the
    reference runs' own Sokal windows; every run's value is its running
    autocorrelation sum there, which equals its own estimate where its own
    window is that window; on AR(1) rows (a = 0.8) the estimates centre on
    (1 + a) / (2 (1 - a)) = 4.5; a null run drawn like the references lies
    within the band, a run with a = 0.9 (tau_int 9.5) does not."""
    rng = np.random.default_rng(26)
    K, nc, n, a = 16, 192, 4000, 0.8
    N, P, T = np.full(nc, 100), np.full(nc, 0.5), np.full(nc, 1.0)
    L = np.full((nc, n), 200.0) + 0.01 * rng.standard_normal((nc, n))

    def one(aa):
        return stats.run_values(_ar1_rows(aa, nc, n, rng), L, N, P, T, workers=2, acf_lags=300)

    ref = [one(a) for _ in range(K)]
    null, slow = one(a), one(0.9)
    own = {id(r): r["tau_U"].copy() for r in ref + [null]}
    w = stats.common_window_tau(ref, ref + [null, slow])
    med = np.floor(np.median(np.stack([r["_tau_window"] for r in ref]), axis=0)).astype(int)
    assert np.array_equal(w, med + 1)
    for r in ref + [null]:
        same = r["_tau_window"] == med
        assert same.any()
        assert np.allclose(r["tau_U"][same], own[id(r)][same], rtol=1e-6)
        k = np.minimum(med, r["_tau_cum"].shape[1] - 1)
        assert np.array_equal(r["tau_U"], np.maximum(r["_tau_cum"][np.arange(nc), k].astype(np.float64), 0.5))
    assert np.median(np.concatenate([r["tau_U"] for r in ref])) == pytest.approx((1 + a) / (2 * (1 - a)), rel=0.05)
    ok = stats.replica_compare(ref, null)["tau_U"]
    assert ok["outside"] <= ok["allowed_outside"], ok
    bad = stats.replica_compare(ref, slow)["tau_U"]
    assert bad["frac_within_band"] < 0.5, bad
