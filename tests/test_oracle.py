"""Pins for the CPU oracle (oracle/oracle.c) from outside itself: published
known-answer vectors, closed forms, finite differences, brute-force sums, and
exact properties of the schedule.  A plausible slip in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails one of these.
"""
import math

import numpy as np
import pytest

from helpers import CASES, hrsw, lj, oracle_spec, rel, yukawa

from jmmonedmc_amd import _native as nat
from jmmonedmc_amd.engine import Chain


# ------------------------------------------------------------------ Philox

# Random123 kat_vectors (philox4x32, 10 rounds): ctr, key -> out
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, out):
    assert tuple(oracle.philox(ctr, key)) == out


def test_u53_range(oracle):
    assert oracle.u53(0, 0) == 0.0
    assert oracle.u53(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    assert oracle.u53(0x80000000, 0) == 0.5
    # monotone in the high word
    v = [oracle.u53(a, 0) for a in range(0, 1 << 32, 1 << 27)]
    assert all(x < y for x, y in zip(v, v[1:]))


def test_u32_range(oracle):
    assert oracle.u32(0) == 2.0 ** -33
    assert oracle.u32(0xFFFFFFFF) == 1.0 - 2.0 ** -33
    assert oracle.u32(0x80000000) == 0.5 + 2.0 ** -33
    assert oracle.u32(12345) == (12345 + 0.5) / 2.0 ** 32


def test_trial_stream_is_pairwise_philox(oracle):
    """orc_trial's uniforms: for a block pair (t even, t+1) of particles
    (i, i+1), one Philox block at counter (i, sweep, chain_id, 0); words 0/1
    drive particle i (displacement, acceptance), words 2/3 particle i+1.
    Replayed by hand through orc_trial_with on a copy of the chain."""
    from jmmonedmc_amd.engine import Chain
    from helpers import lj, oracle_spec
    ch = Chain(N=37, L0=74.0, seed=0x1234567890AB, chain_id=9, bcap=4, **lj())
    spec = oracle_spec(oracle, ch)
    a, b = oracle.Chain(spec), oracle.Chain(spec)
    B = a.counter("B")
    q, rem = ch.N // B, ch.N % B
    key = (0x567890AB, 0x1234)
    sweep = 0
    for t in range(q + 1):
        for blk in range(B):
            size = q + (blk < rem)
            if t >= size:
                continue
            i = blk * q + min(blk, rem) + t
            r = oracle.philox((i - (t & 1), sweep, 9, 0), key)
            p = 2 * (t & 1)
            md = ((oracle.u32(r[p]) - 0.5) * 2.0) * ch.max_step
            acc_a, dEa = a.trial(i, sweep)
            acc_b, dEb = b.trial_with(i, md, oracle.u32(r[p + 1]))
            assert (acc_a, dEa) == (acc_b, dEb), (t, blk, i)
    assert np.array_equal(a.state()[0], b.state()[0])


def test_det_exp_against_libm(oracle):
    xs = np.concatenate([np.linspace(-0.02, 0.02, 2001), np.linspace(-700, 700, 2001),
                         [0.0, 1e-300, -1e-300, 0.5 * math.log(2), -0.5 * math.log(2)]])
    for x in xs:
        e = oracle.det_exp(float(x))
        ref = math.exp(float(x))
        assert abs(e - ref) <= 2 * math.ulp(ref), (x, e, ref)


# ------------------------------------------------------------ potentials

def spec_of(**kw):
    import pyoracle
    ch = Chain(N=8, L0=100.0, **kw)
    return pyoracle.make_spec(**ch.as_dict())


def test_lj_closed_forms(oracle):
    s = spec_of(**lj(shifted=0))
    assert oracle.phi(s, 1.0)[0] == 0.0
    rmin = 2 ** (1 / 6)
    o = oracle.phi(s, rmin)
    assert abs(o[0] + 1.0) < 1e-14
    assert abs(o[1]) < 1e-13  # virial vanishes at the minimum
    assert abs(oracle.phi(s, 3.0)[0] - 4 * (3.0 ** -12 - 3.0 ** -6)) < 1e-18
    assert abs(oracle.phi(s, 3.0)[0] - (-5.4794417e-3)) < 1e-9
    # inclusive cutoff (src/pot.cpp:53)
    assert oracle.phi(s, 3.0)[0] != 0.0
    assert oracle.phi(s, math.nextafter(3.0, 4.0))[0] == 0.0
    # shifted form vanishes at the cutoff and differs by phi(rc) inside
    ss = spec_of(**lj(shifted=1))
    assert oracle.phi(ss, 3.0)[0] == 0.0
    assert abs((oracle.phi(ss, 1.3)[0] - oracle.phi(s, 1.3)[0]) - 5.4794417e-3) < 1e-9
    # eps, sigma scaling: phi(sigma d; eps, sigma) = eps phi(d; 1, 1)
    sg = spec_of(**lj(shifted=0, cutoff=10.0, pot_p=(2.5, 1.7, 0, 0)))
    for d in (1.1, 1.5, 2.2):
        assert rel(oracle.phi(sg, 1.7 * d)[0], 2.5 * oracle.phi(s, d)[0]) < 1e-13


def test_lj_matches_reference_layout(oracle):
    """General LJ at eps = sigma = 1, truncated, reproduces the reference's
    phiLJcut components bit for bit (E, W, HV)."""
    import ctypes as C
    s = spec_of(**lj(shifted=0))
    for d in np.linspace(0.6, 3.2, 101):
        ref = np.zeros(9)
        oracle.lib().orc_phi_ljcut_ref(float(d), 3.0, 1, ref.ctypes.data_as(C.POINTER(C.c_double)))
        o = oracle.phi(s, float(d))
        assert o[0] == ref[0] and o[1] == ref[1] and o[6] == ref[6]


def test_harmonic_closed_forms(oracle):
    s = spec_of(potential=nat.POT_HARMONIC, cutoff=2.0, pot_p=(0, 0, 0, 0), shifted=0)
    assert oracle.phi(s, 0.0)[0] == 10e10 and oracle.phi(s, -0.5)[0] == 10e10
    assert oracle.phi(s, 1.0)[0] == 0.0
    assert oracle.phi(s, 1.5)[0] == 0.25
    assert oracle.phi(s, 2.0)[0] == 0.0            # strict cutoff (src/pot.cpp:120)
    assert oracle.phi(s, math.nextafter(2.0, 0))[0] > 0.99


def test_hrsw_and_yukawa_closed_forms(oracle):
    s = spec_of(**hrsw())
    assert math.isinf(oracle.phi(s, 0.999)[0])
    assert oracle.phi(s, 1.0)[0] == -1.0
    assert oracle.phi(s, 1.4999)[0] == -1.0
    assert oracle.phi(s, 1.5)[0] == 0.0
    y = spec_of(**yukawa())
    assert math.isinf(oracle.phi(y, 0.49)[0])
    for d in (0.5, 1.0, 7.3, 20.0):
        assert rel(oracle.phi(y, d)[0], math.exp(-0.2 * d) / d) < 1e-15
    assert oracle.phi(y, 20.000001)[0] == 0.0


@pytest.mark.parametrize("kw,ds", [
    (lj(shifted=1, cutoff=5.0, pot_p=(1.3, 0.9, 0, 0)), (0.8, 1.0, 1.4, 2.9)),
    (yukawa(pot_p=(1.7, 0.35, 0.5, 0)), (0.7, 2.0, 9.0)),
    (dict(potential=nat.POT_HARMONIC, cutoff=5.0, pot_p=(0, 0, 0, 0), shifted=0), (0.5, 1.3, 2.0)),
])
def test_virial_and_theta_gradients_by_finite_differences(oracle, kw, ds):
    s = spec_of(**kw)
    h = 1e-6
    for d in ds:
        o = oracle.phi(s, d)
        dphi = (oracle.phi(s, d + h)[0] - oracle.phi(s, d - h)[0]) / (2 * h)
        assert abs(o[1] - (-d * dphi)) <= 1e-6 * max(1, abs(o[1]))
        # hypervirial HV = -d dW/dd
        dW = (oracle.phi(s, d + h)[1] - oracle.phi(s, d - h)[1]) / (2 * h)
        assert abs(o[6] - (-d * dW)) <= 1e-6 * max(1, abs(o[6])), (d, o[6], -d * dW)
        if kw["potential"] == nat.POT_HARMONIC:
            continue
        for k in (0, 1):
            p = list(kw["pot_p"])
            up, dn = list(p), list(p)
            up[k] += h
            dn[k] -= h
            su = spec_of(**{**kw, "pot_p": tuple(up)})
            sd = spec_of(**{**kw, "pot_p": tuple(dn)})
            fd_e = (oracle.phi(su, d)[0] - oracle.phi(sd, d)[0]) / (2 * h)
            fd_w = (oracle.phi(su, d)[1] - oracle.phi(sd, d)[1]) / (2 * h)
            assert abs(o[2 + k] - fd_e) <= 1e-6 * max(1, abs(fd_e)), (d, k)
            assert abs(o[4 + k] - fd_w) <= 1e-6 * max(1, abs(fd_w)), (d, k)


# -------------------------------------------------------------- schedule

@pytest.mark.parametrize("name", sorted(CASES))
def test_schedule_invariants(oracle, name):
    import pyoracle
    for ch in CASES[name]():
        s = oracle_spec(pyoracle, ch)
        B, M, w = oracle.schedule(s)
        # a power of two, or whole waves when bcap is not a power of two
        assert B >= 1 and (B & (B - 1) == 0 or (ch.bcap & (ch.bcap - 1) != 0 and B % 64 == 0))
        N = ch.N
        if B > 1:
            assert N // B >= M                      # every block at least M_min long
        g = {nat.POT_LJ: ch.pot_p[2] if ch.pot_p[2] > 0 else 0.5 * ch.pot_p[1], nat.POT_HRSW: ch.pot_p[0],
             nat.POT_YUKAWA: ch.pot_p[2]}.get(ch.potential, 0.0)
        if ch.nbn >= 0:
            assert w <= ch.nbn and M <= ch.nbn + 1
        if g > 0 and math.isfinite(ch.cutoff) and ch.nbn < 0 and M <= N:
            # partner M (and w + 1) lies beyond rc by at least the guard,
            # every partner that can lie within rc is walked
            assert M == w + 1 and M * g >= ch.cutoff + 1e-3 * g
            assert w >= min(N - 1, math.floor(ch.cutoff / g))
        # engine and oracle compute the same schedule
        from jmmonedmc_amd.engine import schedule
        assert schedule(ch) == (B, M, w)


def _phase_members(N, B, t):
    q, r = divmod(N, B)
    out = []
    for b in range(B):
        size = q + (b < r)
        start = b * q + min(b, r)
        if t < size:
            out.append(start + t)
    return out


@pytest.mark.parametrize("name", ["c1_periodic_nvt", "hrsw_npt", "yukawa_npt", "lj_inf_nbn2_walls",
                                  "npt_periodic_log"])
def test_phase_trials_commute(oracle, name):
    """Trials of one phase run in forward and in reverse order give the same
    state bit for bit: the independence the parallel schedule relies on."""
    import pyoracle
    ch = CASES[name]()[0]
    s = oracle_spec(pyoracle, ch)
    B, M, w = oracle.schedule(s)
    a = oracle.Chain(s)
    a.sweeps(5)
    x0, L0, _, sw = a.state()
    rng = np.random.default_rng(0)
    for t in range((ch.N + B - 1) // B):
        members = _phase_members(ch.N, B, t)
        md = rng.uniform(-ch.max_step, ch.max_step, len(members))
        u = rng.uniform(0, 1, len(members))
        fwd = oracle.Chain(s)
        fwd.set_state(x0, L0, sw)
        rev = oracle.Chain(s)
        rev.set_state(x0, L0, sw)
        for k in range(len(members)):
            fwd.trial_with(members[k], md[k], u[k])
        for k in reversed(range(len(members))):
            rev.trial_with(members[k], md[k], u[k])
        assert np.array_equal(fwd.state()[0], rev.state()[0]), f"phase {t}"
        x0 = fwd.state()[0]


# ------------------------------------------------------------ energies

def _random_config(rng, N, L, gmin, walls):
    """Uniform points with every gap >= gmin (and inside the walls)."""
    free = L - N * gmin - (gmin if not walls else 0.0)
    g = np.sort(rng.uniform(0, free, N))
    x = g + gmin * np.arange(N)
    x = x - L / 2 + (gmin / 2 if walls else 0.0)
    return x


@pytest.mark.parametrize("kw,gmin", [(lj(), 0.5), (lj(shifted=0), 0.5), (hrsw(), 1.0),
                                     (yukawa(cutoff=6.0), 0.5),
                                     (dict(potential=nat.POT_HARMONIC, cutoff=1.8,
                                           pot_p=(0, 0, 0, 0), shifted=0), 0.05)])
@pytest.mark.parametrize("boundary", [nat.WALLS, nat.PERIODIC])
@pytest.mark.parametrize("nbn", [-1, 1, 3])
def test_window_energy_equals_brute_force(oracle, kw, gmin, boundary, nbn):
    import pyoracle
    rng = np.random.default_rng(1)
    for N in (2, 3, 5, 17, 64):
        L = max(N * (gmin + 1.0), 2 * (kw["cutoff"] + 0.5) + 1)
        ch = Chain(N=N, L0=L, boundary=boundary, nbn=nbn, **kw)
        s = oracle_spec(pyoracle, ch)
        x = _random_config(rng, N, L, gmin, boundary == nat.WALLS)
        ub = oracle.brute_energy_x(s, x, L)
        uw = oracle.full_energy_x(s, x, L)
        assert rel(uw, ub) <= 1e-13, (N, uw, ub)


@pytest.mark.parametrize("reach", [1.0, 4.0])
@pytest.mark.parametrize("name", sorted(CASES))
def test_trial_energy_change_equals_brute_force(oracle, name, reach):
    """dE of single trials equals brute(after) - brute(before) on every
    accepted trial, and rejected trials leave the state untouched; also
    for displacements up to four times the chain's step (the partner window
    rests on the hard core alone, not on the step size)."""
    import pyoracle
    ch = CASES[name]()[0]
    s = oracle_spec(pyoracle, ch)
    c = oracle.Chain(s)
    rng = np.random.default_rng(2)
    for _ in range(300):
        i = int(rng.integers(ch.N))
        md = float(rng.uniform(-reach * ch.max_step, reach * ch.max_step))
        x0, L, U0, _ = c.state()
        b0 = c.brute_energy()
        acc, dE = c.trial_with(i, md, 0.0)   # u = 0: every finite dE is accepted
        x1, _, U1, _ = c.state()
        b1 = c.brute_energy()
        if acc:
            assert abs((b1 - b0) - dE) <= 1e-9 * max(1.0, abs(b0)), (i, md)
            assert x1[i] == x0[i] + md
        else:
            assert np.array_equal(x0, x1)


def _inside(ch, d):
    """The cutoff test as the potentials define it: inclusive for LJ and
    Yukawa, strict for the harmonic spring and the square well's edge."""
    if ch.potential == nat.POT_HRSW:
        return d < ch.pot_p[2]
    if ch.potential == nat.POT_HARMONIC:
        return d < ch.cutoff
    return d <= ch.cutoff


@pytest.mark.parametrize("name", sorted(CASES))
def test_pair_term_counter_is_n_old_plus_n_new(oracle, name):
    """The pair-term counter of a trial that reaches the energy sum grows by
    the partners within the cutoff of the old position plus those within
    the cutoff of the new one, counted here over every partner (and every
    periodic image on either side) within the NBN limit."""
    import pyoracle
    ch = CASES[name]()[0]
    c = oracle.Chain(oracle_spec(pyoracle, ch))
    rng = np.random.default_rng(5)
    N = ch.N
    nb = N - 1 if ch.nbn < 0 else min(ch.nbn, N - 1)
    checked = 0
    for _ in range(200):
        i = int(rng.integers(N))
        md = float(rng.uniform(-ch.max_step, ch.max_step))
        x, L, _, _ = c.state()
        before = [c.counter(k) for k in ("pair_terms", "wall_rej", "hc_rej")]
        c.trial_with(i, md, 0.5)
        after = [c.counter(k) for k in ("pair_terms", "wall_rej", "hc_rej")]
        if after[1] != before[1] or after[2] != before[2]:
            assert after[0] == before[0]
            continue
        xi, xn = x[i], x[i] + md
        want = 0
        for k in range(1, nb + 1):
            for j, sign in ((i - k, -1), (i + k, 1)):
                if ch.boundary == nat.WALLS and not 0 <= j < N:
                    continue
                xj = x[j % N] + (L if j >= N else -L if j < 0 else 0.0)
                o, n = sign * (xj - xi), sign * (xj - xn)
                want += int(_inside(ch, o)) + int(_inside(ch, n))
        assert after[0] - before[0] == want, (i, md)
        checked += 1
    assert checked > 50


def test_accept_rule_hand_cases(oracle):
    L = oracle.lib()
    beta = 1.0 / 0.7
    assert L.orc_accept(0.0, beta, 0.999999) == 1          # dE <= 0 always accepted
    assert L.orc_accept(-3.0, beta, 0.5) == 1
    dE = 0.4
    bf = math.exp(-dE * beta)
    assert L.orc_accept(dE, beta, bf) == 0                 # tie: strict '>' rejects
    assert L.orc_accept(dE, beta, math.nextafter(bf, 0)) == 1
    assert L.orc_accept(dE, beta, math.nextafter(bf, 1)) == 0
    assert L.orc_accept(1e4, beta, 0.0) == 0               # underflowed factor never accepted
    # box-length trials: accept when bf >= 1, else bf > u
    assert L.orc_vol_accept(-1.0, 0.0, 0.0, 1.0, 1.0, 10.0, 0.999) == 1
    assert L.orc_vol_accept(0.0, 0.5, 0.0, 1.0, 1.0, 10.0, math.exp(-0.5)) == 0
    assert L.orc_vol_accept(0.0, 0.5, 0.0, 1.0, 1.0, 10.0, math.nextafter(math.exp(-0.5), 0)) == 1
    # Jacobian sign: growing the box (ln s > 0) helps acceptance
    assert L.orc_vol_accept(0.3, 0.0, 0.05, 0.0, 1.0, 10.0, 0.95) == 1


def test_wall_and_core_rejections(oracle):
    import pyoracle
    ch = Chain(N=4, L0=8.0, boundary=nat.WALLS, **lj(shifted=0))
    c = oracle.Chain(oracle_spec(pyoracle, ch))
    x, L, _, _ = c.state()                 # lattice -3, -1, 1, 3
    acc, _ = c.trial_with(3, 1.2, 0.0)     # 4.2 > L/2: wall
    assert acc == 0 and c.counter("wall_rej") == 1
    acc, _ = c.trial_with(1, 1.6, 0.0)     # gap to particle 2 becomes 0.4 < 0.5: core
    assert acc == 0 and c.counter("hc_rej") == 1
    assert np.array_equal(c.state()[0], x)


def test_volume_trial_energy_and_scaling_identity(oracle):
    """Accepted box-length trials land on the brute-force energy of the scaled
    configuration; with no cutoff the LJ energy obeys
    E(s x) = s^-12 E12 - s^-6 E6 (the identity qavLJ uses,
    src/jmmMCState.cpp:1658-1664)."""
    import ctypes as C
    import pyoracle
    ch = Chain(N=12, L0=30.0, boundary=nat.WALLS, ensemble=nat.NPT, vol_move=nat.VOL_LINEAR,
               max_dv=2.0, **lj(cutoff=math.inf, shifted=0))
    s = oracle_spec(pyoracle, ch)
    c = oracle.Chain(s)
    c.sweeps(10)
    x, L, U, _ = c.state()
    e12 = e6 = 0.0
    for i in range(12):
        for j in range(i + 1, 12):
            ref = np.zeros(9)
            oracle.lib().orc_phi_ljcut_ref(x[j] - x[i], math.inf, 1,
                                           ref.ctypes.data_as(C.POINTER(C.c_double)))
            e12 += ref[2]
            e6 += ref[4]
    for dv in (-1.5, 0.7, 1.9):
        sc = (L + dv) / L
        want = sc ** -12 * e12 - sc ** -6 * e6
        got = oracle.brute_energy_x(s, x * sc, L + dv)
        assert rel(got, want) <= 1e-12
    acc, dU = c.volume_trial_with(0.7, 0.0)
    assert acc == 1
    x2, L2, U2, _ = c.state()
    assert L2 == L + 0.7
    assert rel(U2, c.brute_energy()) <= 1e-12


@pytest.mark.parametrize("name", sorted(CASES))
def test_incremental_energy_equals_recompute(oracle, name):
    import pyoracle
    for ch in CASES[name]():
        c = oracle.Chain(oracle_spec(pyoracle, ch))
        c.sweeps(60)
        assert rel(c.state()[2], c.brute_energy()) <= 1e-10


@pytest.mark.parametrize("name", ["c1_walls_truncated", "npt_walls_linear", "lj_inf_walls_npt",
                                  "harmonic_nbn1_walls", "hrsw_npt", "npt_periodic_log", "yukawa_npt",
                                  "lj_exclusion_npt", "c1_periodic_nvt"])
def test_windowed_sweeps_equal_all_partner_sweeps(oracle, name):
    """The reference's cost model (all N-1 partners per trial) and the window
    form make the same decisions and reach the same energies."""
    import pyoracle
    ch = CASES[name]()[0]
    a = oracle.Chain(oracle_spec(pyoracle, ch))
    b = oracle.Chain(oracle_spec(pyoracle, ch))
    sa = a.sweeps(30)
    sb = b.sweeps_allpairs(30)
    assert np.array_equal(sa["acc_d"], sb["acc_d"])
    assert np.array_equal(a.state()[0], b.state()[0])
    assert np.allclose(sa["U"], sb["U"], rtol=1e-10, atol=1e-10)


def test_observables_by_finite_differences(oracle):
    """dU/dx and dU/dtheta of a random LJ configuration against central
    differences of the brute-force energy."""
    import pyoracle
    rng = np.random.default_rng(3)
    for boundary in (nat.WALLS, nat.PERIODIC):
        ch = Chain(N=20, L0=40.0, boundary=boundary, **lj(pot_p=(1.1, 0.95, 0, 0)))
        s = oracle_spec(pyoracle, ch)
        x = _random_config(rng, 20, 40.0, 0.8, boundary == nat.WALLS)
        obs, g = oracle.observables_x(s, x, 40.0)
        assert rel(obs[0], oracle.brute_energy_x(s, x, 40.0)) <= 1e-13
        h = 1e-6
        for i in range(20):
            xp, xm = x.copy(), x.copy()
            xp[i] += h
            xm[i] -= h
            fd = (oracle.brute_energy_x(s, xp, 40.0) - oracle.brute_energy_x(s, xm, 40.0)) / (2 * h)
            assert abs(g[i] - fd) <= 1e-6 * max(1, abs(fd))
        for k, dp in ((0, (h, 0)), (1, (0, h))):
            sp = oracle_spec(pyoracle, Chain(N=20, L0=40.0, boundary=boundary,
                                            **lj(pot_p=(1.1 + dp[0], 0.95 + dp[1], 0, 0))))
            sm = oracle_spec(pyoracle, Chain(N=20, L0=40.0, boundary=boundary,
                                            **lj(pot_p=(1.1 - dp[0], 0.95 - dp[1], 0, 0))))
            fd = (oracle.brute_energy_x(sp, x, 40.0) - oracle.brute_energy_x(sm, x, 40.0)) / (2 * h)
            assert abs(obs[2 + k] - fd) <= 1e-6 * max(1, abs(fd))
        # virial: W = -dU/d ln(scale) at fixed box for a periodic system
        if boundary == nat.PERIODIC:
            e = 1e-6
            up = oracle.brute_energy_x(s, x * (1 + e), 40.0 * (1 + e))
            dn = oracle.brute_energy_x(s, x * (1 - e), 40.0 * (1 - e))
            assert abs(obs[1] + (up - dn) / (2 * e)) <= 1e-6 * max(1, abs(obs[1]))


def _hist_spec(oracle, N, L, boundary):
    from jmmonedmc_amd.engine import Chain
    from jmmonedmc_amd import _native as nat
    from helpers import lj, oracle_spec
    return oracle_spec(oracle, Chain(N=N, L0=L, boundary=boundary, **lj()))


def test_hist_hand_binned_walls(oracle):
    """Three particles between walls (fgrho, src/jmmMCState.cpp:1069-1128):
    rho bins floor(x/rbw + rhonb/2), pair bins floor(d/gbw) in each member's
    segment floor(x/gsw + gns/2); binned by hand."""
    from jmmonedmc_amd import _native as nat
    s = _hist_spec(oracle, 3, 10.0, nat.WALLS)
    rho, g = oracle.hist_x(s, [-2.5, 0.2, 1.3], 10.0, 10, 1.0, 2, 5.0, 4, 1.0)
    assert rho.tolist() == [0, 0, 1, 0, 0, 1, 1, 0, 0, 0]
    # pairs: (0,1) d=2.7 bin 2, (0,2) d=3.8 bin 3, (1,2) d=1.1 bin 1;
    # segments: x=-2.5 -> 0, x=0.2 -> 1, x=1.3 -> 1
    assert g.tolist() == [[0, 0, 1, 1], [0, 2, 1, 1]]


def test_hist_hand_binned_periodic(oracle):
    """Four particles on a ring of length 10: every pair once at its
    separation below L/2, the pairs across the end included."""
    from jmmonedmc_amd import _native as nat
    s = _hist_spec(oracle, 4, 10.0, nat.PERIODIC)
    x = [-4.75, -1.0, 2.0, 4.5]
    rho, g = oracle.hist_x(s, x, 10.0, 10, 1.0, 2, 5.0, 5, 1.0)
    assert rho.tolist() == [1, 0, 0, 0, 1, 0, 0, 1, 0, 1]
    # (0,1) 3.75 -> 3, (1,2) 3.0 -> 3, (2,3) 2.5 -> 2, across the end:
    # (3,0) 0.75 -> 0, (3,1) 4.5 -> 4, (2,0) 3.25 -> 3; segments 0,0,1,1
    assert g.tolist() == [[1, 0, 0, 4, 1], [1, 0, 2, 2, 1]]
    # a rigid shift that carries particles past +L/2: positions fold back,
    # separations and so the segment-summed pair counts are unchanged
    rho2, g2 = oracle.hist_x(s, [v + 3.0 for v in x], 10.0, 10, 1.0, 2, 5.0, 5, 1.0)
    assert int(rho2.sum()) == 4
    assert g2.sum(axis=0).tolist() == g.sum(axis=0).tolist()


# ------------------------------------------------ FP32 error model and replay

def _phi32_lj(d, eps=1.0, sig=1.0, shift=0.0):
    """The kernel's FP32 LJ sequence (phi32) in numpy float32: d rounded to
    single, 1/d^2, s6 = (1/d^2)^3, s6 (4 eps sigma^12 s6 - 4 eps sigma^6) -
    shift in two fused steps (rounded separately here, which only widens the
    error)."""
    f = np.float32
    d = f(d)
    ir2 = f(1.0) / (d * d)
    s6 = ir2 * ir2 * ir2
    return s6 * (f(4.0 * eps * sig ** 12) * s6 - f(4.0 * eps * sig ** 6)) - f(shift)


def _phi32_yukawa(d, A=1.0, k=0.2):
    f = np.float32
    d = f(d)
    return f(A) * np.exp(-f(k) * d) * (f(1.0) / d)


@pytest.mark.parametrize("kind", ["lj", "yukawa"])
def test_fp32_error_bound_covers_single_precision_evaluation(oracle, kind):
    """e32(d) bounds |phi32(d) - phi64(d)| for a float32 evaluation of the
    kernel's sequence over the whole range a chain visits (hard core to
    cutoff), with a factor of 4 to spare for the hardware reciprocal and
    exponential (one and two ulp); and it is far below the change a dropped
    pair term would make, so a replay with a lost term cannot pass as a
    tie."""
    rng = np.random.default_rng(5)
    if kind == "lj":
        s = oracle.make_spec(N=4, L0=100.0, **lj())
        ds = rng.uniform(0.5, 3.0, 4000)
        sh = 4.0 / 3.0 ** 12 - 4.0 / 3.0 ** 6
        f32 = [float(_phi32_lj(d, shift=sh)) for d in ds]
    else:
        s = oracle.make_spec(N=4, L0=100.0, **yukawa())
        ds = rng.uniform(0.5, 20.0, 4000)
        f32 = [float(_phi32_yukawa(d)) for d in ds]
    for d, v in zip(ds, f32):
        exact = oracle.phi(s, d)[0]
        e = oracle.e32_pair(s, d)
        assert abs(v - exact) <= e / 4, (d, v, exact, e)
        assert e < 1e-3 * max(abs(exact), 1e-3) or e < 1e-4


def test_fp32_box_identity_within_bound(oracle):
    """The FP32 box-length pass's LJ form, i6 (c12 i6 - c6) with c12, c6 the
    scaled constants, against phi(s d) - phi(d) in FP64, within the bound
    of the pair's two separations; and the Yukawa form, the difference of
    the two FP32 energies."""
    f = np.float32
    rng = np.random.default_rng(6)
    s = oracle.make_spec(N=4, L0=100.0, **lj())
    for _ in range(2000):
        d0 = rng.uniform(0.5, 3.0)
        lns = rng.uniform(-0.01, 0.01)
        sc = math.exp(lns)
        d1 = sc * d0
        c12, c6 = f(4.0 * math.expm1(-12 * lns)), f(4.0 * math.expm1(-6 * lns))
        ir2 = f(1.0) / (f(d0) * f(d0))
        i6 = ir2 * ir2 * ir2
        got = float(i6 * (c12 * i6 - c6))
        if d1 > 3.0:
            continue
        want = oracle.phi(s, d1)[0] - oracle.phi(s, d0)[0]
        assert abs(got - want) <= (oracle.e32_pair(s, d0) + oracle.e32_pair(s, d1)) / 4
    y = oracle.make_spec(N=4, L0=100.0, **yukawa())
    for _ in range(2000):
        d0 = rng.uniform(0.5, 19.5)
        d1 = math.exp(rng.uniform(-0.01, 0.01)) * d0
        got = float(_phi32_yukawa(d1) - _phi32_yukawa(d0))
        want = oracle.phi(y, d1)[0] - oracle.phi(y, d0)[0]
        assert abs(got - want) <= (oracle.e32_pair(y, d0) + oracle.e32_pair(y, d1)) / 4


def test_follow_replay_records_only_real_disagreements(oracle):
    """orc_sweeps_follow: replaying the oracle's own decisions records no
    flip and ends in the same state; a decision planted into the log (a
    single flipped byte) is recorded with a margin far beyond its bound, and
    a box-length decision planted likewise is recorded as a box flip."""
    s = oracle.make_spec(N=120, rho=0.5, T=1.0, P=0.5, ensemble=oracle.NPT, max_dv=0.05, seed=9)
    a = oracle.Chain(s)
    smp, log = a.sweeps(30, log=True)
    b = oracle.Chain(s)
    smp2, fl, nf = b.sweeps_follow(30, log, smp["acc_v"])
    assert nf == 0 and np.array_equal(a.state()[0], b.state()[0]) and a.state()[1] == b.state()[1]
    np.testing.assert_array_equal(smp2["acc_d"], smp["acc_d"])
    lg = log.reshape(30, 120).copy()
    t = np.flatnonzero(lg[7])[0]
    lg[7, t] = 0
    c = oracle.Chain(s)
    _, fl, nf = c.sweeps_follow(30, lg, smp["acc_v"])
    assert nf >= 1 and fl[0]["sweep"] == 7 and fl[0]["index"] == t and fl[0]["margin"] > 1e3 * fl[0]["bound"]
    va = smp["acc_v"].astype(np.uint8).copy()
    va[3] ^= 1
    d = oracle.Chain(s)
    _, fl, nf = d.sweeps_follow(30, log, va)
    assert nf >= 1 and fl[0]["sweep"] == 3 and fl[0]["index"] == np.iinfo(np.uint64).max
    assert fl[0]["margin"] > fl[0]["bound"]


def test_virials_of_a_configuration_by_hand(oracle):
    """orc_virials_x sums W and HV over the pairs inside the cutoff: three
    LJ particles at gaps 1.0 and 1.2 (the outer pair at 2.2 inside rc = 3),
    and the reference's closed forms 48/d^12 - 24/d^6, 576/d^12 - 144/d^6."""
    s = oracle.make_spec(N=3, L0=100.0, boundary=oracle.WALLS, **lj())
    x = np.array([-1.1, -0.1, 1.1])
    W, HV = oracle.virials_x(s, x, 100.0)
    ds = (1.0, 1.2, 2.2)
    assert W == pytest.approx(sum(48 / d ** 12 - 24 / d ** 6 for d in ds), rel=1e-13)
    assert HV == pytest.approx(sum(576 / d ** 12 - 144 / d ** 6 for d in ds), rel=1e-13)


def test_sweeps_virial_is_sweeps_plus_virials(oracle):
    """orc_sweeps_virial runs the same sweeps as orc_sweeps (the same samples
    and final state) and records after every sweep the virials
    orc_virials_x gives for the configuration at that point."""
    s = oracle.make_spec(N=80, rho=0.6, T=1.0, P=0.5, ensemble=oracle.NPT, max_dv=0.02, seed=17, **lj())
    a, b = oracle.Chain(s), oracle.Chain(s)
    smp = a.sweeps(12)
    vsmp, vir = b.sweeps_virial(12)
    np.testing.assert_array_equal(smp, vsmp)
    xa, La, _, _ = a.state()
    xb, Lb, _, _ = b.state()
    assert np.array_equal(xa, xb) and La == Lb
    assert tuple(vir[-1]) == tuple(oracle.virials_x(s, xb, Lb))
    c = oracle.Chain(s)
    for k in range(3):
        c.sweeps(1)
        x, L, _, _ = c.state()
        assert tuple(vir[k]) == tuple(oracle.virials_x(s, x, L))
