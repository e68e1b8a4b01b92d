"""The FP32 form against the FP64 oracle at the sizes it is measured on.

The FP32 form evaluates pair energies in single precision, so a decision can
differ from the oracle's where the decision's margin (|ln u + beta dE| for a
displacement trial, |ln u - arg| for a box-length trial) is within the
single-precision error of the energy change: a near-threshold tie.  Each
test runs the FP32 form with its accept log, then the oracle replays the
same sweeps taking the FP32 form's decisions (orc_sweeps_follow): every
decision that differs from the oracle's own is a flip, and each must be an
energy decision whose margin is within the error bound of that trial
(oracle/oracle.c, single-precision error model).  Because the oracle
follows the FP32 form's decision at a flip, both chains stay in the same
state, every later trial is compared again, and the final positions and
lengths must agree bit for bit (positions and lengths are FP64 state in both
forms).  Reference acceptance rule: src/jmmMCState.cpp:1367-1377 (displacement),
:2249-2255 (box length).
"""
import concurrent.futures as cf

import numpy as np
import pytest

from jmmonedmc_amd import _native as nat
from jmmonedmc_amd import configs
from jmmonedmc_amd.engine import Engine

from helpers import CASES, oracle_spec, rel

COUNTERS = (("trials", "trials"), ("acc_d", "acc_d"), ("acc_v", "acc_v"), ("wall_rejects", "wall_rej"),
            ("core_rejects", "hc_rej"), ("pair_terms", "pair_terms"), ("recomputes", "recomputes"))


def check_flips(flips, n_flips, label):
    """Every flip an energy decision within its trial's bound."""
    assert n_flips == len(flips), f"{label}: more flips than recorded ({n_flips})"
    for f in flips:
        kind = "box" if f["index"] == np.iinfo(np.uint64).max else f"particle {f['index']}"
        assert not f["gated"], f"{label}: sweep {f['sweep']} {kind}: the oracle rejected outright"
        assert f["margin"] <= f["bound"], (f"{label}: sweep {f['sweep']} {kind}: margin {f['margin']:.3e} "
                                           f"beyond the FP32 bound {f['bound']:.3e}")
    worst = float((flips["margin"] / flips["bound"]).max()) if len(flips) else 0.0
    print(f"[ties] {label}: {n_flips} flips, largest margin/bound {worst:.3f}", flush=True)
    return worst


def run_fp32_follow(oracle, chains, sweeps, chunk, block_sweeps, **opts):
    """FP32 engine run in chunks with the accept log; the oracle replays
    every chunk following the FP32 decisions.  Returns per chain: engine
    state, counters, samples, oracle chain, oracle samples, flips, n_flips."""
    eng = Engine(chains, form="batched_fp32", log_accepts=True, sample_capacity=chunk,
                 block_sweeps=block_sweeps, **opts)
    orcs = [oracle.Chain(oracle_spec(oracle, ch)) for ch in chains]
    got = [dict(smp=[], osmp=[], flips=[], nflip=0) for _ in chains]

    def follow(c, log, vacc, n):
        osmp, fl, nf = orcs[c].sweeps_follow(n, log, vacc, max_flips=4096)
        return c, osmp, fl, nf

    with cf.ThreadPoolExecutor(min(len(chains), 8)) as ex:
        done = 0
        while done < sweeps:
            n = min(chunk, sweeps - done)
            eng.run(n)
            logs = eng.read_accept_log(chunk)
            smp = eng.read_samples(chunk)
            futs = [ex.submit(follow, c, logs[c], smp[c]["acc_v"].astype(np.uint8), n)
                    for c in range(len(chains))]
            for fu in futs:
                c, osmp, fl, nf = fu.result()
                got[c]["smp"].append(smp[c].copy())
                got[c]["osmp"].append(osmp)
                got[c]["flips"].append(fl)
                got[c]["nflip"] += nf
            done += n
    for c in range(len(chains)):
        g = got[c]
        g["state"] = eng.state(c)
        g["cnt"] = eng.counters(c)
        g["orc"] = orcs[c]
        g["smp"] = np.concatenate(g["smp"])
        g["osmp"] = np.concatenate(g["osmp"])
        g["flips"] = np.concatenate(g["flips"])
    eng.close()
    return got


def check_replay(got, label, sweeps, u_tol=1e-5):
    """Same state bit for bit, same per-sweep decisions and lengths, every
    counter equal, running energies within u_tol relative."""
    total = 0
    for c, g in enumerate(got):
        lab = f"{label} chain {c}"
        total += g["nflip"]
        check_flips(g["flips"], g["nflip"], lab)
        x, L, U, sw = g["state"]
        ox, oL, oU, osw = g["orc"].state()
        assert sw == osw == sweeps, lab
        assert np.array_equal(x, ox), f"{lab}: positions differ"
        assert L == oL, lab
        np.testing.assert_array_equal(g["smp"]["acc_d"], g["osmp"]["acc_d"])
        np.testing.assert_array_equal(g["smp"]["acc_v"], g["osmp"]["acc_v"])
        np.testing.assert_array_equal(g["smp"]["L"], g["osmp"]["L"])
        du = np.abs(g["smp"]["U"] - g["osmp"]["U"]) / np.maximum(1.0, np.abs(g["osmp"]["U"]))
        assert du.max() <= u_tol, f"{lab}: running energy {du.max():.2e} from the oracle's"
        assert rel(U, oU) <= u_tol
        for k, ok in COUNTERS:
            assert g["cnt"][k] == g["orc"].counter(ok), (lab, k)
    return total


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_fp32_small_cases_ties(gpu, oracle, case):
    """Every small parity case, 40 sweeps: decisions equal to the oracle's
    except documented ties, the same state bit for bit."""
    chains = CASES[case]()
    print(f"[gpu] fp32 ties {case}", flush=True)
    got = run_fp32_follow(oracle, chains, 40, 40, 16)
    check_replay(got, case, 40)


@pytest.mark.gpu
def test_c3_sample_fp32_full_length_ties(gpu, oracle):
    """The six C3 chains of the FP64 full-length replay (the smallest,
    median and largest of each potential, N = 32 to 8176), FP32 form at the
    workload's full 1e4 sweeps: every flip a tie within its bound, state
    bit for bit equal to the oracle's after every chunk of 2500 sweeps."""
    sweeps, chunk = 10_000, 2_500
    c3 = configs.c3()
    pick = []
    for p in (nat.POT_LJ, nat.POT_HRSW):
        idx = sorted((i for i, c in enumerate(c3) if c.potential == p), key=lambda i: c3[i].N)
        pick += [idx[0], idx[len(idx) // 2], idx[-1]]
    chains = [c3[i] for i in pick]
    print(f"[gpu] C3 sample FP32 ties N={[c.N for c in chains]}", flush=True)
    got = run_fp32_follow(oracle, chains, sweeps, chunk, 500)
    n = check_replay(got, "c3 sample", sweeps)
    trials = sum(c.N for c in chains) * sweeps
    print(f"[ties] c3 sample: {n} flips in {trials:.3g} trials", flush=True)
    assert n <= 1e-5 * trials


@pytest.mark.gpu
def test_c3_random_chains_fp32_ties(gpu, oracle):
    """64 chains drawn at random (fixed seed) from the C3 workload, FP32
    form, 2000 sweeps: every flip a tie within its bound, state bit for bit
    equal to the oracle's after every chunk of 1000 sweeps."""
    c3 = configs.c3()
    pick = np.random.default_rng(41).choice(len(c3), 64, replace=False)
    chains = [c3[i] for i in pick]
    print(f"[gpu] C3 random 64 chains FP32 ties", flush=True)
    got = run_fp32_follow(oracle, chains, 2000, 1000, 500)
    n = check_replay(got, "c3 random", 2000)
    trials = sum(c.N for c in chains) * 2000
    print(f"[ties] c3 random: {n} flips in {trials:.3g} trials", flush=True)
    assert n <= 1e-5 * trials


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", ["c2", "c2x", "c4"])
def test_config_chain_fp32_full_length_ties(gpu, oracle, cfg):
    """One chain of C2 (4096 particles, NPT, 512 blocks, the 512-lane
    kernel), of C2 with the LJ exclusion distance at 0.75 sigma, or of C4
    (1024 particles, NVT at rho = 0.6, 128 blocks), FP32 form at the full
    1e4 sweeps: every flip a tie within its bound, state bit for bit equal
    to the oracle's after every chunk of 2500 sweeps."""
    if cfg == "c2x":
        ch = configs.c2(n_chains=1, exclusion=0.75)[0]
    else:
        ch = getattr(configs, cfg)(n_chains=1)[0]
    print(f"[gpu] {cfg} chain FP32 ties", flush=True)
    got = run_fp32_follow(oracle, [ch], 10_000, 2_500, 500)
    n = check_replay(got, f"{cfg} chain", 10_000)
    assert n <= 1e-5 * 10_000 * ch.N


@pytest.mark.gpu
def test_c5_chain_fp32_ties(gpu, oracle):
    """One chain of the C5 workload (65536 Yukawa particles, split into 4
    workgroups of 256 blocks), FP32 form, 200 sweeps in two launches: every
    flip a tie within its bound, state bit for bit equal to the oracle's."""
    ch = configs.c5(1)[0]
    print("[gpu] C5 chain FP32 ties", flush=True)
    got = run_fp32_follow(oracle, [ch], 200, 200, 100)
    n = check_replay(got, "c5 chain", 200)
    assert n <= 1e-5 * 200 * ch.N


def _full_size_properties(chains, sweeps, label):
    """Every chain of a workload in the FP32 form: the running energy within
    1e-5 relative of the FP64 recompute whenever one runs (max_drift), and of
    a fresh FP64 energy at the end; counters consistent with the samples."""
    eci = chains[0].eci
    assert eci and sweeps % eci != 0
    with Engine(chains, form="batched_fp32", sample_capacity=sweeps, block_sweeps=50) as e:
        e.run(sweeps)
        smp = e.read_samples(sweeps)
        cnt = e.all_counters()
        Uf = e.full_energy()
        _, L, U = e.states()
    Ns = np.array([c.N for c in chains], dtype=np.uint64)
    scale = np.maximum(1.0, np.abs(Uf))
    drift = cnt["max_drift"] / scale
    print(f"[gpu] {label}: largest relative drift at a recompute {drift.max():.2e}, "
          f"at the end {(np.abs(U - Uf) / scale).max():.2e}", flush=True)
    assert np.all(drift <= 1e-5), f"{label}: drift {drift.max():.2e}"
    assert np.all(np.abs(U - Uf) <= 1e-5 * scale)
    assert np.all(cnt["sweeps"] == sweeps)
    assert np.all(cnt["trials"] == Ns * np.uint64(sweeps))
    assert np.all(cnt["recomputes"] == sweeps // eci)
    assert np.all(cnt["vol_trials"] == sweeps)
    assert np.array_equal(cnt["acc_d"], smp["acc_d"].sum(axis=1, dtype=np.uint64))
    assert np.array_equal(cnt["acc_v"], smp["acc_v"].sum(axis=1, dtype=np.uint64))
    assert np.all(smp["acc_d"] <= Ns[:, None].astype(np.uint32))
    assert np.all(cnt["acc_d"] + cnt["wall_rejects"] + cnt["core_rejects"] <= cnt["trials"])
    assert np.array_equal(smp["L"][:, -1], L) and np.all(np.isfinite(smp["U"]))
    assert np.all(cnt["pair_terms"] > 0)


@pytest.mark.gpu
def test_c3_all_chains_fp32_properties(gpu):
    """All 4096 C3 chains at their measured precision for 250 sweeps
    (recomputes at 100 and 200)."""
    _full_size_properties(configs.c3(), 250, "c3 all chains")


@pytest.mark.gpu
def test_c5_all_chains_fp32_properties(gpu):
    """All 64 C5 chains (split chains, four workgroups each) for 150 sweeps
    (recompute at 100)."""
    _full_size_properties(configs.c5(), 150, "c5 all chains")
