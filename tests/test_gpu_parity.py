"""GPU parity: the HIP kernels against the CPU oracle on the same seeded
inputs (same Philox stream, same block schedule).

FP64 forms: accept/reject decisions equal move for move, positions equal bit
for bit, energies within 1e-10 relative.  FP32 form: decisions equal on these
sizes (a flip needs a near-threshold tie, |ln u + beta dE| below the FP32
error of dE, about 1e-7 per trial), energies within 1e-5 relative.
"""
import numpy as np
import pytest

from jmmonedmc_amd.engine import Engine

from helpers import CASES, oracle_spec, rel

SWEEPS = 40


def run_pair(oracle, chains, form, sweeps=SWEEPS):
    print(f"[gpu] form={form} chains={len(chains)} N={[c.N for c in chains][:6]} sweeps={sweeps}",
          flush=True)
    eng = Engine(chains, form=form, log_accepts=True, sample_capacity=sweeps, block_sweeps=16)
    eng.run(sweeps)
    logs = eng.read_accept_log(sweeps)
    smp = eng.read_samples(sweeps)
    out = []
    for c, ch in enumerate(chains):
        orc = oracle.Chain(oracle_spec(oracle, ch))
        osmp, olog = orc.sweeps(sweeps, log=True)
        out.append((eng.state(c), orc.state(), logs[c].ravel(), olog, smp[c], osmp,
                    eng.counters(c), orc))
    return eng, out


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain_fp64", "batched_fp64"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_fp64_move_for_move(gpu, oracle, form, case):
    chains = CASES[case]()
    eng, res = run_pair(oracle, chains, form)
    for c, ((x, L, U, sw), (ox, oL, oU, osw), lg, olg, smp, osmp, cnt, orc) in enumerate(res):
        assert sw == osw == SWEEPS
        mism = np.flatnonzero(lg != olg)
        assert mism.size == 0, f"chain {c}: first decision mismatch at trial {mism[:5]}"
        assert np.array_equal(x, ox), f"chain {c}: positions differ"
        assert L == oL
        assert rel(U, oU) <= 1e-10
        np.testing.assert_array_equal(smp["acc_d"], osmp["acc_d"])
        np.testing.assert_array_equal(smp["acc_v"], osmp["acc_v"])
        np.testing.assert_array_equal(smp["L"], osmp["L"])
        assert np.all(np.abs(smp["U"] - osmp["U"]) <= 1e-10 * np.maximum(1, np.abs(osmp["U"])))
        assert cnt["trials"] == orc.counter("trials")
        assert cnt["acc_d"] == orc.counter("acc_d")
        assert cnt["core_rejects"] == orc.counter("hc_rej")
        assert cnt["wall_rejects"] == orc.counter("wall_rej")
        assert cnt["pair_terms"] == orc.counter("pair_terms")
    eng.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_fp32_decisions_and_energy(gpu, oracle, case):
    chains = CASES[case]()
    eng, res = run_pair(oracle, chains, "batched_fp32")
    for c, ((x, L, U, sw), (ox, oL, oU, osw), lg, olg, smp, osmp, cnt, orc) in enumerate(res):
        mism = np.flatnonzero(lg != olg)
        assert mism.size == 0, f"chain {c}: FP32 decision flips at {mism[:5]}"
        assert np.array_equal(x, ox)
        assert rel(U, oU) <= 1e-5
        # cutoff decisions are FP64 in both forms: the same pair terms
        assert cnt["pair_terms"] == orc.counter("pair_terms")
    eng.close()


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain_fp64", "batched_fp64", "batched_fp32"])
def test_full_energy_matches_brute_force(gpu, oracle, form):
    chains = CASES["mixed_batch"]()
    eng = Engine(chains, form=form, sample_capacity=64)
    eng.run(30)
    eng.sync()
    Uf = eng.full_energy()
    tol_running = 1e-10 if form != "batched_fp32" else 1e-5
    for c, ch in enumerate(chains):
        x, L, U, _ = eng.state(c)
        ub = oracle.brute_energy_x(oracle_spec(oracle, ch), x, L)
        assert rel(Uf[c], ub) <= 1e-12
        assert rel(U, ub) <= tol_running
    eng.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["c1_periodic_nvt", "c1_walls_truncated"])
def test_c1_full_length_replays_oracle(gpu, oracle, case):
    """C1 at its full 1e5 sweeps (6.4e6 trials), batched FP64: every
    sweep's accepted-move count, the final positions bit for bit, the
    running energy and every counter equal to the oracle's."""
    sweeps, chunk = 100_000, 10_000
    ch = CASES[case]()[0]
    print(f"[gpu] C1 full length {case}", flush=True)
    eng = Engine([ch], form="batched_fp64", sample_capacity=chunk, block_sweeps=1000)
    acc = []
    for _ in range(sweeps // chunk):
        eng.run(chunk)
        acc.append(eng.read_samples(chunk)[0]["acc_d"].copy())
    x, L, U, sw = eng.state(0)
    cnt = eng.counters(0)
    eng.close()
    orc = oracle.Chain(oracle_spec(oracle, ch))
    osmp = orc.sweeps(sweeps)
    ox, oL, oU, osw = orc.state()
    assert sw == osw == sweeps
    np.testing.assert_array_equal(np.concatenate(acc), osmp["acc_d"])
    assert np.array_equal(x, ox) and L == oL
    assert rel(U, oU) <= 1e-10
    for k, ok in (("trials", "trials"), ("acc_d", "acc_d"), ("wall_rejects", "wall_rej"),
                  ("core_rejects", "hc_rej"), ("pair_terms", "pair_terms"), ("recomputes", "recomputes")):
        assert cnt[k] == orc.counter(ok), k


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", ["c2", "c2x", "c4"])
def test_config_chain_full_length_replays_oracle(gpu, oracle, cfg):
    """One chain of C2 (4096 particles, NPT on a ring, ln L moves), of C2
    with the LJ exclusion distance at 0.75 sigma (512 blocks, the 512-lane
    kernel) or of C4 (1024 particles, NVT at rho = 0.6) at the
    configuration's full 1e4 sweeps, batched FP64: every sweep's accepts and
    box length, the final positions bit for bit, every counter."""
    from jmmonedmc_amd import configs
    sweeps, chunk = 10_000, 2_500
    if cfg == "c2x":
        ch = configs.c2(n_chains=1, exclusion=0.75, bcap=1024)[0]
    else:
        ch = getattr(configs, cfg)(n_chains=1)[0]
    print(f"[gpu] {cfg} chain full length", flush=True)
    eng = Engine([ch], form="batched_fp64", sample_capacity=chunk, block_sweeps=500)
    if cfg == "c2x":
        assert eng.layout(0)["B"] == 512
    parts = []
    for _ in range(sweeps // chunk):
        eng.run(chunk)
        parts.append(eng.read_samples(chunk)[0].copy())
    smp = np.concatenate(parts)
    x, L, U, sw = eng.state(0)
    cnt = eng.counters(0)
    eng.close()
    orc = oracle.Chain(oracle_spec(oracle, ch))
    osmp = orc.sweeps(sweeps)
    ox, oL, oU, osw = orc.state()
    np.testing.assert_array_equal(smp["acc_d"], osmp["acc_d"])
    np.testing.assert_array_equal(smp["acc_v"], osmp["acc_v"])
    np.testing.assert_array_equal(smp["L"], osmp["L"])
    assert osmp["acc_v"].sum() > 0 or cfg == "c4"
    assert np.array_equal(x, ox) and L == oL
    assert rel(U, oU) <= 1e-10
    for k, ok in (("trials", "trials"), ("acc_d", "acc_d"), ("acc_v", "acc_v"), ("core_rejects", "hc_rej"),
                  ("pair_terms", "pair_terms"), ("recomputes", "recomputes")):
        assert cnt[k] == orc.counter(ok), k


@pytest.mark.gpu
def test_c3_sample_full_length_replays_oracle(gpu, oracle):
    """Six chains of the C3 workload (the smallest, median and largest of
    each potential, N = 32 to 8176) in one engine at the workload's full 1e4
    sweeps, batched FP64: every sweep's accepts and box length, the final
    positions bit for bit, every counter."""
    import concurrent.futures as cf
    from jmmonedmc_amd import _native as nat
    from jmmonedmc_amd import configs
    sweeps, chunk = 10_000, 2_500
    c3 = configs.c3()
    pick = []
    for p in (nat.POT_LJ, nat.POT_HRSW):
        idx = sorted((i for i, c in enumerate(c3) if c.potential == p), key=lambda i: c3[i].N)
        pick += [idx[0], idx[len(idx) // 2], idx[-1]]
    chains = [c3[i] for i in pick]
    print(f"[gpu] C3 sample full length N={[c.N for c in chains]}", flush=True)
    eng = Engine(chains, form="batched_fp64", sample_capacity=chunk, block_sweeps=500)
    parts = []
    for _ in range(sweeps // chunk):
        eng.run(chunk)
        parts.append(eng.read_samples(chunk).copy())
    smp = np.concatenate(parts, axis=1)
    states = [eng.state(c) for c in range(len(chains))]
    cnts = [eng.counters(c) for c in range(len(chains))]
    eng.close()

    def replay(ch):
        orc = oracle.Chain(oracle_spec(oracle, ch))
        return orc, orc.sweeps(sweeps)

    with cf.ThreadPoolExecutor(len(chains)) as ex:
        ref = list(ex.map(replay, chains))
    for c, (orc, osmp) in enumerate(ref):
        x, L, U, sw = states[c]
        ox, oL, oU, osw = orc.state()
        np.testing.assert_array_equal(smp[c]["acc_d"], osmp["acc_d"])
        np.testing.assert_array_equal(smp[c]["acc_v"], osmp["acc_v"])
        np.testing.assert_array_equal(smp[c]["L"], osmp["L"])
        assert np.array_equal(x, ox) and L == oL, c
        assert rel(U, oU) <= 1e-10
        for k, ok in (("trials", "trials"), ("acc_d", "acc_d"), ("acc_v", "acc_v"),
                      ("core_rejects", "hc_rej"), ("pair_terms", "pair_terms"), ("recomputes", "recomputes")):
            assert cnts[c][k] == orc.counter(ok), (c, k)


@pytest.mark.gpu
def test_c3_random_chains_replay_oracle(gpu, oracle):
    """64 chains drawn at random (fixed seed) from the C3 workload, so the
    whole (N, T, P) range of both potentials is sampled, in one engine for
    2000 sweeps, batched FP64: every sweep's accepts and box length, the
    final positions bit for bit, the energies and every counter."""
    import concurrent.futures as cf
    from jmmonedmc_amd import configs
    sweeps = 2000
    c3 = configs.c3()
    pick = np.random.default_rng(41).choice(len(c3), 64, replace=False)
    chains = [c3[i] for i in pick]
    print(f"[gpu] C3 random 64 chains, {sum(c.N for c in chains)} particles", flush=True)
    with Engine(chains, form="batched_fp64", sample_capacity=sweeps, block_sweeps=500) as eng:
        eng.run(sweeps)
        smp = eng.read_samples(sweeps)
        states = [eng.state(c) for c in range(len(chains))]
        cnts = [eng.counters(c) for c in range(len(chains))]

    def replay(ch):
        orc = oracle.Chain(oracle_spec(oracle, ch))
        return orc, orc.sweeps(sweeps)

    with cf.ThreadPoolExecutor(8) as ex:
        ref = list(ex.map(replay, chains))
    for c, (orc, osmp) in enumerate(ref):
        x, L, U, sw = states[c]
        ox, oL, oU, osw = orc.state()
        np.testing.assert_array_equal(smp[c]["acc_d"], osmp["acc_d"])
        np.testing.assert_array_equal(smp[c]["acc_v"], osmp["acc_v"])
        np.testing.assert_array_equal(smp[c]["L"], osmp["L"])
        assert np.array_equal(x, ox) and L == oL, c
        assert rel(U, oU) <= 1e-10
        for k, ok in (("trials", "trials"), ("acc_d", "acc_d"), ("acc_v", "acc_v"),
                      ("core_rejects", "hc_rej"), ("pair_terms", "pair_terms"), ("recomputes", "recomputes")):
            assert cnts[c][k] == orc.counter(ok), (c, k)
