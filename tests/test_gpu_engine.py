"""Engine services on the GPU, through the C ABI: restart from a saved state
(printCoords/restart), the drift guard (ECheck) with an injected energy
error, the sample buffer's full condition, graph replay against eager
launches, and the bulk counters against the per-chain ones."""
import numpy as np
import pytest

from jmmonedmc_amd._native import JmmError
from jmmonedmc_amd.engine import Chain, Engine

from helpers import CASES, lj, oracle_spec, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["plain_fp64", "batched_fp64"])
def test_restart_from_saved_state_replays_oracle(gpu, oracle, form):
    """Run 12 sweeps, save, restore into a fresh engine at sweep 12 and run 20
    more: positions and decisions equal the oracle's uninterrupted run."""
    print(f"[gpu] restart {form}", flush=True)
    chains = CASES["npt_periodic_log"]()
    with Engine(chains, form=form, sample_capacity=64, block_sweeps=8) as a:
        a.run(12)
        saved = [a.state(c) for c in range(len(chains))]
    with Engine(chains, form=form, sample_capacity=64, block_sweeps=8, log_accepts=True) as b:
        for c, (x, L, U, sw) in enumerate(saved):
            assert sw == 12
            b.set_state(c, x, L, sw)
        b.run(20)
        logs = b.read_accept_log(64)
        for c, ch in enumerate(chains):
            orc = oracle.Chain(oracle_spec(oracle, ch))
            orc.sweeps(12)
            _, olog = orc.sweeps(20, log=True)
            x, L, U, sw = b.state(c)
            ox, oL, oU, osw = orc.state()
            assert sw == osw == 32
            assert np.array_equal(logs[c].ravel(), olog)
            assert np.array_equal(x, ox) and L == oL
            assert rel(U, oU) <= 1e-10


def test_set_state_rejects_invalid_configurations(gpu):
    ch = Chain(N=16, L0=32.0, seed=1, **lj())
    with Engine([ch], form="batched_fp64") as e:
        x, L, _, _ = e.state(0)
        bad = x.copy()
        bad[3], bad[4] = bad[4], bad[3]
        with pytest.raises(JmmError, match="unsorted or inside the hard core"):
            e.set_state(0, bad, L)
        with pytest.raises(JmmError, match="bad length"):
            e.set_state(0, x, 6.0)                    # periodic length must exceed 2 (rc + step)
        e.set_state(0, x + 0.25, L, 5)
        x2, L2, U2, sw2 = e.state(0)
        assert np.array_equal(x2, x + 0.25) and sw2 == 5


@pytest.mark.parametrize("form", ["plain_fp64", "batched_fp64", "batched_fp32"])
def test_drift_guard_repairs_injected_energy_error(gpu, form):
    """eci = 10: an error of +0.5 injected into chain 1's running energy after
    sweep 10 is found by the recompute at sweep 20 and replaced by the FP64
    energy; the other chains see only rounding-level drift."""
    print(f"[gpu] drift guard {form}", flush=True)
    chains = [Chain(N=96, rho=0.5, T=1.0, seed=2, chain_id=c, eci=10, **lj()) for c in range(3)]
    with Engine(chains, form=form, sample_capacity=64, block_sweeps=10) as e:
        e.run(10)
        e.sync()
        e.perturb_energy(1, 0.5)
        e.run(10)
        smp = e.read_samples(64)
        cnt = e.all_counters()
        U = e.full_energy()
        assert cnt["recomputes"].tolist() == [2, 2, 2]
        assert 0.5 - 1e-4 <= cnt["max_drift"][1] <= 0.5 + 1e-4
        tol = 1e-9 if form != "batched_fp32" else 1e-4
        assert cnt["max_drift"][0] <= tol and cnt["max_drift"][2] <= tol
        for c in range(3):
            assert rel(cnt["U"][c], U[c]) <= 1e-12
        # sweeps 11..19 carry the injected offset; sweep 20 is the repaired value
        assert abs(smp["U"][1, 19] - U[1]) <= 1e-9 * max(1.0, abs(U[1]))


def test_sample_buffer_full_and_drain(gpu):
    ch = Chain(N=32, L0=64.0, seed=4, **lj())
    with Engine([ch], form="batched_fp64", sample_capacity=20, block_sweeps=10) as e:
        e.run(20)
        with pytest.raises(JmmError, match="read samples first"):
            e.run(1)
        s = e.read_samples(20)
        assert s.shape == (1, 20)
        e.run(5)
        s2 = e.read_samples(20)
        assert s2.shape == (1, 5)
        assert e.counters(0)["sweeps"] == 25


@pytest.mark.parametrize("form", ["plain_fp64", "batched_fp64", "batched_fp32"])
def test_graph_replay_equals_eager(gpu, form):
    """use_graph replays the captured sweep block: after four blocks with a
    sample read and host synchronisation between them, every replay's
    samples and the final state equal the eager engine's bit for bit."""
    print(f"[gpu] graph replay {form}", flush=True)
    chains = CASES["mixed_batch"]()
    out = []
    for graph in (False, True):
        with Engine(chains, form=form, sample_capacity=16, block_sweeps=16, use_graph=graph) as e:
            rows = []
            for _ in range(4):
                e.run(16)
                e.sync()
                rows.append(e.read_samples(16))
            xs, Ls, Us = e.states()
            out.append((rows, np.concatenate(xs), Ls, Us))
    (ra, xa, La, Ua), (rb, xb, Lb, Ub) = out
    for a, b in zip(ra, rb):
        assert np.array_equal(a, b)
    assert np.array_equal(xa, xb) and np.array_equal(La, Lb) and np.array_equal(Ua, Ub)


def test_bulk_counters_equal_per_chain(gpu):
    chains = CASES["mixed_batch"]()
    with Engine(chains, form="batched_fp32", sample_capacity=32) as e:
        e.run(24)
        allc = e.all_counters()
        xs, Ls, Us = e.states()
        for c in range(len(chains)):
            one = e.counters(c)
            for k, v in one.items():
                assert allc[k][c] == v, (c, k)
            x, L, U, _ = e.state(c)
            assert np.array_equal(xs[c], x) and Ls[c] == L and Us[c] == U


@pytest.mark.parametrize("case", ["c1_walls_truncated", "npt_periodic_log", "mixed_batch"])
def test_histograms_equal_oracle_counts(gpu, oracle, case):
    """rho(x)/g(x) sampling every 5 sweeps: the summed counts equal the
    oracle's all-pairs binning of the same configurations, bit for bit
    (the oracle replays the chains move for move in FP64)."""
    print(f"[gpu] histograms {case}", flush=True)
    chains = CASES[case]()
    hs = dict(rhonb=40, rbw=2.5, gns=3, gsw=20.0, gnb=30, gbw=0.25)
    with Engine(chains, form="batched_fp64", sample_capacity=64, block_sweeps=16) as e:
        e.hist_configure(5, **hs)
        e.run(20)
        rho, g, ns = e.hist_read()
        assert ns == 4
        for c, ch in enumerate(chains):
            orc = oracle.Chain(oracle_spec(oracle, ch))
            want_r = np.zeros(hs["rhonb"], dtype=np.uint64)
            want_g = np.zeros((hs["gns"], hs["gnb"]), dtype=np.uint64)
            for _ in range(4):
                orc.sweeps(5)
                x, L, _, _ = orc.state()
                r1, g1 = oracle.hist_x(orc.spec, x, L, hs["rhonb"], hs["rbw"], hs["gns"],
                                       hs["gsw"], hs["gnb"], hs["gbw"])
                want_r += r1
                want_g += g1
            assert np.array_equal(rho[c], want_r), c
            assert np.array_equal(g[c], want_g), c
            assert int(rho[c].sum()) <= 4 * ch.N
        rho2, g2, ns2 = e.hist_read(reset=True)
        assert ns2 == 4 and np.array_equal(rho2, rho)
        _, _, ns3 = e.hist_read()
        assert ns3 == 0


@pytest.mark.parametrize("form", ["plain_fp64", "batched_fp64"])
def test_step_sizes_changed_between_sweeps_replay_oracle(gpu, oracle, form):
    """jmm_set_steps (step adaptation support): new displacement and box-move
    half-widths after 10 sweeps; the next 15 sweeps equal the oracle's with
    the same change, move for move.  A step above the creation value is
    refused (the schedule is built for it)."""
    print(f"[gpu] set_steps {form}", flush=True)
    chains = CASES["npt_periodic_log"]()
    n = len(chains)
    new_step = [0.2 + 0.05 * c for c in range(n)]
    new_dv = [0.003 * (c + 1) for c in range(n)]
    with Engine(chains, form=form, sample_capacity=32, block_sweeps=8, log_accepts=True) as e:
        e.run(10)
        e.read_accept_log(32)
        e.read_samples(32)
        e.set_steps(new_step, new_dv)
        with pytest.raises(JmmError, match="creation value"):
            e.set_steps([ch.max_step * 1.5 for ch in chains])
        e.run(15)
        logs = e.read_accept_log(32)
        for c, ch in enumerate(chains):
            orc = oracle.Chain(oracle_spec(oracle, ch))
            orc.sweeps(10)
            orc.set_steps(new_step[c], new_dv[c])
            _, olog = orc.sweeps(15, log=True)
            x, L, U, _ = e.state(c)
            ox, oL, oU, _ = orc.state()
            assert np.array_equal(logs[c].ravel(), olog), c
            assert np.array_equal(x, ox) and L == oL
            assert rel(U, oU) <= 1e-10


@pytest.mark.parametrize("case", ["c1_walls_truncated", "npt_periodic_log", "mixed_batch"])
def test_scaled_energy_and_set_lengths(gpu, oracle, case):
    """jmm_scaled_energy (calculateEnergyOfTrialVolumeChange) and
    jmm_set_lengths (moveVolume): energies of configurations scaled by 0.97
    and 1.04 equal the oracle's all-pairs energy of x * s at length L s;
    after set_lengths the positions are x * (L'/L) bit for bit."""
    print(f"[gpu] scaled energy {case}", flush=True)
    chains = CASES[case]()
    with Engine(chains, form="batched_fp64", sample_capacity=32, block_sweeps=8) as e:
        e.run(12)
        xs, Ls, Us = e.states()
        for s in (0.97, 1.04):
            U = e.scaled_energy(Ls * s)
            for c, ch in enumerate(chains):
                spec = oracle_spec(oracle, ch)
                sc = (Ls[c] * s) / Ls[c]
                want = oracle.brute_energy_x(spec, xs[c] * sc, Ls[c] * s)
                if np.isinf(want):                      # a hard core overlaps
                    assert U[c] == want, (c, s, U[c])
                    continue
                assert rel(U[c], want) <= 1e-12 or abs(U[c] - want) <= 1e-12, (c, s, U[c], want)
        _, Ls2, Us2 = e.states()
        assert np.array_equal(Ls2, Ls) and np.array_equal(Us2, Us)     # state untouched
        Lnew = Ls * 1.02
        e.set_lengths(Lnew)
        xs3, Ls3, Us3 = e.states()
        for c, ch in enumerate(chains):
            assert np.array_equal(xs3[c], xs[c] * (Lnew[c] / Ls[c]))
            assert Ls3[c] == Lnew[c]
            spec = oracle_spec(oracle, ch)
            assert rel(Us3[c], oracle.brute_energy_x(spec, xs3[c], Lnew[c])) <= 1e-12
        with pytest.raises(JmmError, match="hard core|bad length|walls"):
            e.set_lengths(Lnew * 0.2)


def test_autotuner_candidates_give_identical_chains(gpu, tmp_path):
    """The tuned parameters only change packing and launch: every candidate
    produces bit-identical samples and states (FP32 form, mixed batch), and
    the tuner returns one of them, cached by the chain-mix signature."""
    from jmmonedmc_amd.tune import candidates_for, tune
    print("[gpu] autotuner invariance", flush=True)
    chains = CASES["mixed_batch"]() + [Chain(N=2000, L0=4000.0, seed=9, chain_id=77, **lj())]
    cands = candidates_for(chains)
    assert any(c["wg_threads"] == 1024 for c in cands)
    ref = None
    for cand in cands:
        with Engine(chains, form="batched_fp32", sample_capacity=40, block_sweeps=20, **cand) as e:
            e.run(40)
            smp = e.read_samples(40)
            xs, Ls, Us = e.states()
            got = (smp, np.concatenate(xs), Ls, Us)
        if ref is None:
            ref = got
            continue
        assert np.array_equal(got[0], ref[0]), cand
        assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2]), cand
    cache = str(tmp_path / "tune.json")
    r1 = tune(chains, block_sweeps=20, blocks=1, cache=cache)
    assert r1.best in cands and len(r1.ms_per_sweep) == len(cands)
    r2 = tune(chains, block_sweeps=20, blocks=1, cache=cache)
    assert r2.best == r1.best and r2.ms_per_sweep == r1.ms_per_sweep


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["batched_fp64", "batched_fp32"])
def test_observables_on_device_equal_host(gpu, form):
    """jmm_observables_device, queued after a sweep block on the engine's
    stream, writes the bits jmm_observables copies to the host."""
    import torch
    chains = CASES["mixed_batch"]()
    print(f"[gpu] observables on device {form}", flush=True)
    # torch's device first (its own HIP runtime), then the engine's
    obs_d = torch.full((len(chains) * 6,), float("nan"), dtype=torch.float64, device="cuda")
    grad_d = torch.full((int(sum(c.N for c in chains)),), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with Engine(chains, form=form, sample_capacity=10, block_sweeps=10) as e:
        e.run(10)
        e.observables_device(obs_d.data_ptr(), grad_d.data_ptr())
        e.sync()
        obs, grads = e.observables(want_grad=True)
    assert np.array_equal(obs_d.cpu().numpy().reshape(-1, 6), obs)
    assert np.array_equal(grad_d.cpu().numpy(), np.concatenate(grads))
