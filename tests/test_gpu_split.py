"""Chains split over several cooperating workgroups (JMM_MODE_SPLIT), the
path for chains whose LDS image exceeds one workgroup (65536-particle
Yukawa chains).  The split is forced on small chains with the engine's
seg_lds option, at segment widths on both sides of one wavefront (the
cross-segment sums take a different order of operations there), and the
split chains are checked

  * against the same chains kept whole: samples, positions, box lengths,
    energies and counters bit-identical (FP64 and FP32 forms), over several
    launches, with whole chains sharing the engine;
  * against the oracle, move for move (FP64 form): acceptance logs and
    positions equal, energies to 1e-10.

Cases cover periodic and wall boundaries, NVT and NPT (box-length moves
across segment boundaries and across the periodic seam), the drift-guard
recompute, every potential, and one or two lanes per block."""
import math

import numpy as np
import pytest

from jmmonedmc_amd import _native as nat
from jmmonedmc_amd.engine import Chain, Engine

from helpers import hrsw, lj, oracle_spec, rel, yukawa

pytestmark = pytest.mark.gpu

SWEEPS = 24
BLOCK = 8          # three launches


def split_case(name):
    if name == "lj_periodic_nvt":
        return Chain(N=4096, rho=0.6, T=1.0, seed=21, chain_id=5, **lj())
    if name == "lj_periodic_npt_log":
        return Chain(N=3000, rho=0.5, T=1.0, P=0.5, ensemble=nat.NPT, vol_move=nat.VOL_LOG,
                     max_dv=0.01, seed=22, chain_id=6, eci=7, **lj())
    if name == "lj_walls_npt_linear":
        return Chain(N=2500, L0=5000.0, T=1.0, P=0.5, ensemble=nat.NPT, boundary=nat.WALLS,
                     vol_move=nat.VOL_LINEAR, max_dv=1.0, seed=23, eci=5, **lj(shifted=0))
    if name == "hrsw_npt":
        return Chain(N=3000, rho=0.5, T=1.0, P=1.0, ensemble=nat.NPT, max_dv=0.01, seed=24,
                     chain_id=8, **hrsw())
    if name == "yukawa_npt":
        return Chain(N=4096, rho=0.8, T=1.5, P=1.0, ensemble=nat.NPT, vol_move=nat.VOL_LOG,
                     max_dv=0.01, seed=25, chain_id=9, eci=10, **yukawa())
    if name == "lj_periodic_npt_b1024":
        # 1024 blocks: whole in one 1024-lane workgroup, split into 256-lane
        # segments of four waves each
        return Chain(N=16384, rho=0.6, T=1.0, P=0.5, ensemble=nat.NPT, vol_move=nat.VOL_LOG,
                     max_dv=0.01, seed=27, chain_id=10, eci=7, bcap=1024, **lj())
    if name == "harmonic_walls":
        return Chain(N=1024, L0=1024.0, T=0.4, P=0.7, ensemble=nat.NPT, boundary=nat.WALLS,
                     vol_move=nat.VOL_LINEAR, max_dv=1.0, max_step=0.1, nbn=1, seed=26,
                     potential=nat.POT_HARMONIC, cutoff=math.inf, pot_p=(0, 0, 0, 0), shifted=0)
    raise KeyError(name)


NAMES = ["lj_periodic_nvt", "lj_periodic_npt_log", "lj_walls_npt_linear", "hrsw_npt",
         "yukawa_npt", "harmonic_walls", "lj_periodic_npt_b1024"]


def seg_lds_for(oracle, ch, S):
    """seg_lds (bytes) under which the chain is split into exactly S
    segments: the LDS image of one segment of B/S blocks."""
    B, _, w = oracle.schedule(oracle_spec(oracle, ch))
    qb = ch.N // B
    smax = qb + (1 if ch.N % B else 0)
    Sp = smax | 1
    wg = max(w, 1)
    assert B % S == 0 and B // S <= 256, (B, S)
    return 8 * (2 * wg + (B // S) * Sp), B


def run(chains, form, seg_lds=0, log=False, lanes=0):
    with Engine(chains, form=form, sample_capacity=SWEEPS, block_sweeps=BLOCK,
                seg_lds=seg_lds, log_accepts=log, seg_lanes=lanes) as e:
        lay = [e.layout(c) for c in range(len(chains))]
        e.run(SWEEPS)
        logs = e.read_accept_log(SWEEPS) if log else None
        smp = e.read_samples(SWEEPS)
        xs, Ls, Us = e.states()
        cnt = e.all_counters()
    return dict(layout=lay, samples=smp, x=xs, L=Ls, U=Us, counters=cnt, logs=logs)


def assert_same(a, b):
    assert np.array_equal(a["samples"], b["samples"])
    for xa, xb in zip(a["x"], b["x"]):
        assert np.array_equal(xa, xb)
    assert np.array_equal(a["L"], b["L"]) and np.array_equal(a["U"], b["U"])
    for k in a["counters"]:
        assert np.array_equal(a["counters"][k], b["counters"][k]), k


def segment_counts(oracle, ch):
    """Split widths to test: 256 and 128 lanes (segment sums are sequences
    of wave sums) and 16 lanes (several segments per wave); for B <= 64 the
    tree-only case."""
    B, _, _ = oracle.schedule(oracle_spec(oracle, ch))
    out = []
    for Bs in (256, 128, 16):
        if Bs < B and B // Bs >= 2:
            out.append(B // Bs)
    if B <= 64 and B >= 4:
        out.append(4)
    return sorted(set(out))


@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("form", ["batched_fp64", "batched_fp32"])
@pytest.mark.parametrize("name", NAMES)
def test_split_chain_equals_whole_chain(gpu, oracle, name, form, lanes):
    """lanes = 2: two lanes per block share every trial (one walks the left
    partners, one the right)."""
    ch = split_case(name)
    # a whole chain in the same engine (its image stays under every seg_lds used)
    small = Chain(N=32, rho=0.5, T=1.0, seed=99, chain_id=99, **lj())
    chains = [small, ch]
    whole = run(chains, form)
    assert whole["layout"][1]["mode"] == "lds"
    counts = segment_counts(oracle, ch)
    assert counts, name
    for S in counts:
        seg_lds, B = seg_lds_for(oracle, ch, S)
        print(f"[gpu] split {name} {form} B={B} S={S} lanes={lanes}", flush=True)
        got = run(chains, form, seg_lds=seg_lds, lanes=lanes)
        lay = got["layout"][1]
        assert lay["mode"] == "split" and lay["segments"] == S and lay["lanes"] == lanes, lay
        assert got["layout"][0]["mode"] == "lds"
        assert_same(got, whole)


@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("name", NAMES)
def test_split_chain_replays_oracle(gpu, oracle, name, lanes):
    ch = split_case(name)
    S = segment_counts(oracle, ch)[-1]
    seg_lds, B = seg_lds_for(oracle, ch, S)
    print(f"[gpu] split oracle {name} B={B} S={S} lanes={lanes}", flush=True)
    got = run([ch], "batched_fp64", seg_lds=seg_lds, log=True, lanes=lanes)
    assert got["layout"][0]["mode"] == "split"
    orc = oracle.Chain(oracle_spec(oracle, ch))
    _, olog = orc.sweeps(SWEEPS, log=True)
    ox, oL, oU, osw = orc.state()
    assert np.array_equal(got["logs"][0].ravel(), olog)
    assert np.array_equal(got["x"][0], ox) and got["L"][0] == oL
    assert rel(got["U"][0], oU) <= 1e-10
    assert got["counters"]["sweeps"][0] == osw == SWEEPS
    assert got["counters"]["trials"][0] == orc.counter("trials")
    assert got["counters"]["acc_d"][0] == orc.counter("acc_d")


def test_c5_chain_replays_oracle(gpu, oracle):
    """A chain of the 65536-particle Yukawa workload, split because its LDS
    image (0.53 MB) exceeds a workgroup's: 4 segments of 256 blocks, FP64
    form, two launches, equal to the oracle move for move."""
    from jmmonedmc_amd.configs import c5
    ch = c5(1)[0]
    print("[gpu] split oracle c5 chain N=65536", flush=True)
    with Engine([ch], form="batched_fp64", sample_capacity=6, block_sweeps=3, log_accepts=True) as e:
        lay = e.layout(0)
        assert lay["mode"] == "split" and lay["segments"] == 4 and lay["B"] == 1024, lay
        e.run(6)
        log = e.read_accept_log(6)[0].ravel()
        smp = e.read_samples(6)[0]
        x, L, U, sw = e.state(0)
    orc = oracle.Chain(oracle_spec(oracle, ch))
    osmp, olog = orc.sweeps(6, log=True)
    ox, oL, oU, osw = orc.state()
    assert np.array_equal(log, olog)
    assert np.array_equal(x, ox) and L == oL and sw == osw == 6
    assert rel(U, oU) <= 1e-10
    assert np.array_equal(smp["acc_v"], osmp["acc_v"])


def test_c5_chain_long_replays_oracle(gpu, oracle):
    """The same split C5 chain for 200 sweeps in two 100-sweep launches
    (about 600 neighbour syncs per segment, box moves accepted), FP64 form:
    every sweep's accepts and box length, the final positions bit for bit,
    every counter equal to the oracle's."""
    from jmmonedmc_amd.configs import c5
    ch = c5(1)[0]
    print("[gpu] split oracle c5 chain N=65536, 200 sweeps", flush=True)
    with Engine([ch], form="batched_fp64", sample_capacity=200, block_sweeps=100) as e:
        assert e.layout(0)["mode"] == "split"
        e.run(200)
        smp = e.read_samples(200)[0]
        x, L, U, sw = e.state(0)
        cnt = e.counters(0)
    orc = oracle.Chain(oracle_spec(oracle, ch))
    osmp = orc.sweeps(200)
    ox, oL, oU, osw = orc.state()
    assert np.array_equal(smp["acc_d"], osmp["acc_d"])
    assert np.array_equal(smp["acc_v"], osmp["acc_v"]) and osmp["acc_v"].sum() > 0
    assert np.array_equal(smp["L"], osmp["L"])
    assert np.array_equal(x, ox) and L == oL and sw == osw == 200
    assert rel(U, oU) <= 1e-10
    for k, ok in (("trials", "trials"), ("acc_d", "acc_d"), ("acc_v", "acc_v"), ("core_rejects", "hc_rej"),
                  ("pair_terms", "pair_terms"), ("recomputes", "recomputes")):
        assert cnt[k] == orc.counter(ok), k


def test_every_placement_in_one_engine_replays_oracle(gpu, oracle):
    """One engine holding a chain that runs from HBM (24000 particles with
    one block: too large for a workgroup and nothing to split), a chain in
    LDS and a split chain, FP64 form: each equal to the oracle move for
    move (the three launch groups run one after another on one stream)."""
    hbm = Chain(N=24000, rho=0.5, T=1.0, P=0.5, ensemble=nat.NPT, vol_move=nat.VOL_LOG,
                max_dv=0.01, seed=31, chain_id=31, bcap=1, **lj())
    whole = Chain(N=64, L0=128.0, T=1.0, seed=32, chain_id=32, **lj())
    split = split_case("lj_periodic_npt_log")
    seg_lds, _ = seg_lds_for(oracle, split, 16)
    chains = [hbm, whole, split]
    print("[gpu] placements: hbm + lds + split", flush=True)
    with Engine(chains, form="batched_fp64", sample_capacity=4, block_sweeps=2, seg_lds=seg_lds,
                log_accepts=True) as e:
        modes = [e.layout(c)["mode"] for c in range(3)]
        assert modes == ["hbm", "lds", "split"], modes
        e.run(4)
        logs = e.read_accept_log(4)
        smp = e.read_samples(4)
        states = [e.state(c) for c in range(3)]
    for c, ch in enumerate(chains):
        orc = oracle.Chain(oracle_spec(oracle, ch))
        osmp, olog = orc.sweeps(4, log=True)
        ox, oL, oU, osw = orc.state()
        x, L, U, sw = states[c]
        assert np.array_equal(logs[c].ravel(), olog), c
        assert np.array_equal(x, ox) and L == oL and sw == osw == 4, c
        assert rel(U, oU) <= 1e-10, c
        assert np.array_equal(smp[c]["acc_v"], osmp["acc_v"]), c


def test_whole_wave_chains_stay_whole(gpu, oracle):
    """Chains whose block count is a multiple of 64 but not a power of two
    (bcap 448) are never split: one too large for a workgroup's LDS runs
    from HBM (40000 particles, 448 blocks), one that a small split
    threshold would split otherwise stays in LDS (4096 particles); FP64
    form, each equal to the oracle move for move."""
    big = Chain(N=40000, rho=0.5, T=1.0, P=0.5, ensemble=nat.NPT, vol_move=nat.VOL_LOG,
                max_dv=0.01, seed=41, chain_id=41, bcap=448, **lj())
    mid = Chain(N=4096, rho=0.5, T=1.0, P=0.5, ensemble=nat.NPT, vol_move=nat.VOL_LOG,
                max_dv=0.01, seed=42, chain_id=42, bcap=448, **lj())
    chains = [big, mid]
    print("[gpu] whole-wave chains: hbm + lds under a 16 KiB split threshold", flush=True)
    with Engine(chains, form="batched_fp64", sample_capacity=4, block_sweeps=2, seg_lds=16384,
                log_accepts=True) as e:
        lay = [e.layout(c) for c in range(2)]
        assert [l["mode"] for l in lay] == ["hbm", "lds"], lay
        assert [l["B"] for l in lay] == [448, 448], lay
        e.run(4)
        logs = e.read_accept_log(4)
        smp = e.read_samples(4)
        states = [e.state(c) for c in range(2)]
    for c, ch in enumerate(chains):
        orc = oracle.Chain(oracle_spec(oracle, ch))
        osmp, olog = orc.sweeps(4, log=True)
        ox, oL, oU, osw = orc.state()
        x, L, U, sw = states[c]
        assert np.array_equal(logs[c].ravel(), olog), c
        assert np.array_equal(x, ox) and L == oL and sw == osw == 4, c
        assert rel(U, oU) <= 1e-10, c
        assert np.array_equal(smp[c]["acc_v"], osmp["acc_v"]), c
