"""Benchmark: trial moves/s of the batched FP32 engine on one MI355X, beside
the plain FP64 GPU form and the FP64 CPU oracle, on the 4096-chain mixed
workload (C3).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3]

A step is one sweep block (--block-sweeps sweeps, default 100) of every chain;
the default K=100 steps is the workload's full 1e4 sweeps.  Prints one JSON
line.  Other workloads (--workload): c2 (and its variants c2x, c2b), c5,
and c4, whose steps also run the backward (U, W, dU/dtheta, dW/dtheta,
dU/dx) of every chain after each block into device arrays, timed with the
sweeps; "backward" gives its own rate, on the device and with the results
read back to the host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "trial moves/s on 1 MI355X, batched FP32 vs plain FP64, at ≤3σ ⟨U⟩ agreement"
PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}
# v_exp_f32 / v_rcp_f32 / v_sqrt_f32 issue at a quarter of the FP32 lane rate:
# 256 CUs x 64 lanes / 4 per clock x 2.4 GHz
PEAK_TRANS_TOPS = 256 * 64 / 4 * 2.4e9 / 1e12


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--block-sweeps", type=int, default=100)
    p.add_argument("--workload", default="c3", choices=["c3", "c2", "c2x", "c2b", "c4", "c5"])
    p.add_argument("--plain-steps", type=int, default=1, help="sweep blocks timed for the plain FP64 form")
    p.add_argument("--forms", default="fp32,fp64,plain",
                   help="comma list of forms to time: fp32 (batched FP32, the value), fp64 "
                        "(batched FP64), plain (plain FP64)")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="only run a device-to-device copy of MIB MiB (HBM counter calibration)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-sweeps", type=int, default=0, help="CPU baseline sweeps (0 = auto ~15 s)")
    p.add_argument("--graph", type=int, default=1)
    p.add_argument("--tune", type=int, default=1,
                   help="choose the batched forms' launch parameters with the autotuner first")
    p.add_argument("--acc-burnin", type=int, default=-1,
                   help="FP64 sweeps that equilibrate the common start of the accuracy runs "
                        "(-1 = the workload's sweep count: 1e4, C5 2e3)")
    p.add_argument("--acc-sweeps", type=int, default=-1,
                   help="sweeps of each accuracy run (0 = no accuracy check, -1 = as --acc-burnin)")
    p.add_argument("--acc-replicas", type=int, default=16,
                   help="FP64 reference runs of the accuracy check")
    p.add_argument("--acc-dump", default="",
                   help="write every accuracy run's per-chain values to this .npz file")
    return p.parse_args()


def workload(name):
    from jmmonedmc_amd import configs
    if name == "c3":
        return configs.c3(), ("C3: 4096 chains, N log-uniform [32, 8192], T~U[0.5,2], P~U[0.1,2], "
                              "1/2 shifted LJ rc=3 + 1/2 hard rod 1 + square well 1.5, NPT ln L "
                              "+-0.01, step +-0.5, density 0.5 start, FP64 recompute every 100 sweeps")
    if name == "c2":
        return configs.c2(), ("C2: 256 chains, N=4096, NPT T=1 P=0.5, shifted LJ rc=3 "
                              "(512 blocks of 8 per chain)")
    if name == "c2x":
        return configs.c2(exclusion=0.75), ("C2 with the LJ exclusion distance at 0.75 sigma "
                                            "(0.5 sigma in C2): a window of 4 partners instead of 6")
    if name == "c2b":
        return configs.c2(bcap=448), ("C2 with block counts in whole waves up to 448 "
                                      "(448 blocks of 9-10 particles per chain; 512 in C2)")
    if name == "c4":
        return configs.c4(), ("C4: 1024 chains, N=1024, NVT T=1, rho=0.6, shifted LJ rc=3, "
                              "(eps, sigma) = (1, 1); after every sweep block the backward "
                              "(U, W, dU/dtheta, dW/dtheta, dU/dx) of every chain's configuration "
                              "into device arrays")
    return configs.c5(), "C5: 64 chains, N=65536, NPT T=1.5 P=1, Yukawa rc=20 + hard core 0.5"


def algorithmic_flops(eng, chains, before):
    """sum over chains of f_pair*(n_old+n_new) + (n_old+n_new) + 8 per trial,
    with n_old+n_new the pair evaluations the kernel counted; and the
    transcendental instructions those need (per pair by potential, plus the
    Metropolis exponential of every trial)."""
    from jmmonedmc_amd.configs import flop_per_pair, trans_per_pair
    cnt = eng.all_counters()
    terms = (cnt["pair_terms"] - before["pair_terms"]).astype(np.float64)
    trials = (cnt["trials"] - before["trials"]).astype(np.float64)
    fp = np.array([flop_per_pair(ch.potential) for ch in chains], dtype=np.float64)
    tot = float((fp * terms + terms + 8.0 * trials).sum())
    tp = np.array([trans_per_pair(ch.potential) for ch in chains], dtype=np.float64)
    trans = float((tp * terms + trials).sum())
    return tot, int(terms.sum()), int(trials.sum()), trans


def run_form(chains, form, steps, warmup, bs, graph, label, opts=None, backward=False):
    """Time `steps` sweep blocks; with `backward`, each block is followed by
    the backward of every chain's configuration (jmm_observables_device:
    U, W, dU/dtheta, dW/dtheta and dU/dx into device arrays, queued after
    the block on the engine's stream), inside the timed region.  The
    backward's own rate is timed after it, on the device and with the
    results read back to the host (jmm_observables)."""
    from jmmonedmc_amd.engine import Engine
    cap = max(steps, warmup, 1) * bs
    o = dict(use_graph=bool(graph))
    o.update(opts or {})
    eng = Engine(chains, form=form, sample_capacity=cap, block_sweeps=bs, **o)
    if backward:
        import torch
        obs_d = torch.empty(len(chains) * 6, dtype=torch.float64, device="cuda")
        grad_d = torch.empty(int(sum(c.N for c in chains)), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
    for _ in range(warmup):
        eng.run(bs)
    eng.sync()
    eng.read_samples(cap)
    eng.kernel_time(reset=True)
    before = eng.all_counters()
    print(f"[bench] {label}: timing {steps} blocks of {bs} sweeps", file=sys.stderr, flush=True)
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.run(bs)
        if backward:
            eng.observables_device(obs_d.data_ptr(), grad_d.data_ptr())
    eng.sync()
    t1 = time.perf_counter()
    launches, ms, sweeps = eng.kernel_time(reset=True)
    flops, terms, trials, trans = algorithmic_flops(eng, chains, before)
    if backward:
        nb = max(steps, 5)
        tb = time.perf_counter()
        for _ in range(nb):
            eng.observables_device(obs_d.data_ptr(), grad_d.data_ptr())
        eng.sync()
        dev_s = time.perf_counter() - tb
        tb = time.perf_counter()
        for _ in range(nb):
            eng.observables(want_grad=True)
        host_s = time.perf_counter() - tb
    t2 = time.perf_counter()
    smp = eng.read_samples(cap)
    readback = time.perf_counter() - t2
    eng.close()
    wall = t1 - t0
    out = dict(wall_s=wall, trials=trials, terms=terms, flops=flops, trans=trans, launches=launches,
               kernel_ms=ms, sweeps=sweeps, samples=smp, readback_s=readback,
               trials_per_s=trials / wall, ms_per_step=1e3 * wall / steps)
    if backward:
        out["backward"] = {"calls": nb, "configurations_per_call": len(chains),
                           "configurations_per_s": nb * len(chains) / dev_s,
                           "ms_per_call": 1e3 * dev_s / nb,
                           "includes": "U, W, dU/dtheta, dW/dtheta and dU/dx of every chain into device "
                                       "arrays (in the timed region after every block)",
                           "host_readback": {"configurations_per_s": nb * len(chains) / host_s,
                                             "ms_per_call": 1e3 * host_s / nb,
                                             "includes": "the same read back to host arrays over PCIe "
                                                         "(jmm_observables), not in the value"}}
    return out


# autocorrelation lags kept per run for the common-window tau_int (windows
# of up to 2000 sweeps, tau_int up to about 330 sweeps at c = 6)
ACF_LAGS = 2000


def accuracy_check(chains, opts, burnin, sweeps, bs, replicas, dump=""):
    """FP32 against FP64 per chain and observable at stated accuracy.

    A burn-in of `burnin` FP64 sweeps from the lattice start gives every
    chain a common starting configuration.  From it, `replicas` + 2 runs of
    `sweeps` sweeps with independent seeds: `replicas` FP64 reference runs,
    one more FP64 run (the null) and one FP32 run.  Per run and chain, the
    values of <U>, <L>, <rho>, cP, betaT and tau_int of U (stats.run_values;
    tau_int at one window per chain for every run, the median of the
    reference runs' own windows, stats.common_window_tau, and compared
    through its logarithm).
    The null and the FP32 run are each compared with the reference runs
    (stats.replica_compare): t = (x - mean) / (sd sqrt(1 + 1/K)) per chain,
    Student's t with K - 1 degrees of freedom if the run is one more draw
    of the reference runs' distribution, judged against the t band of the
    probability of a normal 3 sigma band (99.73 %).  The standard errors
    come from the runs' spread, so slow modes that an autocorrelation
    estimate from one run misses are counted.  The gate is valid when the
    null has at least 99 % of its chains within the band for every
    observable (with few chains: no more outside than the 99 % quantile of
    the binomial count the band's probability gives, stats.band_allowance),
    and passes when the FP32 run does too.  The FP32 error per
    observable: mean t (bias, with its standard error), the median
    difference in units of one run's spread, and relative differences, each
    beside the null's."""
    import dataclasses
    from jmmonedmc_amd import stats
    from jmmonedmc_amd import _native as nat
    from jmmonedmc_amd.engine import Engine
    t0 = time.perf_counter()
    workers = max(1, min(16, len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "16"))))
    print(f"[bench] accuracy: {burnin} FP64 burn-in sweeps", file=sys.stderr, flush=True)
    with Engine(chains, form="batched_fp64", sample_capacity=bs, block_sweeps=bs, **opts) as e:
        for _ in range(max(1, burnin // bs)):
            e.run(bs)
            e.read_samples(bs)
        x, L0, _ = e.states()
    N = np.array([c.N for c in chains])
    P = np.array([c.P for c in chains])
    T = np.array([c.T for c in chains])
    pot = np.array([c.potential for c in chains])
    names = {nat.POT_LJ: "lj", nat.POT_HRSW: "hrsw", nat.POT_YUKAWA: "yukawa", nat.POT_HARMONIC: "harmonic"}
    groups = {names[p]: pot == p for p in np.unique(pot)}
    runs = [(f"ref{j}", "batched_fp64") for j in range(replicas)] + [("null", "batched_fp64"),
                                                                     ("fp32", "batched_fp32")]
    vals = {}
    pending = None
    for j, (name, form) in enumerate(runs):
        print(f"[bench] accuracy: run {name} ({form}, {sweeps} sweeps)", file=sys.stderr, flush=True)
        chs = [dataclasses.replace(c, seed=c.seed + (j + 1) * 1_000_003) for c in chains]
        with Engine(chs, form=form, sample_capacity=sweeps, block_sweeps=bs, **opts) as e:
            e.set_states(x, L0)
            e.run(sweeps)
            if pending is not None:  # the previous run's values while this one runs
                vals[pending[0]] = stats.run_values(pending[1], pending[2], N, P, T, workers=workers,
                                                    acf_lags=ACF_LAGS)
            smp = e.read_samples(sweeps)
        pending = (name, np.ascontiguousarray(smp["U"]), np.ascontiguousarray(smp["L"]))
        del smp
    vals[pending[0]] = stats.run_values(pending[1], pending[2], N, P, T, workers=workers, acf_lags=ACF_LAGS)
    del pending
    ref = [vals[f"ref{j}"] for j in range(replicas)]
    own_tau = {r: d["tau_U"] for r, d in vals.items()}
    windows = stats.common_window_tau(ref, list(vals.values()))
    for d in vals.values():
        del d["_tau_cum"], d["_tau_window"]
    if dump:
        np.savez_compressed(dump, N=N, P=P, T=T, potential=pot, tau_window=windows,
                            **{f"{r}/{k}": v for r, d in vals.items() for k, v in d.items()},
                            **{f"{r}/tau_U_own_window": v for r, v in own_tau.items()})
    null = stats.replica_compare(ref, vals["null"], groups=groups)
    fp32 = stats.replica_compare(ref, vals["fp32"], groups=groups)
    # observables constant in the ensemble (the length at constant length)
    # are reported but not gated
    gated = [k for k in stats.OBSERVABLES if null[k]["chains_constant"] < len(chains)]

    def within(r):
        return r["chains_compared"] > 0 and r["outside"] <= r["allowed_outside"]
    valid = all(within(null[k]) for k in gated)
    passed = valid and all(within(fp32[k]) for k in gated)
    per_obs = {k: {"fp32_vs_fp64": fp32[k], "null_fp64_vs_fp64": null[k]} for k in stats.OBSERVABLES}
    return {"method": f"common start after an FP64 burn-in; {replicas} FP64 reference runs, one FP64 null run "
                      f"and one FP32 run with independent seeds; per chain t = (x - mean_ref) / "
                      f"(sd_ref sqrt(1 + 1/{replicas})), Student's t with {replicas - 1} degrees of freedom "
                      f"under the hypothesis, band = its two-sided 99.73 % quantile (the probability of 3 sigma); "
                      f"tau_U of every run at one window per chain, the median of the reference runs' "
                      f"Sokal windows (c = 6), compared through its logarithm",
            "tau_window_sweeps": {"median": float(np.median(windows)), "max": int(windows.max()),
                                  "chains_beyond_kept_lags": int((windows > ACF_LAGS).sum()),
                                  "kept_lags": ACF_LAGS},
            "burnin_sweeps": burnin, "sweeps_per_run": sweeps, "reference_runs": replicas,
            "chains": len(chains), "band": null["U"]["band"], "gate_observables": gated,
            "gate_valid_null_ge_0.99": bool(valid), "gate_passed": bool(passed), "observables": per_obs,
            "seconds": time.perf_counter() - t0}


def cpu_baseline(chains, sweeps, threads, allpairs=False):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    specs = [pyoracle.make_spec(**ch.as_dict()) for ch in chains]
    dt, _ = pyoracle.run_batch(specs, sweeps, threads, allpairs=allpairs)
    trials = sweeps * sum(ch.N for ch in chains)
    return trials / dt, dt, trials


def cpu_extra(chains, n_trials_per_sweep):
    """The oracle on one core over the same workload, and the reference's
    algorithm (every other particle a partner of every trial, as qad2,
    src/jmmMCState.cpp:1213-1348) on one core for C1 and one C2 chain."""
    from jmmonedmc_amd import configs
    sw1 = max(1, int(10.0 * 2.7e7 / n_trials_per_sweep))  # about 10 s at 2.7e7 trials/s
    v1, dt1, tr1 = cpu_baseline(chains, sw1, 1)
    out = {"single_core": {"value": v1, "unit": "trial moves/s", "cores": 1,
                           "sample": f"FP64 oracle, all {len(chains)} chains x {sw1} sweeps "
                                     f"({tr1:.3g} trials, {dt1:.1f} s)"}}
    ref = {"unit": "trial moves/s", "cores": 1,
           "note": "FP64 oracle with every other particle a partner of every trial "
                   "(the reference's O(N) loop), not the cutoff window"}
    for name, chs, sw in (("c1", configs.c1(), 10000), ("c2_one_chain", configs.c2(n_chains=1), 100)):
        v, dt, tr = cpu_baseline(chs, sw, 1, allpairs=True)
        ref[name] = {"value": v, "sample": f"N={chs[0].N}, {sw} sweeps ({tr:.3g} trials, {dt:.1f} s)"}
    out["reference_algorithm"] = ref
    return out


def calibrate_copy(mib):
    """A device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE
    readings scale the sweep kernel's counters to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}))
    return 0


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this "
                                   "build (chains are independent and would shard across GPUs, one engine per GPU)",
                          "n_gpus": a.gpus}))
        return 2
    if a.calibrate_copy:
        return calibrate_copy(a.calibrate_copy)
    if a.workload == "c4":
        # the backward's device arrays come from torch, whose HIP runtime
        # must take the device before the engine library's does
        import torch
        torch.cuda.init()
    forms = set(a.forms.split(","))
    chains, desc = workload(a.workload)
    n_trials_per_sweep = sum(ch.N for ch in chains)
    T0 = time.perf_counter()
    timeline = []

    tuned = None
    opts = {}
    if a.tune:
        from jmmonedmc_amd.tune import tune
        print("[bench] autotuning launch parameters", file=sys.stderr, flush=True)
        tuned = tune(chains, form="batched_fp32", block_sweeps=a.block_sweeps)
        opts = dict(tuned.best)
    bwd = a.workload == "c4"

    def leg(name, fn):
        t = time.perf_counter() - T0
        r = fn()
        timeline.append({"leg": name, "start_s": round(t, 3), "end_s": round(time.perf_counter() - T0, 3)})
        return r

    if tuned:
        timeline.append({"leg": "autotune (GPU)", "start_s": 0.0, "end_s": round(time.perf_counter() - T0, 3)})
    fp32 = leg("batched_fp32 (GPU, the value)",
               lambda: run_form(chains, "batched_fp32", a.steps, a.warmup, a.block_sweeps, a.graph,
                                "batched_fp32", opts, backward=bwd))
    samples_bytes = int(fp32["samples"].size * 24)
    fp32.pop("samples")
    res = {}
    if "fp64" in forms:
        res["batched_fp64"] = leg("batched_fp64 (GPU)",
                                  lambda: run_form(chains, "batched_fp64", a.steps, a.warmup, a.block_sweeps,
                                                   a.graph, "batched_fp64", opts, backward=bwd))
        res["batched_fp64"].pop("samples")
    if "plain" in forms and a.plain_steps > 0:
        res["plain_fp64"] = leg("plain_fp64 (GPU)",
                                lambda: run_form(chains, "plain_fp64", a.plain_steps, 0, a.block_sweeps, 0,
                                                 "plain_fp64"))
        res["plain_fp64"].pop("samples")

    accuracy = None
    acc_burn = a.acc_burnin if a.acc_burnin >= 0 else (2000 if a.workload == "c5" else 10000)
    acc_sw = a.acc_sweeps if a.acc_sweeps >= 0 else acc_burn
    if acc_sw > 0:
        accuracy = leg("accuracy runs (GPU, untimed)",
                       lambda: accuracy_check(chains, opts, acc_burn, acc_sw, a.block_sweeps,
                                                      a.acc_replicas, a.acc_dump))

    # roofline of the dominant kernel (the batched FP32 sweep kernel)
    achieved = fp32["flops"] / (fp32["kernel_ms"] * 1e-3) / 1e12
    roof = {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
            "frac": achieved / PEAK_TFLOPS["fp32"], "traffic": None,
            "flops_per_launch": fp32["flops"] / max(1, fp32["launches"]),
            "avg_launch_ms": fp32["kernel_ms"] / max(1, fp32["launches"]),
            "pair_evals_per_trial": fp32["terms"] / max(1, fp32["trials"])}
    trans_achieved = fp32["trans"] / (fp32["kernel_ms"] * 1e-3) / 1e12
    roof["secondary"] = {"bound": "transcendental", "achieved": trans_achieved, "peak": PEAK_TRANS_TOPS,
                         "unit": "Top/s", "frac": trans_achieved / PEAK_TRANS_TOPS,
                         "per_trial": fp32["trans"] / max(1, fp32["trials"])}
    traffic_file = os.path.join(ROOT, "profiles", f"{a.workload}_sweep_traffic.txt")
    if os.path.exists(traffic_file):
        for line in open(traffic_file):
            if line.startswith("hbm_bytes_per_launch:"):
                roof["traffic"] = float(line.split(":", 1)[1])

    cpu = None
    if not a.no_cpu:
        # every CPU the process may run on; a host that allots a CPU share per
        # GPU job states it in OMP_NUM_THREADS, which then caps the threads
        host = os.cpu_count() or 1
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else host
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = min(avail, omp) if omp > 0 else avail
        sw = a.cpu_sweeps or max(1, int(15.0 * 2.4e7 * threads / n_trials_per_sweep))  # ~15 s
        v, dt, tr = leg("cpu baseline, all threads (CPU)", lambda: cpu_baseline(chains, sw, threads))
        cpu = {"value": v, "unit": "trial moves/s", "cores": threads, "kind": "port",
               "host_cpus": host, "cpus_available": avail, "omp_num_threads": omp or None,
               "sample": f"FP64 oracle, all {len(chains)} chains x {sw} sweeps ({tr:.3g} trials, "
                         f"{dt:.1f} s), one chain per thread on {threads} threads"}
        cpu.update(leg("cpu baseline, one core and the all-partner algorithm (CPU)",
                       lambda: cpu_extra(chains, n_trials_per_sweep)))

    plain = res.get("plain_fp64")
    out = {
        "metric": METRIC,
        "value": fp32["trials_per_s"],
        "unit": "trial moves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": fp32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": desc, "n_chains": len(chains), "particles": n_trials_per_sweep,
                   "sweeps_timed": a.steps * a.block_sweeps, "block_sweeps": a.block_sweeps,
                   "form": "batched_fp32", "parallelism": "1 GPU, chains packed into workgroups",
                   "launch": opts or {"use_graph": bool(a.graph)}},
        "autotune_ms_per_sweep": tuned and tuned.ms_per_sweep,
        "chain_sweeps_per_s": len(chains) * a.steps * a.block_sweeps / fp32["wall_s"],
        "sample_readback": {"seconds": fp32["readback_s"],
                            "bytes": samples_bytes,
                            "value_including_readback": fp32["trials"] / (fp32["wall_s"] + fp32["readback_s"])},
        "plain_fp64": plain and {"value": plain["trials_per_s"], "sweeps_timed": plain["sweeps"],
                                 "ms_per_block": plain["ms_per_step"]},
        "speedup_vs_plain_fp64": plain and fp32["trials_per_s"] / plain["trials_per_s"],
        "speedup_vs_batched_fp64": ("batched_fp64" in res) and
                                   fp32["trials_per_s"] / res["batched_fp64"]["trials_per_s"] or None,
        "value_status": (None if accuracy is None else
                         "accuracy gate passed" if accuracy["gate_passed"] else
                         "accuracy gate failed" if accuracy["gate_valid_null_ge_0.99"] else
                         "accuracy gate not valid (the FP64 null is below 99 % within the 3 sigma band)"),
        "roofline": roof,
        "cpu_baseline": cpu,
        "accuracy_fp32_vs_fp64": accuracy,
        "timeline": timeline,
    }
    if "backward" in fp32:
        out["backward"] = fp32["backward"]
    if "batched_fp64" in res:
        b = res["batched_fp64"]
        out["batched_fp64"] = {"value": b["trials_per_s"],
                               "achieved_tflops": b["flops"] / (b["kernel_ms"] * 1e-3) / 1e12,
                               "peak": PEAK_TFLOPS["fp64"]}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
