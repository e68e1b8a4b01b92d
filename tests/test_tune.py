"""Autotuner candidates (host logic, no GPU)."""
from jmmonedmc_amd import configs
from jmmonedmc_amd.tune import DEFAULT_CANDIDATES, candidates_for, signature


def test_candidates_cover_widths_and_split_parameters():
    """Every workload gets both workgroup widths; only workloads with chains
    that may be split get split thresholds and two lanes per block; every
    candidate names only engine options."""
    opts = {"lds_budget", "wg_threads", "use_graph", "seg_lds", "seg_lanes"}
    small = candidates_for(configs.c1())
    assert {c["wg_threads"] for c in small} == {256, 512, 1024}
    assert all("seg_lds" not in c and "seg_lanes" not in c for c in small)
    assert all(c in small for c in DEFAULT_CANDIDATES)
    big = candidates_for(configs.c5(2))
    assert any(c.get("seg_lanes") == 2 for c in big)
    assert {c.get("seg_lds") for c in big} >= {65536, 32768}
    for c in small + big:
        assert set(c) <= opts
    assert len({str(sorted(c.items())) for c in big}) == len(big)


def test_signature_ignores_seeds():
    a = configs.c3(64)
    b = [c.__class__(**{**c.as_dict(), "seed": c.seed + 5}) for c in a]
    assert signature(a, "batched_fp32") == signature(b, "batched_fp32")
    assert signature(a, "batched_fp32") != signature(a, "batched_fp64")
