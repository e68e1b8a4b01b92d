"""Per-sweep virial and hypervirial (engine option sample_virial) against
the oracle: W = sum -d phi'(d) and HV = sum -d W'(d) of the configuration
after every sweep, the per-trial Vir/HV sums of the reference's
updateThermo (src/jmmMCState.cpp:1941-1961) taken once per sweep.  Both are
FP64 sums over the same pairs; they differ from the oracle's only in the
order of addition."""
import numpy as np
import pytest

from jmmonedmc_amd._native import JmmError
from jmmonedmc_amd.engine import Engine

from helpers import CASES, oracle_spec

SWEEPS = 30
VIRIAL_CASES = ["c1_periodic_nvt", "c1_walls_truncated", "npt_walls_linear", "npt_periodic_log",
                "harmonic_nbn1_walls", "yukawa_npt", "hrsw_npt", "lj_inf_walls_npt", "mixed_batch"]


def close(a, b, n):
    return np.all(np.abs(a - b) <= 1e-11 * np.maximum(1.0, np.abs(b)) + 1e-12 * n)


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain_fp64", "batched_fp64"])
@pytest.mark.parametrize("case", VIRIAL_CASES)
def test_virials_per_sweep_match_oracle(gpu, oracle, form, case):
    chains = CASES[case]()
    print(f"[gpu] virials {form} {case}", flush=True)
    with Engine(chains, form=form, sample_capacity=SWEEPS, block_sweeps=7, sample_virial=True) as e:
        e.run(SWEEPS)
        smp, vir = e.read_samples_virial(SWEEPS)
    for c, ch in enumerate(chains):
        orc = oracle.Chain(oracle_spec(oracle, ch))
        osmp, ovir = orc.sweeps_virial(SWEEPS)
        np.testing.assert_array_equal(smp[c]["L"], osmp["L"])
        assert close(vir[c, :, 0], ovir[:, 0], ch.N), (c, vir[c, :3, 0], ovir[:3, 0])
        assert close(vir[c, :, 1], ovir[:, 1], ch.N), (c, vir[c, :3, 1], ovir[:3, 1])


@pytest.mark.gpu
@pytest.mark.parametrize("case", VIRIAL_CASES)
def test_fp32_virials_are_those_of_its_configuration(gpu, oracle, case):
    """FP32 form: the virials are FP64 sums over the chain's own
    configuration; the last sweep's equal the oracle's virials of the
    engine's final positions."""
    chains = CASES[case]()
    with Engine(chains, form="batched_fp32", sample_capacity=SWEEPS, block_sweeps=7, sample_virial=True) as e:
        e.run(SWEEPS)
        smp, vir = e.read_samples_virial(SWEEPS)
        states = [e.state(c) for c in range(len(chains))]
    assert np.all(np.isfinite(vir))
    for c, ch in enumerate(chains):
        x, L, _, _ = states[c]
        want = oracle.virials_x(oracle_spec(oracle, ch), x, L)
        assert close(vir[c, -1], want, ch.N), (c, vir[c, -1], want)


@pytest.mark.gpu
def test_split_chain_virials_match_oracle(gpu, oracle):
    """A chain split over 16 workgroups (segment syncs carry the partial
    sums), FP64 form: per-sweep virials equal to the oracle's."""
    from test_gpu_split import seg_lds_for, split_case
    ch = split_case("lj_periodic_npt_log")
    seg_lds, _ = seg_lds_for(oracle, ch, 16)
    with Engine([ch], form="batched_fp64", sample_capacity=SWEEPS, block_sweeps=10, seg_lds=seg_lds,
                sample_virial=True) as e:
        assert e.layout(0)["mode"] == "split"
        e.run(SWEEPS)
        smp, vir = e.read_samples_virial(SWEEPS)
    orc = oracle.Chain(oracle_spec(oracle, ch))
    osmp, ovir = orc.sweeps_virial(SWEEPS)
    np.testing.assert_array_equal(smp[0]["L"], osmp["L"])
    assert close(vir[0, :, 0], ovir[:, 0], ch.N) and close(vir[0, :, 1], ovir[:, 1], ch.N)


@pytest.mark.gpu
def test_virial_read_needs_the_option(gpu):
    chains = CASES["c1_periodic_nvt"]()
    with Engine(chains, form="batched_fp64", sample_capacity=4) as e:
        e.run(2)
        with pytest.raises(JmmError, match="sample_virial"):
            e.read_samples_virial(4)
