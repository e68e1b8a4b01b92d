"""jmm_read_samples_noalloc: the sample read-back staged in the library's
fixed per-GPU device array instead of a buffer it allocates.  Two engines
on the same chains, one read with jmm_read_samples and one with the
no-allocation entry point, must return the same records byte for byte, on
two reads in a row (the second smaller); a read larger than the staging
array is refused and drains nothing."""
import ctypes as C

import numpy as np
import pytest

from jmmonedmc_amd import _native as nat
from jmmonedmc_amd.engine import SAMPLE_DTYPE, Engine

from helpers import CASES, lj


def read_noalloc(eng, m):
    out = np.empty((eng.n, m), dtype=SAMPLE_DTYPE)
    n = C.c_uint32()
    rc = nat.lib().jmm_read_samples_noalloc(eng.h, out.ctypes.data_as(C.POINTER(nat.Sample)), m, C.byref(n))
    return rc, out[:, : n.value]


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["batched_fp32", "batched_fp64"])
def test_noalloc_read_equals_allocating_read(gpu, form):
    chains = CASES["mixed_batch"]()
    print(f"[gpu] no-allocation read {form}", flush=True)
    a = Engine(chains, form=form, sample_capacity=64, block_sweeps=16)
    b = Engine(chains, form=form, sample_capacity=64, block_sweeps=16)
    try:
        for sweeps in (50, 20):
            a.run(sweeps)
            b.run(sweeps)
            want = a.read_samples(64)
            rc, got = read_noalloc(b, 64)
            assert rc == 0, nat.lib().jmm_last_error()
            assert got.shape == want.shape == (len(chains), sweeps)
            assert got.tobytes() == want.tobytes()
        assert np.array_equal(a.states()[1], b.states()[1])
    finally:
        a.close()
        b.close()


@pytest.mark.gpu
def test_noalloc_read_refuses_beyond_its_staging(gpu):
    """8192 chains x 100 sweeps of records (19.7 MB) exceed the 16 MiB
    staging array: refused with a message, the samples stay pending and the
    allocating read returns them."""
    from jmmonedmc_amd.engine import Chain
    chains = [Chain(N=2, L0=8.0, T=1.0, seed=k, chain_id=k, **lj()) for k in range(8192)]
    with Engine(chains, form="batched_fp32", sample_capacity=100, block_sweeps=100) as e:
        e.run(100)
        rc, _ = read_noalloc(e, 100)
        assert rc == -1  # JMM_EINVAL
        assert b"staging" in nat.lib().jmm_last_error()
        smp = e.read_samples(100)
        assert smp.shape == (8192, 100)
        rc, got = read_noalloc(e, 100)
        assert rc == 0 and got.shape == (8192, 0)
