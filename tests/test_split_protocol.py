"""The synchronisation protocol of split chains, modelled on the CPU.

A chain split into S segments (k_sweep_seg in jmmonedmc_amd/csrc/kernels.hip,
one workgroup per segment) keeps its neighbours' boundary particles in LDS
halos.  Segments run freely between synchronisations; a neighbour sync comes
after phases K-1, 2K-1, ... (K = qb - wg) and after a sweep's last phase,
then the halo is re-read from what the neighbours published.  A segment
publishes its boundary particles as it moves them.  The claim is that every
position a trial, a box-length pass or a drift check reads is then the one
the sequential schedule (for t: for b:) reads.  On the GPU the segments run
nearly in lock-step, so the GPU tests cannot show that the claim survives
the largest skews the protocol allows; this model runs the protocol's steps
under random, bursty interleavings.

Each particle carries a version: how many times it was moved or rescaled.
The model mirrors run_segment's sequence: initial publication and neighbour
sync; per sweep the phases (one step per block), the neighbour syncs, the
box-length pass (the periodic chain's last segment reads segment 0's
published head directly), the all-segment sync, on acceptance the rescale,
publication and neighbour sync, and on drift-check sweeps the check's reads
and an all-segment sync; finally the counters' all-segment sync.  A halo is
refreshed one element per step, so a neighbour can move between two of its
loads.  A read outside the halo is itself a failure.  Planted flaws (one
phase more between syncs, no sync at a sweep's end, no refresh after an
accepted box move) must each be caught."""
import random

import pytest


def blocks(N, B):
    qb, rem = N // B, N % B
    sizes = [qb + (1 if b < rem else 0) for b in range(B)]
    starts = [b * qb + min(b, rem) for b in range(B)]
    return qb, sizes, starts


def right(i, N, wg, per):
    return [(i + k) % N for k in range(1, wg + 1) if per or i + k < N]


def window(i, N, wg, per):
    """Indices a displacement trial of particle i may read."""
    left = [(i - k) % N for k in range(1, wg + 1) if per or i - k >= 0]
    return left + right(i, N, wg, per)


def plan(n_sweeps, seed):
    """Box-length acceptances and drift-check sweeps, as every segment of a
    chain sees them (they follow from the chain's totals)."""
    rng = random.Random(10_000 + seed)
    accept = [rng.random() < 0.5 for _ in range(n_sweeps)]
    guard = [(sw + 1) % 2 == 0 for sw in range(n_sweeps)]
    return accept, guard


def sequential_reads(N, B, wg, per, n_sweeps, seed):
    _, sizes, starts = blocks(N, B)
    accept, guard = plan(n_sweeps, seed)
    ver = [0] * N
    reads = {}
    for sw in range(n_sweeps):
        for t in range(max(sizes)):
            for b in range(B):
                if t < sizes[b]:
                    i = starts[b] + t
                    reads[(sw, "trial", i)] = {j: ver[j] for j in window(i, N, wg, per)}
                    ver[i] += 1
        for i in range(N):
            reads[(sw, "box", i)] = {j: ver[j] for j in right(i, N, wg, per)}
        if accept[sw]:
            ver = [v + 1 for v in ver]
        if guard[sw]:
            for i in range(N):
                reads[(sw, "guard", i)] = {j: ver[j] for j in right(i, N, wg, per)}
    return reads


class Chain:
    def __init__(self, N, B, S, wg, per):
        self.N, self.B, self.S, self.wg, self.per = N, B, S, wg, per
        self.qb, self.sizes, self.starts = blocks(N, B)
        self.smax = max(self.sizes)
        self.Bs = B // S
        self.seg_of = [0] * N
        for b in range(B):
            for t in range(self.sizes[b]):
                self.seg_of[self.starts[b] + t] = b // self.Bs
        self.ver = [0] * N
        self.ctr = [0] * S
        self.nbrs, self.halo_idx = [], []
        for s in range(S):
            has_l, has_r = per or s > 0, per or s + 1 < S
            nb, idx = set(), []
            if has_l:
                ln = (s - 1) % S
                nb.add(ln)
                bl = ln * self.Bs + self.Bs - 1
                e = self.starts[bl] + self.sizes[bl]
                idx += list(range(e - wg, e))
            if has_r:
                rn = (s + 1) % S
                nb.add(rn)
                f = self.starts[rn * self.Bs]
                idx += list(range(f, f + wg))
            self.nbrs.append(sorted(nb - {s}))
            self.halo_idx.append(idx)
        self.halo = [dict() for _ in range(S)]
        self.reads = {}

    def seg_program(self, s, K, n_sweeps, accept, guard, flaw):
        """Generator: one step of segment s per next(); yields "blocked" while
        it waits on a partner's counter."""
        sp = 0
        others = [x for x in range(self.S) if x != s]
        own_blocks = range(s * self.Bs, (s + 1) * self.Bs)
        own = [i for i in range(self.N) if self.seg_of[i] == s]

        def sync(partners):
            nonlocal sp
            sp += 1
            self.ctr[s] = sp
            while not all(self.ctr[p] >= sp for p in partners):
                yield "blocked"

        def refresh():
            for j in self.halo_idx[s]:
                self.halo[s][j] = self.ver[j]
                yield "step"

        def read(j, live_wrap=False):
            if self.seg_of[j] == s or live_wrap:
                return self.ver[j]
            assert j in self.halo[s], f"segment {s} reads {j} outside its halo"
            return self.halo[s][j]

        yield from sync(self.nbrs[s])
        yield from refresh()
        for sw in range(n_sweeps):
            for t in range(self.smax):
                for b in own_blocks:
                    if t < self.sizes[b]:
                        i = self.starts[b] + t
                        self.reads[(sw, "trial", i)] = {j: read(j) for j in window(i, self.N, self.wg, self.per)}
                        self.ver[i] += 1
                        yield "step"
                last = t + 1 == self.smax
                if (t + 1) % K == 0 or (last and flaw != "no_end_sync"):
                    yield from sync(self.nbrs[s])
                    yield from refresh()
            wrap = self.per and s + 1 == self.S
            for i in own:
                self.reads[(sw, "box", i)] = {
                    j: read(j, live_wrap=wrap and self.seg_of[j] == 0) for j in right(i, self.N, self.wg, self.per)}
            yield "step"
            yield from sync(others)
            if accept[sw]:
                for i in own:
                    self.ver[i] += 1
                yield "step"
                yield from sync(self.nbrs[s])
                if flaw != "no_refresh_after_accept":
                    yield from refresh()
            if guard[sw]:
                for i in own:
                    self.reads[(sw, "guard", i)] = {j: read(j) for j in right(i, self.N, self.wg, self.per)}
                yield "step"
                yield from sync(others)
        yield from sync(others)


def protocol_reads(N, B, S, wg, per, n_sweeps, seed, flaw=None):
    ch = Chain(N, B, S, wg, per)
    K = ch.qb - wg + (1 if flaw == "K+1" else 0)
    accept, guard = plan(n_sweeps, seed)
    progs = [ch.seg_program(s, K, n_sweeps, accept, guard, flaw) for s in range(S)]
    live = set(range(S))
    rng = random.Random(seed)
    idle = 0
    while live:
        s = rng.choice(sorted(live))
        # a burst: run segment s for up to 3K steps or until it blocks
        for _ in range(rng.randint(1, 3 * K)):
            try:
                st = next(progs[s])
            except StopIteration:
                live.discard(s)
                break
            if st == "blocked":
                idle += 1
                break
            idle = 0
        assert idle < 1000 * S, "deadlock"
    return ch.reads


CASES = [
    # N, B, S, wg, periodic
    (96, 8, 2, 3, True),
    (100, 8, 4, 4, True),
    (130, 16, 4, 5, False),
    (203, 16, 8, 7, True),
    (64, 4, 4, 1, False),
    (300, 16, 16, 9, True),
    (170, 8, 2, 12, False),
]


def mismatches(N, B, S, wg, per, seed, flaw=None, n_sweeps=4):
    ref = sequential_reads(N, B, wg, per, n_sweeps, seed)
    got = protocol_reads(N, B, S, wg, per, n_sweeps, seed, flaw)
    assert got.keys() == ref.keys()
    return [key for key in ref if got[key] != ref[key]]


@pytest.mark.parametrize("N,B,S,wg,per", CASES)
def test_split_protocol_reads_match_sequential(N, B, S, wg, per):
    assert N // B > wg
    for seed in range(25):
        bad = mismatches(N, B, S, wg, per, seed)
        assert not bad, (seed, bad[:3])


@pytest.mark.parametrize("flaw", ["K+1", "no_end_sync", "no_refresh_after_accept"])
@pytest.mark.parametrize("N,B,S,wg,per", [CASES[1], CASES[3], CASES[6]])
def test_planted_flaws_are_caught(N, B, S, wg, per, flaw):
    qb = N // B
    smax = qb + (1 if N % B else 0)
    if flaw == "no_end_sync" and smax % (qb - wg) == 0:
        pytest.skip("the sweep's last phase is a K boundary: the flaw changes nothing")
    for seed in range(300):
        try:
            if mismatches(N, B, S, wg, per, seed, flaw):
                return
        except AssertionError as e:  # a read outside the halo counts too
            if "outside its halo" in str(e):
                return
            raise
    pytest.fail(f"flaw {flaw} not caught in 300 schedules")
