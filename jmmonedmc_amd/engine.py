"""Python view of the engine (include/jmm_engine.h), for tests, benchmarks and
the autograd binding.  Every call goes through the C ABI; nothing here
computes the chain on the CPU.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as nat
from ._native import ChainSpec, Counters, EngineOpts, HistSpec, Sample, check, lib

# numpy view of jmm_sample
SAMPLE_DTYPE = np.dtype([("U", "f8"), ("L", "f8"), ("acc_d", "u4"), ("acc_v", "u4")])
assert SAMPLE_DTYPE.itemsize == C.sizeof(Sample)


@dataclass
class Chain:
    """One chain (the hot-path fields of the reference's MCInput,
    src/jmmMCState.h:66-101, plus the engine's additive fields)."""
    N: int
    T: float = 1.0
    P: float = 0.0
    L0: float | None = None
    rho: float | None = None
    potential: int = nat.POT_LJ
    boundary: int = nat.PERIODIC
    ensemble: int = nat.NLT
    vol_move: int = nat.VOL_LOG
    nbn: int = -1
    shifted: int = 1
    max_step: float = 0.5
    max_dv: float = 0.01
    cutoff: float = 3.0
    pot_p: tuple = (1.0, 1.0, 0.0, 0.0)
    seed: int = 1
    chain_id: int = 0
    eci: int = 0
    bcap: int = 256

    def length(self) -> float:
        return self.L0 if self.L0 is not None else self.N / self.rho

    def to_c(self) -> ChainSpec:
        s = ChainSpec()
        s.potential, s.boundary, s.ensemble, s.vol_move = (self.potential, self.boundary,
                                                           self.ensemble, self.vol_move)
        s.nbn, s.shifted, s.N = self.nbn, self.shifted, self.N
        s.T, s.P, s.L0 = self.T, self.P, self.length()
        s.max_step, s.max_dv, s.cutoff = self.max_step, self.max_dv, self.cutoff
        for k in range(4):
            s.pot_p[k] = self.pot_p[k] if k < len(self.pot_p) else 0.0
        s.seed, s.chain_id, s.eci, s.bcap = self.seed, self.chain_id, self.eci, self.bcap
        return s

    def as_dict(self) -> dict:
        d = dict(self.__dict__)
        d["L0"] = self.length()
        d.pop("rho")
        return d


def schedule(chain: Chain):
    B, M, w = C.c_uint64(), C.c_uint64(), C.c_uint64()
    s = chain.to_c()
    check(lib().jmm_schedule(C.byref(s), C.byref(B), C.byref(M), C.byref(w)), "jmm_schedule")
    return B.value, M.value, w.value


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Engine:
    """Many chains on one GPU."""

    def __init__(self, chains, form="batched_fp32", device=0, sample_capacity=1024,
                 block_sweeps=100, wg_threads=0, use_graph=False, log_accepts=False,
                 lds_budget=0, seg_lds=0, seg_lanes=0, sample_virial=False):
        self.chains = list(chains)
        self.n = len(self.chains)
        self.form = nat.FORMS[form] if isinstance(form, str) else int(form)
        self.Ns = np.array([c.N for c in self.chains], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.Ns)[:-1]]).astype(np.int64)
        self.cap = sample_capacity
        specs = (ChainSpec * self.n)(*[c.to_c() for c in self.chains])
        o = EngineOpts()
        o.form, o.device, o.sample_capacity, o.block_sweeps = self.form, device, sample_capacity, block_sweeps
        o.wg_threads, o.use_graph, o.log_accepts, o.lds_budget = (wg_threads, int(use_graph),
                                                                  int(log_accepts), lds_budget)
        o.seg_lds = seg_lds
        o.seg_lanes = seg_lanes
        o.sample_virial = int(sample_virial)
        h = C.c_void_p()
        check(lib().jmm_create(specs, self.n, C.byref(o), C.byref(h)), "jmm_create")
        self.h = h
        self.log_accepts = log_accepts

    def close(self):
        if getattr(self, "h", None):
            lib().jmm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, n_sweeps: int):
        check(lib().jmm_run_sweeps(self.h, n_sweeps), "jmm_run_sweeps")

    def sync(self):
        check(lib().jmm_synchronize(self.h), "jmm_synchronize")

    def read_samples(self, max_sweeps: int | None = None):
        """Structured array [chain, sweep] with fields U, L, acc_d, acc_v
        (the library writes the rows straight into it)."""
        m = max_sweeps or self.cap
        out = np.empty((self.n, m), dtype=SAMPLE_DTYPE)
        n = C.c_uint32()
        check(lib().jmm_read_samples(self.h, out.ctypes.data_as(C.POINTER(Sample)), m, C.byref(n)),
              "jmm_read_samples")
        return out if n.value == m else out[:, : n.value].copy()

    def read_samples_virial(self, max_sweeps: int | None = None):
        """(samples, virials): samples as read_samples, virials a float64
        array [chain, sweep, 2] of (W, HV) after each sweep (engines created
        with sample_virial)."""
        m = max_sweeps or self.cap
        out = np.empty((self.n, m), dtype=SAMPLE_DTYPE)
        vir = np.empty((self.n, m, 2))
        n = C.c_uint32()
        check(lib().jmm_read_samples_virial(self.h, out.ctypes.data_as(C.POINTER(Sample)), _dp(vir), m,
                                            C.byref(n)), "jmm_read_samples_virial")
        if n.value == m:
            return out, vir
        return out[:, : n.value].copy(), vir[:, : n.value].copy()

    def read_accept_log(self, max_sweeps: int | None = None):
        """List (per chain) of uint8 arrays [sweep, particle]; read before read_samples."""
        m = max_sweeps or self.cap
        buf = np.zeros(int(self.Ns.sum()) * m, dtype=np.uint8)
        n = C.c_uint32()
        check(lib().jmm_read_accept_log(self.h, buf.ctypes.data_as(C.POINTER(C.c_uint8)), m,
                                        C.byref(n)), "jmm_read_accept_log")
        out = []
        for c in range(self.n):
            o = int(self.offsets[c]) * m
            out.append(buf[o: o + n.value * int(self.Ns[c])].reshape(n.value, int(self.Ns[c])))
        return out

    def state(self, chain: int):
        x = np.zeros(int(self.Ns[chain]))
        L, U, sw = C.c_double(), C.c_double(), C.c_uint64()
        check(lib().jmm_get_state(self.h, chain, _dp(x), C.byref(L), C.byref(U), C.byref(sw)),
              "jmm_get_state")
        return x, L.value, U.value, sw.value

    def states(self):
        """All chains at once: (list of position arrays, L[n], U[n])."""
        x = np.zeros(int(self.Ns.sum()))
        L = np.zeros(self.n)
        U = np.zeros(self.n)
        check(lib().jmm_get_states(self.h, _dp(x), _dp(L), _dp(U)), "jmm_get_states")
        return [x[o: o + n] for o, n in zip(self.offsets, self.Ns)], L, U

    def set_state(self, chain: int, x, L: float, sweep: int = 0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        check(lib().jmm_set_state(self.h, chain, _dp(x), L, sweep), "jmm_set_state")

    def set_states(self, x, L, U=None, sweep=None):
        """Every chain at once: x (list of position arrays or one flat
        array in chain order), L[n]; U[n] installed as given, or recomputed
        in FP64 when None; sweep[n] (None keeps the counters)."""
        x = np.ascontiguousarray(np.concatenate(x) if isinstance(x, (list, tuple)) else x, dtype=np.float64)
        assert x.size == int(self.Ns.sum())
        L = np.ascontiguousarray(L, dtype=np.float64)
        Ua = None if U is None else np.ascontiguousarray(U, dtype=np.float64)
        sw = None if sweep is None else np.ascontiguousarray(sweep, dtype=np.uint64)
        check(lib().jmm_set_states(self.h, _dp(x), _dp(L), _dp(Ua) if Ua is not None else None,
                                   sw.ctypes.data_as(C.POINTER(C.c_uint64)) if sw is not None else None),
              "jmm_set_states")

    def set_steps(self, max_step=None, max_dv=None):
        """Per-chain displacement and box-move half-widths (None = unchanged)."""
        ms = None if max_step is None else np.ascontiguousarray(max_step, dtype=np.float64)
        mv = None if max_dv is None else np.ascontiguousarray(max_dv, dtype=np.float64)
        check(lib().jmm_set_steps(self.h, _dp(ms) if ms is not None else None,
                                  _dp(mv) if mv is not None else None), "jmm_set_steps")

    def scaled_energy(self, L_try):
        """FP64 energy of every chain rescaled to L_try[c] (state unchanged)."""
        Lt = np.ascontiguousarray(L_try, dtype=np.float64)
        U = np.zeros(self.n)
        check(lib().jmm_scaled_energy(self.h, _dp(Lt), _dp(U)), "jmm_scaled_energy")
        return U

    def set_lengths(self, L_new):
        Ln = np.ascontiguousarray(L_new, dtype=np.float64)
        check(lib().jmm_set_lengths(self.h, _dp(Ln)), "jmm_set_lengths")

    def perturb_energy(self, chain: int, delta: float):
        check(lib().jmm_perturb_energy(self.h, chain, delta), "jmm_perturb_energy")

    def counters(self, chain: int) -> dict:
        c = Counters()
        check(lib().jmm_get_counters(self.h, chain, C.byref(c)), "jmm_get_counters")
        return {k: getattr(c, k) for k, _ in Counters._fields_}

    def layout(self, chain: int) -> dict:
        """Placement of a chain: blocks B, window w, memory mode ("lds",
        "hbm" or "split"), workgroups spanned, LDS bytes, lanes per block."""
        l = nat.Layout()
        check(lib().jmm_chain_layout(self.h, chain, C.byref(l)), "jmm_chain_layout")
        return dict(B=l.B, w=l.w, mode=nat.MODES[l.mode], segments=l.segments, lds_bytes=l.lds_bytes,
                    lanes=l.lanes)

    def all_counters(self) -> dict:
        """Every chain's counters in one call: dict of arrays indexed by chain."""
        buf = (Counters * self.n)()
        check(lib().jmm_get_counters(self.h, -1, buf), "jmm_get_counters")
        raw = np.frombuffer(buf, dtype=[(k, "u8" if t is C.c_uint64 else "f8")
                                        for k, t in Counters._fields_])
        return {k: raw[k].copy() for k, _ in Counters._fields_}

    def full_energy(self):
        U = np.zeros(self.n)
        check(lib().jmm_full_energy(self.h, _dp(U)), "jmm_full_energy")
        return U

    def observables(self, want_grad: bool = False):
        obs = np.empty(self.n * 6)
        g = np.empty(int(self.Ns.sum())) if want_grad else None
        check(lib().jmm_observables(self.h, _dp(obs), _dp(g) if want_grad else None),
              "jmm_observables")
        obs = obs.reshape(self.n, 6)
        if not want_grad:
            return obs
        return obs, [g[o: o + n] for o, n in zip(self.offsets, self.Ns)]

    def observables_device(self, obs_ptr: int, grad_ptr: int | None = None):
        """observables() into device arrays (obs [n*6], grad [sum N]),
        queued on the engine's stream after the sweeps before it; sync()
        waits for it."""
        check(lib().jmm_observables_device(self.h, C.c_void_p(obs_ptr),
                                           C.c_void_p(grad_ptr) if grad_ptr else None),
              "jmm_observables_device")

    def hist_configure(self, every: int, rhonb: int = 0, rbw: float = 0.0, gns: int = 0,
                       gsw: float = 0.0, gnb: int = 0, gbw: float = 0.0):
        """Sample rho(x) and g(x) every `every` sweeps (0 = off)."""
        h = HistSpec()
        h.rhonb, h.rbw, h.gns, h.gsw, h.gnb, h.gbw = rhonb, rbw, gns, gsw, gnb, gbw
        self._hist = (rhonb, gns, gnb)
        check(lib().jmm_hist_configure(self.h, C.byref(h), every), "jmm_hist_configure")

    def hist_sample(self):
        check(lib().jmm_hist_sample(self.h), "jmm_hist_sample")

    def hist_read(self, reset: bool = False):
        """(rho[chain, rhonb], g[chain, gns, gnb], n_samples) counts."""
        rhonb, gns, gnb = self._hist
        rho = np.zeros(self.n * rhonb, dtype=np.uint64)
        g = np.zeros(max(self.n * gns * gnb, 1), dtype=np.uint64)
        ns = C.c_uint64()
        u64p = C.POINTER(C.c_uint64)
        check(lib().jmm_hist_read(self.h, rho.ctypes.data_as(u64p), g.ctypes.data_as(u64p),
                                  C.byref(ns), int(reset)), "jmm_hist_read")
        return (rho.reshape(self.n, rhonb), g[: self.n * gns * gnb].reshape(self.n, gns, gnb),
                ns.value)

    def kernel_time(self, reset: bool = False):
        n, ms, sw = C.c_uint64(), C.c_double(), C.c_uint64()
        check(lib().jmm_kernel_time(self.h, C.byref(n), C.byref(ms), C.byref(sw), int(reset)),
              "jmm_kernel_time")
        return n.value, ms.value, sw.value


def pair_observables(chain: Chain, x_ptr: int, L_ptr: int, obs_ptr: int, grad_ptr: int | None,
                     n_configs: int, precision: int = 64, stream: int | None = None):
    """Backward kernel on caller-owned device memory (raw pointers)."""
    s = chain.to_c()
    check(lib().jmm_pair_observables(C.byref(s), n_configs, C.c_void_p(x_ptr), C.c_void_p(L_ptr),
                                     C.c_void_p(obs_ptr),
                                     C.c_void_p(grad_ptr) if grad_ptr else None, precision,
                                     C.c_void_p(stream) if stream else None),
          "jmm_pair_observables")
