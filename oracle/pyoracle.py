"""ctypes view of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker.  The product package
never imports it.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

POT_LJ, POT_HARMONIC, POT_HRSW, POT_YUKAWA = 0, 1, 2, 3
WALLS, PERIODIC = 0, 1
NLT, NPT = 0, 1
VOL_LINEAR, VOL_LOG = 0, 1


class Spec(C.Structure):
    _fields_ = [
        ("potential", C.c_int32), ("boundary", C.c_int32), ("ensemble", C.c_int32),
        ("vol_move", C.c_int32), ("nbn", C.c_int32), ("shifted", C.c_int32),
        ("N", C.c_uint64), ("T", C.c_double), ("P", C.c_double), ("L0", C.c_double),
        ("max_step", C.c_double), ("max_dv", C.c_double), ("cutoff", C.c_double),
        ("pot_p", C.c_double * 4), ("seed", C.c_uint64), ("chain_id", C.c_uint32),
        ("eci", C.c_uint32), ("bcap", C.c_uint32), ("pad", C.c_uint32),
    ]


class Sample(C.Structure):
    _fields_ = [("U", C.c_double), ("L", C.c_double), ("acc_d", C.c_uint32), ("acc_v", C.c_uint32)]


class Flip(C.Structure):
    _fields_ = [("sweep", C.c_uint64), ("index", C.c_uint64), ("margin", C.c_double),
                ("bound", C.c_double), ("taken", C.c_int32), ("gated", C.c_int32)]


FLIP_DTYPE = np.dtype([("sweep", "u8"), ("index", "u8"), ("margin", "f8"), ("bound", "f8"),
                       ("taken", "i4"), ("gated", "i4")])


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        u32p = C.POINTER(C.c_uint32)
        u64p = C.POINTER(C.c_uint64)
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_u53.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_u53.restype = C.c_double
        L.orc_hist_x.argtypes = [C.POINTER(Spec), dp, C.c_double, C.c_uint32, C.c_double,
                                 C.c_uint32, C.c_double, C.c_uint32, C.c_double,
                                 C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.orc_chain_set_steps.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_u32.argtypes = [C.c_uint32]
        L.orc_u32.restype = C.c_double
        L.orc_det_exp.argtypes = [C.c_double]
        L.orc_det_exp.restype = C.c_double
        L.orc_phi.argtypes = [C.POINTER(Spec), C.c_double, C.c_double, dp]
        L.orc_phi_ljcut_ref.argtypes = [C.c_double, C.c_double, C.c_int, dp]
        L.orc_phi_harmonic_ref.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        L.orc_schedule.argtypes = [C.POINTER(Spec), u64p, u64p, u64p]
        L.orc_chain_new.argtypes = [C.POINTER(Spec)]
        L.orc_chain_new.restype = C.c_void_p
        L.orc_chain_free.argtypes = [C.c_void_p]
        L.orc_chain_get.argtypes = [C.c_void_p, dp, dp, dp, u64p]
        L.orc_chain_set.argtypes = [C.c_void_p, dp, C.c_double, C.c_uint64]
        L.orc_chain_counter.argtypes = [C.c_void_p, C.c_int]
        L.orc_chain_counter.restype = C.c_uint64
        L.orc_trial.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, dp]
        L.orc_trial_with.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.c_double, dp]
        L.orc_volume_trial.argtypes = [C.c_void_p, C.c_uint64, dp]
        L.orc_volume_trial_with.argtypes = [C.c_void_p, C.c_double, C.c_double, dp]
        L.orc_accept.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_vol_accept.argtypes = [C.c_double] * 7
        L.orc_sweeps.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(Sample), C.POINTER(C.c_uint8)]
        L.orc_sweeps_allpairs.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(Sample)]
        L.orc_sweeps_follow.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8),
                                        C.POINTER(Sample), C.POINTER(Flip), C.c_uint64]
        L.orc_sweeps_follow.restype = C.c_uint64
        L.orc_sweeps_virial.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(Sample), dp]
        L.orc_virials_x.argtypes = [C.POINTER(Spec), dp, C.c_double, dp]
        L.orc_e32_pair.argtypes = [C.POINTER(Spec), C.c_double]
        L.orc_e32_pair.restype = C.c_double
        L.orc_full_energy.argtypes = [C.c_void_p]
        L.orc_full_energy.restype = C.c_double
        L.orc_brute_energy.argtypes = [C.c_void_p]
        L.orc_brute_energy.restype = C.c_double
        L.orc_full_energy_x.argtypes = [C.POINTER(Spec), dp, C.c_double]
        L.orc_full_energy_x.restype = C.c_double
        L.orc_brute_energy_x.argtypes = [C.POINTER(Spec), dp, C.c_double]
        L.orc_brute_energy_x.restype = C.c_double
        L.orc_observables_x.argtypes = [C.POINTER(Spec), dp, C.c_double, dp, dp]
        L.orc_run_batch.argtypes = [C.POINTER(Spec), C.c_int, C.c_uint64, C.c_int, C.c_int, dp]
        L.orc_run_batch.restype = C.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_spec(N, T=1.0, P=0.0, L0=None, rho=None, potential=POT_LJ, boundary=PERIODIC,
              ensemble=NLT, vol_move=VOL_LOG, nbn=-1, shifted=1, max_step=0.5, max_dv=0.01,
              cutoff=3.0, pot_p=(1.0, 1.0, 0.0, 0.0), seed=1, chain_id=0, eci=0, bcap=256) -> Spec:
    s = Spec()
    s.potential, s.boundary, s.ensemble, s.vol_move = potential, boundary, ensemble, vol_move
    s.nbn, s.shifted = nbn, shifted
    s.N = N
    s.T, s.P = T, P
    if L0 is None:
        L0 = N / rho
    s.L0 = L0
    s.max_step, s.max_dv, s.cutoff = max_step, max_dv, cutoff
    for k in range(4):
        s.pot_p[k] = pot_p[k] if k < len(pot_p) else 0.0
    s.seed, s.chain_id, s.eci, s.bcap = seed, chain_id, eci, bcap
    return s


def spec_from_dict(d: dict) -> Spec:
    return make_spec(**d)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def schedule(spec: Spec):
    B, M, w = C.c_uint64(), C.c_uint64(), C.c_uint64()
    lib().orc_schedule(C.byref(spec), C.byref(B), C.byref(M), C.byref(w))
    return B.value, M.value, w.value


def phi(spec: Spec, d: float, L: float = 0.0):
    o = np.zeros(9)
    lib().orc_phi(C.byref(spec), d, L, _dp(o))
    return o


def full_energy_x(spec: Spec, x, L):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_full_energy_x(C.byref(spec), _dp(x), L)


def brute_energy_x(spec: Spec, x, L):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_brute_energy_x(C.byref(spec), _dp(x), L)


def observables_x(spec: Spec, x, L, want_grad=True):
    x = np.ascontiguousarray(x, dtype=np.float64)
    obs = np.zeros(6)
    g = np.zeros(len(x)) if want_grad else None
    lib().orc_observables_x(C.byref(spec), _dp(x), L, _dp(obs), _dp(g) if want_grad else None)
    return obs, g


class Chain:
    """One oracle chain; same spec, same Philox stream, same schedule as the engine."""

    COUNTERS = {"trials": 0, "acc_d": 1, "wall_rej": 2, "hc_rej": 3, "pair_terms": 4,
                "vol_trials": 5, "acc_v": 6, "recomputes": 7, "B": 100, "M_min": 101, "w": 102}

    def __init__(self, spec: Spec):
        self.spec = spec
        self.N = spec.N
        self.h = lib().orc_chain_new(C.byref(spec))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_chain_free(self.h)
            self.h = None

    def state(self):
        x = np.zeros(self.N)
        L, U, sw = C.c_double(), C.c_double(), C.c_uint64()
        lib().orc_chain_get(self.h, _dp(x), C.byref(L), C.byref(U), C.byref(sw))
        return x, L.value, U.value, sw.value

    def set_state(self, x, L, sweep=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        lib().orc_chain_set(self.h, _dp(x), L, sweep)

    def set_steps(self, max_step, max_dv):
        if lib().orc_chain_set_steps(self.h, max_step, max_dv) != 0:
            raise ValueError("max_step above the creation value")

    def counter(self, name):
        return lib().orc_chain_counter(self.h, self.COUNTERS[name])

    def sweeps(self, n, log=False):
        smp = (Sample * max(n, 1))()
        lg = np.zeros(n * self.N, dtype=np.uint8) if log else None
        lib().orc_sweeps(self.h, n, smp, lg.ctypes.data_as(C.POINTER(C.c_uint8)) if log else None)
        arr = np.array([(s.U, s.L, s.acc_d, s.acc_v) for s in smp[:n]],
                       dtype=[("U", "f8"), ("L", "f8"), ("acc_d", "u4"), ("acc_v", "u4")])
        return (arr, lg) if log else arr

    def sweeps_virial(self, n):
        """n sweeps; returns (samples, virials[n, 2]) with the virial and
        hypervirial of the configuration after each sweep."""
        smp = (Sample * max(n, 1))()
        v = np.zeros((max(n, 1), 2))
        lib().orc_sweeps_virial(self.h, n, smp, _dp(v))
        arr = np.array([(s.U, s.L, s.acc_d, s.acc_v) for s in smp[:n]],
                       dtype=[("U", "f8"), ("L", "f8"), ("acc_d", "u4"), ("acc_v", "u4")])
        return arr, v[:n]

    def sweeps_follow(self, n, accept_log, vol_acc=None, max_flips=1024):
        """Replays n sweeps taking the decisions of another form's run:
        accept_log [n, N] (or flat) one byte per displacement trial,
        vol_acc [n] the box-length decisions.  Returns (samples, flips,
        n_flips): flips a structured array (sweep, index, margin, bound,
        taken, gated) of the first max_flips decisions that differ from the
        oracle's own, index = 2**64-1 for a box-length trial."""
        lg = np.ascontiguousarray(np.asarray(accept_log, dtype=np.uint8).ravel())
        assert lg.size == n * self.N
        va = np.zeros(max(n, 1), dtype=np.uint8) if vol_acc is None else \
            np.ascontiguousarray(np.asarray(vol_acc, dtype=np.uint8).ravel())
        assert va.size >= n
        smp = (Sample * max(n, 1))()
        fl = (Flip * max(max_flips, 1))()
        u8p = C.POINTER(C.c_uint8)
        nf = lib().orc_sweeps_follow(self.h, n, lg.ctypes.data_as(u8p), va.ctypes.data_as(u8p), smp, fl,
                                     max_flips)
        arr = np.array([(s.U, s.L, s.acc_d, s.acc_v) for s in smp[:n]],
                       dtype=[("U", "f8"), ("L", "f8"), ("acc_d", "u4"), ("acc_v", "u4")])
        flips = np.array([(f.sweep, f.index, f.margin, f.bound, f.taken, f.gated)
                          for f in fl[:min(nf, max_flips)]], dtype=FLIP_DTYPE)
        return arr, flips, int(nf)

    def sweeps_allpairs(self, n):
        smp = (Sample * max(n, 1))()
        lib().orc_sweeps_allpairs(self.h, n, smp)
        return np.array([(s.U, s.L, s.acc_d, s.acc_v) for s in smp[:n]],
                        dtype=[("U", "f8"), ("L", "f8"), ("acc_d", "u4"), ("acc_v", "u4")])

    def trial(self, i, sweep):
        dE = C.c_double()
        a = lib().orc_trial(self.h, i, sweep, C.byref(dE))
        return a, dE.value

    def trial_with(self, i, md, u):
        dE = C.c_double()
        a = lib().orc_trial_with(self.h, i, md, u, C.byref(dE))
        return a, dE.value

    def volume_trial_with(self, dv, u):
        dU = C.c_double()
        a = lib().orc_volume_trial_with(self.h, dv, u, C.byref(dU))
        return a, dU.value

    def volume_trial(self, sweep):
        dU = C.c_double()
        a = lib().orc_volume_trial(self.h, sweep, C.byref(dU))
        return a, dU.value

    def full_energy(self):
        return lib().orc_full_energy(self.h)

    def brute_energy(self):
        return lib().orc_brute_energy(self.h)


def run_batch(specs, n_sweeps, n_threads, allpairs=False):
    arr = (Spec * len(specs))(*specs)
    U = np.zeros(len(specs))
    dt = lib().orc_run_batch(arr, len(specs), n_sweeps, n_threads, 1 if allpairs else 0, _dp(U))
    return dt, U


def u53(a, b):
    return lib().orc_u53(a, b)


def hist_x(spec: Spec, x, L, rhonb, rbw, gns, gsw, gnb, gbw):
    """(rho[rhonb], g[gns, gnb]) counts of one configuration."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    rho = np.zeros(rhonb, dtype=np.uint64)
    g = np.zeros(max(gns * gnb, 1), dtype=np.uint64)
    u64p = C.POINTER(C.c_uint64)
    lib().orc_hist_x(C.byref(spec), _dp(x), L, rhonb, rbw, gns, gsw, gnb, gbw,
                     rho.ctypes.data_as(u64p), g.ctypes.data_as(u64p))
    return rho, g[: gns * gnb].reshape(gns, gnb)


def virials_x(spec: Spec, x, L):
    """(W, HV) of one configuration."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    o = np.zeros(2)
    lib().orc_virials_x(C.byref(spec), _dp(x), L, _dp(o))
    return o


def e32_pair(spec: Spec, d: float) -> float:
    """Single-precision error bound of one pair-energy evaluation at d."""
    return lib().orc_e32_pair(C.byref(spec), d)


def u32(w):
    return lib().orc_u32(w)


def det_exp(x):
    return lib().orc_det_exp(x)
