"""Host side: the INPUT parser (against the reference's own parser test table,
src/testInput.c:15-63, and the four INPUT files it ships), the C ABI's
exports, and (GPU) the driver replayed against the oracle.
"""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import oracle_spec, rel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "jmmonedmc_amd", "build", "jmm_onedmc")
GOLDEN = os.path.join(ROOT, "tests", "golden", "inputs")


def parse(path):
    r = subprocess.run([DRIVER, "--print-input", path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    fields = dict(line.split("=", 1) for line in r.stdout.splitlines() if "=" in line and
                  not line.startswith("#"))
    return fields, r.stdout


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        from jmmonedmc_amd import build
        build.build_driver()
    return DRIVER


def test_reference_parser_table(driver, tmp_path):
    """The key/value table of src/testInput.c:15-63, written with the same
    spacing; the unknown POT "YT" falls back to LJ and YADA only warns."""
    p = tmp_path / "TEST1INPUT"
    p.write_text("N      5\nP      0.5\nT   0.5\nNUMSTEPS\t\t2\nPOT\t    YT\nNBN\t\t-1\n"
                 "MAXSTEP    0.1\nMAXDV     1\nCPI\t\t   1\nTPI\t\t1\nRBW\t\t0.01\nRHONB\t\t1000\n"
                 "RHOPI\t\t1\nGSW\t\t1\nGNS\t\t5\nGBW\t\t0.01\nGNB\t\t1000\nGPI\t\t1\nSEED\t\t125\n"
                 "YADA\t\t16.543\nENGCHECK\t\t1\nDADJ\t\t1\nVADJ\t\t1")
    f, out = parse(str(p))
    want = dict(N="5", ns="2", nbn="-1", cpi="1", tpi="1", rhonb="1000", rhopi="1", gns="5",
                gnb="1000", gpi="1", seed="125", eci="1", mdai="1", mvai="1", potStr="LJ")
    for k, v in want.items():
        assert f[k] == v, (k, f[k])
    for k, v in dict(P=0.5, T=0.5, maxStep=0.1, maxdl=1.0, rbw=0.01, gsw=1.0, gbw=0.01).items():
        assert float(f[k]) == v, k
    assert f["potCutOff"] == "inf"
    assert "Unknown potential type requested (YT)" in out
    assert "Property command YADA not understood." in out


@pytest.mark.parametrize("name,checks", [
    ("test_INPUT", dict(ensembleStr="NLT", N="2000", L=4000.0, T=0.5, relaxFlag="1", ns="1000",
                        potStr="LJcut", potCutOff=5.0, nbn="-1", maxStep=0.1, maxdl=2.0,
                        rhopi="50000000", seed="774281", eci="1", mdai="10000000")),
    ("test_INPUTstd", dict(N="10", P=0.7, T=0.4, ns="10", potStr="HARMONIC", nbn="1", seed="125",
                           gnb="100000", ensembleStr="NPT")),
    ("test_INPUT_smalltest", dict(N="10", P=1.0, T=0.9, relaxFlag="1", ns="5000000", potStr="LJ",
                                  seed="92847", eci="1000")),
    ("scripts_INPUT", dict(N="2000", P=0.7, T=0.8, ns="5000000000", potStr="LJ", nbn="2",
                           maxdl=5.0, seed="92847", eci="10000")),
])
def test_shipped_inputs_parse(driver, name, checks):
    f, out = parse(os.path.join(GOLDEN, name))
    for k, v in checks.items():
        if isinstance(v, float):
            assert float(f[k]) == v, (k, f[k])
        else:
            assert f[k] == v, (k, f[k])
    if name == "test_INPUTstd":
        assert "Property command YADA not understood." in out


def test_parser_edge_cases(driver, tmp_path):
    p = tmp_path / "INPUT"
    p.write_text("# comment line\n\nN 1E3   # scientific notation for an integer\n"
                 "ENSEMBLE XYZ\nPOT LJshift 2.5\nRESTART false\nBOUNDARY periodic\nVOLMOVE log\n"
                 "POTPARAMS 1.5 0.9\nCHAINS 4\n" + "T " + "1" * 150 + "\n")
    f, out = parse(str(p))
    assert f["N"] == "1000"
    assert f["ensembleStr"] == "NPT" and "Unknown ensemble type requested (XYZ)" in out
    assert f["potStr"] == "LJshift" and float(f["potCutOff"]) == 2.5
    assert f["isRestart"] == "0"
    assert f["boundary"] == "PERIODIC" and f["volMove"] == "LOG" and f["chains"] == "4"
    assert f["potParams"].startswith("1.5,0.9") and f["potParamsSet"] == "2"
    assert float(f["T"]) > 1e149        # a 150-digit value on a long line


def test_library_exports_every_declared_symbol():
    """libjmm_engine.so loads without a GPU and exports every function that
    include/jmm_engine.h declares."""
    import ctypes
    from jmmonedmc_amd import _native
    hdr = open(os.path.join(ROOT, "include", "jmm_engine.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(jmm_\w+)\s*\(", hdr, re.M))
    assert declared >= set(_native.EXPORTED)
    lib = ctypes.CDLL(_native.LIB)
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert lib.jmm_abi_version() == 1


def test_schedule_and_validation_without_gpu():
    """Host-only C ABI calls work without a device: the schedule, and
    argument validation errors surface through jmm_last_error."""
    import ctypes as C
    from jmmonedmc_amd import _native
    from jmmonedmc_amd.configs import c3
    from jmmonedmc_amd.engine import Chain, schedule
    ch = c3(8)
    assert schedule(ch[0])[0] >= 1
    lib = _native.lib()
    bad = Chain(N=10, L0=5.0, cutoff=3.0)           # lattice spacing 0.5 ok, periodic L too short
    s = bad.to_c()
    h = C.c_void_p()
    rc = lib.jmm_create(C.byref(s), 1, None, C.byref(h))
    assert rc != 0
    assert b"periodic length" in lib.jmm_last_error() or b"device" in lib.jmm_last_error()


def test_lj_exclusion_distance_schedule(oracle):
    """The schedule from the hard core alone, engine and oracle alike:
    w = floor(rc / g + 1e-3), M_min = w + 1.  C2's chain at the default LJ
    exclusion 0.5 sigma: w = 6, M_min = 7, 256 blocks (the default cap) or
    512 with bcap 1024; at 0.75 sigma w = 4, M_min = 5.  Hard rod + square
    well (g = 1, rc = 1.5): w = 1, M_min = 2; C5's Yukawa (g = 0.5, rc = 20):
    w = 40, M_min = 41, 1024 blocks.  rc / g just below an integer (within
    the guard) counts as that integer; further below, one partner less.  An
    exclusion outside the cutoff is refused."""
    import ctypes as C
    from jmmonedmc_amd import _native
    from jmmonedmc_amd.configs import c2, c3_params, c5
    from jmmonedmc_amd.engine import Chain, schedule
    from helpers import hrsw, lj, oracle_spec

    def both(ch):
        got = tuple(schedule(ch))
        assert got == tuple(oracle.schedule(oracle_spec(oracle, ch))), got
        return got

    assert both(c2(1, bcap=0)[0]) == (256, 7, 6)
    assert both(c2(1)[0]) == (512, 7, 6)
    assert both(c2(1, exclusion=0.75)[0]) == (512, 5, 4)
    assert both(Chain(N=4096, rho=0.5, **hrsw())) == (256, 2, 1)
    assert both(Chain(N=40, rho=0.5, **hrsw())) == (16, 2, 1)
    assert both(c5(1)[0]) == (1024, 41, 40)
    for gap, w in ((0.5, 6), (0.5 + 4e-5, 6), (0.5 + 0.01, 5)):
        assert both(Chain(N=4096, rho=0.5, cutoff=3.0, pot_p=(1.0, 1.0, gap, 0.0))) [1:] == (w + 1, w)
    lib = _native.lib()
    s = Chain(N=10, L0=40.0, cutoff=3.0, pot_p=(1.0, 1.0, 3.5, 0.0)).to_c()
    h = C.c_void_p()
    assert lib.jmm_create(C.byref(s), 1, None, C.byref(h)) != 0
    assert b"exclusion" in lib.jmm_last_error()


def test_block_counts_in_whole_waves(oracle):
    """A bcap that is not a power of two (a multiple of 64) gives block
    counts that are multiples of 64 (floor(N / M_min / 64) * 64, at most
    bcap), or a power of two below 64; engine and oracle alike.  Other
    bcap values are refused."""
    import ctypes as C
    from jmmonedmc_amd import _native
    from jmmonedmc_amd.engine import Chain, schedule
    from helpers import lj, oracle_spec
    for N, bcap, B in ((4096, 448, 448), (4096, 0, 256), (1400, 448, 192), (2500, 448, 320), (600, 448, 64),
                       (300, 448, 32), (8192, 448, 448), (4096, 512, 512), (6000, 960, 832)):
        ch = Chain(N=N, rho=0.5, bcap=bcap, **lj())
        got = tuple(schedule(ch))
        assert got[0] == B and got == tuple(oracle.schedule(oracle_spec(oracle, ch))), (N, bcap, got)
        assert N // got[0] >= got[1]
    lib = _native.lib()
    for bad in (100, 2048):
        s = Chain(N=64, L0=128.0, bcap=bad, **lj()).to_c()
        h = C.c_void_p()
        assert lib.jmm_create(C.byref(s), 1, None, C.byref(h)) != 0
        assert b"bcap" in lib.jmm_last_error()


@pytest.mark.gpu
def test_driver_replays_oracle(gpu, oracle, driver, tmp_path):
    """The driver on a C1-shaped INPUT (periodic, shifted LJ, NVT): the final
    configuration and energy equal the oracle's replay of the same sweeps."""
    import pyoracle
    from jmmonedmc_amd.engine import Chain
    from helpers import lj
    inp = tmp_path / "INPUT"
    inp.write_text("ENSEMBLE NLT\nN 64\nL 128\nT 1.0\nPOT LJshift 3.0\nNBN -1\nMAXSTEP 0.5\n"
                   "NUMSTEPS 64000\nCPI 64000\nTPI 6400\nSEED 1\nENGCHECK 32000\nBOUNDARY PERIODIC\n"
                   "FORM BATCHED_FP64\n")
    print("[gpu] driver C1 1000 sweeps", flush=True)
    r = subprocess.run([driver, str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "PROGRAM COMPLETED SUCCESSFULLY!" in r.stdout
    assert "Energy was just verified" in r.stdout
    E = float(re.search(r"\nE = (\S+)", r.stdout).group(1))
    ch = Chain(N=64, L0=128.0, T=1.0, seed=1, chain_id=0, **lj())
    c = oracle.Chain(oracle_spec(pyoracle, ch))
    c.sweeps(1000)
    x, L, U, _ = c.state()
    assert rel(E, U) <= 1e-7              # printed with %.8G
    frames = (tmp_path / "config.dat.mcs").read_text().split("Step no.:")
    last = [float(l.split()[3]) for l in frames[-1].splitlines()[1:] if l.strip() and len(l.split()) == 4]
    assert len(last) == 64
    assert np.allclose(last, x, rtol=1e-7, atol=1e-6)
    rows = (tmp_path / "thermo.dat.mcs").read_text().splitlines()
    assert rows[0].startswith("Step") and len(rows) == 1 + 1 + 10   # header, step 0, 10 blocks
    assert len(rows[-1].split("\t")) == 13


@pytest.mark.gpu
def test_driver_team_sizes_give_identical_output(gpu, driver, tmp_path):
    """The driver keeps the reference's OpenMP team structure (collective
    incrementStep/Step/fad/printG, output in single regions): a team of 4
    threads writes the same files as a team of 1."""
    text = ("ENSEMBLE NPT\nN 200\nP 0.5\nT 1.0\nPOT LJshift 3.0\nMAXSTEP 0.5\nMAXDV 0.01\n"
            "NUMSTEPS 402000\nCPI 100500\nTPI 20100\nSEED 3\nENGCHECK 40200\nBOUNDARY PERIODIC\n"
            "VOLMOVE LOG\nLSTART 400\nFORM BATCHED_FP32\n")
    outs = []
    for threads in ("1", "4"):
        d = tmp_path / f"t{threads}"
        d.mkdir()
        (d / "INPUT").write_text(text)
        print(f"[gpu] driver OMP_NUM_THREADS={threads}", flush=True)
        env = dict(os.environ, OMP_NUM_THREADS=threads)
        r = subprocess.run([driver, str(d / "INPUT")], cwd=d, capture_output=True, text=True,
                           timeout=120, env=env)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert f"Running with {threads} threads" in r.stdout
        outs.append(((d / "thermo.dat.mcs").read_text(), (d / "config.dat.mcs").read_text()))
    assert outs[0] == outs[1]
    assert len(outs[0][0].splitlines()) == 1 + 1 + 20


@pytest.mark.gpu
def test_driver_writes_density_and_pair_histograms(gpu, oracle, driver, tmp_path):
    """RHOPI/GPI output (printRho/printG, src/jmmMCState.cpp:1021-1062):
    rho.dat.mcs rows integrate to N (every particle inside the binned range),
    the step-0 row is the lattice, and g<s>.dat.mcs rows equal the oracle's
    binning of the same sweeps (walls, FP64 form, one sample per sweep)."""
    import pyoracle
    from jmmonedmc_amd.engine import Chain
    from jmmonedmc_amd import _native as nat
    from helpers import lj
    inp = tmp_path / "INPUT"
    N = 40
    inp.write_text(f"ENSEMBLE NLT\nN {N}\nL 80\nT 1.0\nPOT LJcut 3.0\nMAXSTEP 0.5\n"
                   f"NUMSTEPS {N * 50}\nCPI {N * 50}\nTPI {N * 10}\nRHOPI {N * 10}\nGPI {N * 25}\n"
                   "RBW 2.0\nRHONB 40\nGSW 40.0\nGNS 2\nGBW 0.5\nGNB 12\nSEED 5\nENGCHECK 0\n"
                   "BOUNDARY WALLS\nFORM BATCHED_FP64\nHISTEVERY 1\n")
    print("[gpu] driver histograms", flush=True)
    r = subprocess.run([driver, str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = [l.split() for l in (tmp_path / "rho.dat.mcs").read_text().splitlines()]
    assert [int(row[0]) for row in rows] == [0, 400, 800, 1200, 1600, 2000]
    for row in rows:
        vals = np.array([float(v) for v in row[1:]])
        assert len(vals) == 40
        assert abs(vals.sum() * 2.0 - N) < 1e-5 * N
    lattice = ((np.arange(N) + 0.5) / N - 0.5) * 80.0
    want0 = np.bincount(np.floor(lattice / 2.0 + 20).astype(int), minlength=40) / 2.0
    assert np.allclose([float(v) for v in rows[0][1:]], want0)
    # g files: replay 50 sweeps in the oracle, one sample per sweep
    ch = Chain(N=N, L0=80.0, T=1.0, seed=5, boundary=nat.WALLS, **lj(shifted=0))
    orc = pyoracle.Chain(oracle_spec(pyoracle, ch))

    def sample():
        x, L, _, _ = orc.state()
        return pyoracle.hist_x(orc.spec, x, L, 40, 2.0, 2, 40.0, 12, 0.5)[1].astype(float)

    g0 = sample()                          # printed at step 0 from one sample
    acc = [np.zeros((2, 12)), np.zeros((2, 12))]
    for sw in range(50):
        orc.sweeps(1)
        acc[0 if sw < 25 else 1] += sample()
    for s in range(2):
        grows = [l.split() for l in (tmp_path / f"g{s}.dat.mcs").read_text().splitlines()]
        assert [int(g[0]) for g in grows] == [0, 1000, 2000]
        for row, want in zip(grows, (g0[s] / 1, acc[0][s] / 25, acc[1][s] / 25)):
            got = np.array([float(v) for v in row[1:]])
            assert np.allclose(got, want / 40.0 / 0.5, rtol=1e-7, atol=0), (s, row[0])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["PUBLISHED", "REFERENCE"])
def test_driver_step_adaptation(gpu, driver, tmp_path, mode):
    """DADJ/VADJ (maxDisAdjust/maxDVAdjust, src/jmmMCState.cpp:2100-2139):
    every adjustment follows Swendsen's rule from the cumulative acceptance
    counts it prints, within the reference's clamps.  ADJUST PUBLISHED (the
    default): ln(a 0.5 + b) / ln(a P + b), box step clamped to [0.002 N,
    0.1 N].  ADJUST REFERENCE: the reference's ln(a 0.5 + b) / ln(a (P + b))
    (:2105, :2128) and its box step set to 0.5 N above 0.1 N (:2132-2133)."""
    inp = tmp_path / "INPUT"
    N = 60
    inp.write_text(f"ENSEMBLE NPT\nN {N}\nP 1.0\nT 0.8\nPOT LJshift 3.0\nMAXSTEP 0.05\nMAXDV 1.0\n"
                   f"NUMSTEPS {(N + 1) * 400}\nCPI {(N + 1) * 400}\nTPI {(N + 1) * 100}\n"
                   f"DADJ {(N + 1) * 50}\nVADJ {(N + 1) * 50}\nSEED 11\nBOUNDARY WALLS\nLSTART 90\n"
                   f"ADJUST {mode}\n")
    print(f"[gpu] driver step adaptation {mode}", flush=True)
    r = subprocess.run([driver, str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    a, b = 0.672924, 0.0644284
    ref = mode == "REFERENCE"
    for key, start, lo, hi in (("max Step", 0.05, 0.002, 0.5), ("max vol.ch", 1.0, 0.002 * N, 0.1 * N)):
        lines = re.findall(r"Updating " + re.escape(key) + r"\.\.\.\.?\w+: (\d+),(\d+)\s+new maxStep: (\S+)",
                           r.stdout)
        assert len(lines) == 8, (key, lines)
        prev = start
        for acc, rej, new in lines:
            ratio = int(acc) / (int(acc) + int(rej))
            den = math.log(a * (ratio + b)) if ref else math.log(a * ratio + b)
            v = prev * math.log(a * 0.5 + b) / den
            if key == "max Step":
                want = min(hi, max(lo, v))
            else:
                want = lo if v < lo else (0.5 * N if ref else hi) if v > hi else v
            assert float(new) == pytest.approx(want, rel=2e-4), (key, prev, ratio, new)
            prev = float(new)


@pytest.mark.gpu
def test_driver_relax_volume(gpu, driver, tmp_path):
    """RELAX (relaxVolume, src/jmmMCState.cpp:2396) on the harmonic
    nearest-neighbour chain between walls, where the relaxed length of the
    scaled lattice solves P - N T / l + 2 (N-1)(l/N - 1)/N = 0 (E = (N-1)
    (l/N - 1)^2 for lattice gaps l/N): the Newton iteration reaches that root
    and reports convergence."""
    from scipy.optimize import brentq
    inp = tmp_path / "INPUT"
    N, P, T = 50, 1.0, 0.3
    inp.write_text(f"ENSEMBLE NPT\nN {N}\nP {P}\nT {T}\nPOT HARMONIC\nNBN 1\nMAXSTEP 0.1\nMAXDV 0.5\n"
                   f"NUMSTEPS {N + 1}\nCPI {N + 1}\nTPI {N + 1}\nSEED 2\nRELAX\nBOUNDARY WALLS\n"
                   "LSTART 75\nFORM BATCHED_FP64\n")
    print("[gpu] driver relax", flush=True)
    r = subprocess.run([driver, str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Relaxing the Volume...Starting Length,Energy: 75" in r.stdout
    m = re.search(r"Relaxation converged\. Length,energy: (\S+),(\S+)", r.stdout)
    assert m, r.stdout[-3000:]
    lf, Ef = float(m.group(1)), float(m.group(2))
    root = brentq(lambda l: P - N * T / l + 2 * (N - 1) * (l / N - 1) / N, 20.0, 75.0)
    assert abs(lf - root) < 0.01, (lf, root)
    assert Ef == pytest.approx((N - 1) * (lf / N - 1) ** 2, rel=1e-8)


@pytest.mark.gpu
def test_driver_thermo_virial_columns_are_block_means(gpu, oracle, driver, tmp_path):
    """thermo.dat.mcs's Virial, Virial2, EconfVir, HV and HV2 columns
    (printThermo, src/jmmMCState.cpp:1916-1918) are block means of the
    per-sweep virial and hypervirial, as the U columns are: against the
    oracle's replay of the same sweeps (FP64 form, walls, truncated LJ)."""
    import pyoracle
    from jmmonedmc_amd.engine import Chain
    from jmmonedmc_amd import _native as nat
    from helpers import lj
    N, blocks, per = 64, 4, 50
    inp = tmp_path / "INPUT"
    inp.write_text(f"ENSEMBLE NLT\nN {N}\nL 110\nT 0.9\nPOT LJcut 3.0\nMAXSTEP 0.4\n"
                   f"NUMSTEPS {N * per * blocks}\nCPI {N * per * blocks}\nTPI {N * per}\nSEED 7\n"
                   "BOUNDARY WALLS\nFORM BATCHED_FP64\n")
    print("[gpu] driver thermo virial columns", flush=True)
    r = subprocess.run([driver, str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = [l.split("\t") for l in (tmp_path / "thermo.dat.mcs").read_text().splitlines()[2:]]
    assert len(rows) == blocks
    ch = Chain(N=N, L0=110.0, T=0.9, seed=7, boundary=nat.WALLS, max_step=0.4, **lj(shifted=0))
    orc = pyoracle.Chain(oracle_spec(pyoracle, ch))
    smp, vir = orc.sweeps_virial(per * blocks)
    for b, row in enumerate(rows):
        sl = slice(b * per, (b + 1) * per)
        U, W, H = smp["U"][sl], vir[sl, 0], vir[sl, 1]
        want = [U.mean(), (U * U).mean(), W.mean(), (W * W).mean(), (U * W).mean(), H.mean(), (H * H).mean()]
        got = [float(row[k]) for k in (1, 2, 8, 9, 10, 11, 12)]
        assert int(row[0]) == N * per * (b + 1)
        for g, w in zip(got, want):
            assert g == pytest.approx(w, rel=2e-7, abs=1e-9), (b, got, want)


def _run(driver, d, text, timeout=180):
    (d / "INPUT").write_text(text)
    r = subprocess.run([driver, str(d / "INPUT")], cwd=d, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    return r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["BATCHED_FP64", "BATCHED_FP32"])
def test_driver_restart_continues_bit_for_bit(gpu, driver, tmp_path, form):
    """RESTART (src/jmmMCState.cpp:572-765) from the binary checkpoint
    written with every config frame: a run stopped past its last frame and
    restarted ends with the same config.dat.mcs, thermo.dat.mcs and
    rho.dat.mcs, byte for byte, as the run that was never stopped; the rows
    the stopped run wrote after its last frame are removed on restart, and
    every chain of a multi-chain run is restored."""
    N, K = 40, 200  # sweeps to the checkpoint
    nt = N + 1
    base = (f"ENSEMBLE NPT\nN {N}\nP 0.6\nT 1.0\nPOT LJshift 3.0\nMAXSTEP 0.5\nMAXDV 0.02\n"
            f"CPI {nt * K}\nTPI {nt * K // 4}\nRHOPI {nt * K // 4}\nRBW 4.0\nRHONB 40\nSEED 13\n"
            f"ENGCHECK {nt * 50}\nBOUNDARY PERIODIC\nVOLMOVE LOG\nLSTART 80\nFORM {form}\nCHAINS 3\n")
    a = tmp_path / "a"
    b = tmp_path / "b"
    a.mkdir()
    b.mkdir()
    print(f"[gpu] driver restart {form}", flush=True)
    _run(driver, a, base + f"NUMSTEPS {nt * 2 * K}\n")
    _run(driver, b, base + f"NUMSTEPS {nt * (K + K // 2)}\n")
    stopped = (b / "thermo.dat.mcs").read_text()
    assert f"\n{nt * (K + K // 4)}\t" in stopped          # rows past the checkpoint exist ...
    out = _run(driver, b, base + f"NUMSTEPS {nt * 2 * K}\nRESTART TRUE\n")
    assert "Checkpoint found for step number" in out
    for name in ("config.dat.mcs", "thermo.dat.mcs", "rho.dat.mcs"):
        assert (a / name).read_text() == (b / name).read_text(), name   # ... and are rewritten


@pytest.mark.gpu
def test_driver_relax_trigger_counts_trials(gpu, driver, tmp_path):
    """With the default STEPUNIT TRIAL a step is one trial, as in the
    reference, so src/Main.cpp:173's getStepNum() % 10000 == 0 relax
    trigger fires at steps 10000, 20000 and 30000 of a 24-trial-per-sweep
    chain (whose sweeps end on no multiple of 10000 until step 30000)."""
    N = 23
    out = _run(driver, tmp_path, f"ENSEMBLE NPT\nN {N}\nP 1.0\nT 0.3\nPOT HARMONIC\nNBN 1\nMAXSTEP 0.1\n"
                                 f"MAXDV 0.5\nNUMSTEPS 30000\nCPI 30000\nTPI 24000\nSEED 2\nRELAX\n"
                                 "BOUNDARY WALLS\nLSTART 35\nFORM BATCHED_FP64\n")
    assert out.count("Relaxing the Volume") == 4, out[-3000:]
    steps = [int(s) for s in re.findall(r"Step: (\d+)\.\.\.", out)]
    assert steps[:4] == [0, 10000, 20000, 30000]
