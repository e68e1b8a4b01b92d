// mcstate.cpp -- the jmmMCState-style facade over the engine's C ABI.
//
// Output files keep the reference's formats: config.dat.mcs frames
// (src/jmmMCState.cpp:1009-1013) and the 13-column thermo.dat.mcs
// (header :566-568, rows :1916-1918).  Thermo block means are over per-sweep
// samples (the reference samples every trial), the virial and hypervirial
// columns included: the engine computes both after every sweep.
//
// Steps (STEPUNIT): TRIAL, the default, counts one step per Step() call as
// the reference does (src/jmmMCState.cpp:1734-1811), so every interval test
// of src/Main.cpp, getStepNum() % 10000 included, fires on the steps it
// fires on in the reference; every numTrialTypes steps queue one sweep.
// SWEEP advances the counter by numTrialTypes per Step() (one host call per
// sweep) and the interval tests fire when a multiple is crossed.
//
// Restart: every printCoords also writes checkpoint.dat.mcs, the engine
// state of every chain in binary (positions, length, running energy, sweep
// counter, step sizes), so a RESTART continues bit for bit where the text
// frame alone (L printed with %.5f, chain 0 only) could not; without a
// checkpoint the last complete config.dat.mcs frame is used as the
// reference does.  The output files are rewound to the restart step
// (src/jmmMCState.cpp:696-756): rows after it are removed.
#include "jmmMCState.h"

#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "jmm_engine.h"

// The reference's fad, incrementStep, Step, printG and relaxVolume are
// collective: src/Main.cpp:50-178 calls them from every thread of an OpenMP
// team and they share the work through orphaned single/barrier constructs.
// Here one thread of the team does the work and every thread returns its
// result; outside a parallel region (or with a team of one) they are plain
// calls.
template <class R, class F>
static R collective(F f)
{
    static R shared;
#pragma omp single
    shared = f();
    const R r = shared;
#pragma omp barrier
    return r;
}

struct MCState {
    MCInput inp;
    jmm_engine *eng = nullptr;
    std::vector<jmm_chain_spec> specs;
    unsigned long N = 0, numTrialTypes = 0, numSteps = 0, sn = 0, lastStep = 0;
    unsigned long pending = 0, cap = 4096, sweepsDone = 0;
    bool trialUnit = true;           // STEPUNIT TRIAL: one step per Step()
    bool refAdjust = false;          // ADJUST REFERENCE
    unsigned long trialInSweep = 0;  // steps of the sweep being counted (TRIAL)
    std::vector<double> vbuf;        // per-sweep virial and hypervirial
    double rhoSysA = 0, rhoSysSA = 0, lA = 0, lSA = 0, EA = 0, ESA = 0, lEA = 0;
    double VirA = 0, VirSA = 0, EVirA = 0, HVA = 0, HVSA = 0;
    unsigned long nSamples = 0;
    double E = 0, l = 0, Vir = 0, HV = 0;
    FILE *cf = nullptr, *tf = nullptr;
    std::vector<jmm_sample> buf;
    // rho(x)/g(x): engine counts pulled into per-file accumulators, so the
    // two print intervals reset independently as in the reference
    bool hist = false;
    FILE *rhof = nullptr;
    std::vector<FILE *> gf;
    std::vector<uint64_t> rhoAcc, gAcc, tmpR, tmpG;
    uint64_t nsRho = 0, nsG = 0;
    // step-size adaptation (DADJ/VADJ)
    std::vector<double> maxStep, maxDv;
    uint64_t vTrialsAtAdjust = 0;
};

static void die_on(int rc, const char *what)
{
    if (rc) std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, jmm_last_error());
}

static jmm_chain_spec spec_from_input(const MCInput &inp, unsigned long k, bool &ok)
{
    jmm_chain_spec s{};
    ok = true;
    const std::string pot = inp.potStr;
    s.shifted = 0;
    s.cutoff = inp.potCutOff;
    s.pot_p[0] = 1.0;
    s.pot_p[1] = 1.0;
    if (pot == "LJ") {
        s.potential = JMM_POT_LJ;
        s.cutoff = INFINITY;  // reference: LJ ignores the cutoff (src/jmmMCState.cpp:292-294)
    } else if (pot == "LJcut") {
        s.potential = JMM_POT_LJ;
    } else if (pot == "LJshift") {
        s.potential = JMM_POT_LJ;
        s.shifted = 1;
    } else if (pot == "HARMONIC") {
        s.potential = JMM_POT_HARMONIC;
    } else if (pot == "HRSW") {
        s.potential = JMM_POT_HRSW;
        s.pot_p[0] = 1.0;
        s.pot_p[1] = 1.0;
        s.pot_p[2] = 1.5;
    } else if (pot == "YUKAWA") {
        s.potential = JMM_POT_YUKAWA;
        s.pot_p[0] = 1.0;
        s.pot_p[1] = 0.2;
        s.pot_p[2] = 0.5;
        if (!std::isfinite(s.cutoff)) s.cutoff = 20.0;
    } else {
        ok = false;
    }
    for (int q = 0; q < inp.potParamsSet; ++q) s.pot_p[q] = inp.potParams[q];
    if (s.potential == JMM_POT_HRSW) s.cutoff = s.pot_p[2];
    const bool npt = std::strncmp(inp.ensembleStr, "NPT", 80) == 0;
    s.ensemble = npt ? JMM_NPT : JMM_NLT;
    s.boundary = inp.boundary == "PERIODIC" ? JMM_PERIODIC : JMM_WALLS;
    s.vol_move = inp.volMove == "LOG" ? JMM_VOL_LOG : JMM_VOL_LINEAR;
    s.nbn = inp.nbn;
    s.N = inp.N;
    s.T = inp.T;
    s.P = inp.P;
    // NPT starts at l = N (src/jmmMCState.cpp:549-555) unless LSTART is given
    s.L0 = npt ? (inp.lStart > 0 ? inp.lStart : (double)inp.N) : inp.L;
    // with DADJ adaptation the step may grow to the reference's clamp 0.5
    // (src/jmmMCState.cpp:2106-2111): the engine's step cap (and periodic
    // length bound) is built for the larger of the two and the INPUT value
    // is set right after creation
    const bool adapt = inp.mdai && inp.mdai < inp.ns;
    s.max_step = adapt ? std::max(inp.maxStep, 0.5) : inp.maxStep;
    s.max_dv = inp.maxdl;
    s.seed = inp.seed + k;
    s.chain_id = (uint32_t)k;
    s.eci = 0;  // the facade's ECheck does the reference's check-and-repair
    s.bcap = (uint32_t)inp.bcap;
    return s;
}

// Runs the queued sweeps and folds their samples (chain 0), the virial and
// hypervirial of every sweep included, into the thermo accumulators
// (updateThermo, src/jmmMCState.cpp:1941-1961, per sweep).
static int flush(MCState *m)
{
    while (m->pending) {
        const uint32_t n = (uint32_t)std::min(m->pending, m->cap);
        int rc = jmm_run_sweeps(m->eng, n);
        if (rc) return rc;
        uint32_t got = 0;
        m->buf.resize((size_t)m->cap * m->specs.size());
        m->vbuf.resize((size_t)m->cap * m->specs.size() * 2);
        rc = jmm_read_samples_virial(m->eng, m->buf.data(), m->vbuf.data(), (uint32_t)m->cap, &got);
        if (rc) return rc;
        for (uint32_t s = 0; s < got; ++s) {
            const jmm_sample &x = m->buf[s];
            m->Vir = m->vbuf[2 * s];
            m->HV = m->vbuf[2 * s + 1];
            const double rho = (double)m->N / x.L;
            m->rhoSysA += rho;
            m->rhoSysSA += rho * rho;
            m->lA += x.L;
            m->lSA += x.L * x.L;
            m->EA += x.U;
            m->ESA += x.U * x.U;
            m->lEA += x.L * x.U;
            m->VirA += m->Vir;
            m->VirSA += m->Vir * m->Vir;
            m->EVirA += x.U * m->Vir;
            m->HVA += m->HV;
            m->HVSA += m->HV * m->HV;
            m->nSamples++;
            m->E = x.U;
            m->l = x.L;
        }
        m->pending -= n;
        m->sweepsDone += n;
    }
    return 0;
}

// Binary checkpoint of every chain: "JMMCKPT1", then u64 chains, N, step;
// per chain u64 sweep counter and f64 length, running energy, max_step,
// max_dv, positions[N].  Written to a temporary file and renamed, so a
// checkpoint on disk is always whole.
static const char CKPT[] = "checkpoint.dat.mcs";

static int write_checkpoint(MCState *m)
{
    const size_t nc = m->specs.size();
    std::vector<double> x(nc * m->N), L(nc), U(nc);
    if (int rc = jmm_get_states(m->eng, x.data(), L.data(), U.data())) return rc;
    std::vector<jmm_counters> c(nc);
    if (int rc = jmm_get_counters(m->eng, -1, c.data())) return rc;
    const std::string tmp = std::string(CKPT) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return -1;
    const uint64_t hdr[3] = {(uint64_t)nc, (uint64_t)m->N, (uint64_t)m->sn};
    bool ok = std::fwrite("JMMCKPT1", 1, 8, f) == 8 && std::fwrite(hdr, 8, 3, f) == 3;
    for (size_t k = 0; k < nc && ok; ++k) {
        const uint64_t sw = c[k].sweeps;
        const double v[4] = {L[k], U[k], m->maxStep[k], m->maxDv[k]};
        ok = std::fwrite(&sw, 8, 1, f) == 1 && std::fwrite(v, 8, 4, f) == 4 &&
             std::fwrite(x.data() + k * m->N, 8, m->N, f) == m->N;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), CKPT) != 0) return -1;
    return 0;
}

struct Checkpoint {
    unsigned long step = 0;
    std::vector<uint64_t> sweep;
    std::vector<double> x, L, U, maxStep, maxDv;
};

static bool read_checkpoint(const MCState *m, Checkpoint &ck)
{
    FILE *f = std::fopen(CKPT, "rb");
    if (!f) return false;
    char magic[8];
    uint64_t hdr[3];
    const size_t nc = m->specs.size();
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "JMMCKPT1", 8) == 0 &&
              std::fread(hdr, 8, 3, f) == 3 && hdr[0] == nc && hdr[1] == m->N;
    if (ok) {
        ck.step = (unsigned long)hdr[2];
        ck.sweep.resize(nc);
        ck.x.resize(nc * m->N);
        ck.L.resize(nc);
        ck.U.resize(nc);
        ck.maxStep.resize(nc);
        ck.maxDv.resize(nc);
        for (size_t k = 0; k < nc && ok; ++k) {
            double v[4];
            ok = std::fread(&ck.sweep[k], 8, 1, f) == 1 && std::fread(v, 8, 4, f) == 4 &&
                 std::fread(ck.x.data() + k * m->N, 8, m->N, f) == m->N;
            ck.L[k] = v[0];
            ck.U[k] = v[1];
            ck.maxStep[k] = v[2];
            ck.maxDv[k] = v[3];
        }
    }
    std::fclose(f);
    return ok;
}

// Removes the rows of a step-keyed output file after the restart step (the
// reference overwrites them in place, src/jmmMCState.cpp:696-756): rows that
// start with a step number above `step` and everything after the first such
// row; rows without a leading number (headers) are kept.
static void rewind_rows(const char *name, unsigned long step)
{
    FILE *f = std::fopen(name, "r");
    if (!f) return;
    char *line = nullptr;
    size_t cap = 0;
    long keep = 0;
    ssize_t n;
    while ((n = getline(&line, &cap, f)) > 0) {
        unsigned long st;
        if (std::sscanf(line, "%lu", &st) == 1 && st > step) break;
        keep = std::ftell(f);
    }
    std::free(line);
    std::fclose(f);
    if (truncate(name, keep) != 0) std::printf("WARNING: could not rewind %s\n", name);
}

// The last complete config.dat.mcs frame at or before `upto` (any step when
// upto is ULONG_MAX): its step, length and positions, and the file offset
// just after it.  Frames are "N", "Step no.: S  Box length: L" and N
// "i 0.0 0.0 x" rows (src/jmmMCState.cpp:1009-1013).
static bool last_frame(const MCState *m, unsigned long upto, unsigned long &step, double &L,
                       std::vector<double> &x, long &end)
{
    FILE *f = std::fopen("config.dat.mcs", "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    std::vector<double> cur;
    while (std::fgets(line, sizeof line, f)) {
        unsigned long n, st;
        double Lf;
        if (std::sscanf(line, "%lu", &n) != 1 || std::strchr(line, ':') || std::strchr(line, '.') || n != m->N)
            continue;
        if (!std::fgets(line, sizeof line, f)) break;
        if (std::sscanf(line, "Step no.: %lu  Box length: %lf", &st, &Lf) != 2) continue;
        cur.clear();
        for (unsigned long i = 0; i < n && std::fgets(line, sizeof line, f); ++i) {
            unsigned long pn;
            double a, b, z;
            if (std::sscanf(line, "%lu %lf %lf %lf", &pn, &a, &b, &z) != 4) break;
            cur.push_back(z);
        }
        if (cur.size() == m->N && st <= upto) {
            x = cur;
            step = st;
            L = Lf;
            end = std::ftell(f);
            found = true;
        }
    }
    std::fclose(f);
    return found;
}

// Restart (src/jmmMCState.cpp:572-765): the checkpoint when there is one,
// else chain 0's last complete text frame; then the output files are cut
// back to the restart step.
static bool restart(MCState *m)
{
    Checkpoint ck;
    unsigned long step = 0;
    double Lf = 0.0;
    std::vector<double> xf;
    long end = 0;
    const size_t nc = m->specs.size();
    if (read_checkpoint(m, ck)) {
        const uint64_t swk = ck.step / m->numTrialTypes;
        if (m->trialUnit ? swk != ck.sweep[0] : false) {
            std::printf("ERROR: checkpoint step %lu does not match its sweep counter %llu\n", ck.step,
                        (unsigned long long)ck.sweep[0]);
            return false;
        }
        if (int rc = jmm_set_steps(m->eng, ck.maxStep.data(), ck.maxDv.data())) {
            die_on(rc, "jmm_set_steps");
            return false;
        }
        m->maxStep = ck.maxStep;
        m->maxDv = ck.maxDv;
        if (int rc = jmm_set_states(m->eng, ck.x.data(), ck.L.data(), ck.U.data(), ck.sweep.data())) {
            die_on(rc, "jmm_set_states");
            return false;
        }
        step = ck.step;
        std::printf("Checkpoint found for step number %lu\nBox length: %.17g\n", step, ck.L[0]);
        if (!last_frame(m, step, step, Lf, xf, end)) end = 0;
    } else {
        if (nc > 1) {
            std::printf("ERROR: restarting %zu chains needs %s (config.dat.mcs holds chain 0 only)\n", nc, CKPT);
            return false;
        }
        if (!last_frame(m, ~0ul, step, Lf, xf, end)) {
            std::printf("ERROR: no complete configuration in config.dat.mcs\n");
            return false;
        }
        std::printf("Complete configuration found for step number %lu\nBox length: %.8g\n", step, Lf);
        if (int rc = jmm_set_state(m->eng, 0, xf.data(), Lf, step / m->numTrialTypes)) {
            die_on(rc, "jmm_set_state");
            return false;
        }
    }
    if (truncate("config.dat.mcs", end) != 0 && end > 0) std::printf("WARNING: could not rewind config.dat.mcs\n");
    rewind_rows("thermo.dat.mcs", step);
    rewind_rows("rho.dat.mcs", step);
    for (int k = 0; k < m->inp.gns; ++k) {
        char name[64];
        std::snprintf(name, sizeof name, "g%d.dat.mcs", k);
        rewind_rows(name, step);
    }
    m->sn = m->lastStep = step;
    m->trialInSweep = m->trialUnit ? step % m->numTrialTypes : 0;
    return true;
}

static bool crossed(const MCState *m, unsigned long interval)
{
    if (!interval || m->sn == m->lastStep) return false;
    return m->sn / interval != m->lastStep / interval;
}

struct MCState *setupMCS(struct MCInput inp)
{
    if (inp.N < 1) {
        std::printf("FATAL ERROR: N must be at least 1.\nABORTING SIMULATION\n\n");
        return nullptr;
    }
    MCState *m = new MCState();
    m->inp = inp;
    m->N = inp.N;
    const bool npt = std::strncmp(inp.ensembleStr, "NPT", 80) == 0;
    m->numTrialTypes = npt ? inp.N + 1 : inp.N;
    m->inp.numTrialTypes = m->numTrialTypes;
    m->numSteps = inp.ns;
    std::printf("ENSEMBLE = %s\n", npt ? "NPT" : "NLT");
    for (unsigned long k = 0; k < std::max(1ul, inp.chains); ++k) {
        bool ok;
        m->specs.push_back(spec_from_input(inp, k, ok));
        if (!ok) {
            std::printf("FATAL ERROR: Unknown potential.\nABORTING SIMULATION\n\n");
            delete m;
            return nullptr;
        }
    }
    m->trialUnit = inp.stepUnit != "SWEEP";
    m->refAdjust = inp.adjust == "REFERENCE";
    if (m->trialUnit && inp.ns % m->numTrialTypes)
        std::printf("Note: NUMSTEPS %lu is not a whole number of sweeps of %lu trials; the last %lu "
                    "steps complete no sweep and are not run.\n",
                    inp.ns, m->numTrialTypes, inp.ns % m->numTrialTypes);
    jmm_engine_opts o{};
    o.form = inp.form == "PLAIN_FP64" ? JMM_FORM_PLAIN_FP64
           : inp.form == "BATCHED_FP32" ? JMM_FORM_BATCHED_FP32 : JMM_FORM_BATCHED_FP64;
    o.device = inp.device;
    o.sample_capacity = (uint32_t)m->cap;
    o.block_sweeps = (uint32_t)std::max(1ul, inp.sweepBlock);
    o.sample_virial = 1;
    int rc = jmm_create(m->specs.data(), (int)m->specs.size(), &o, &m->eng);
    if (rc) {
        die_on(rc, "jmm_create");
        delete m;
        return nullptr;
    }
    m->maxStep.assign(m->specs.size(), inp.maxStep);
    m->maxDv.assign(m->specs.size(), inp.maxdl);
    if (inp.mdai && inp.mdai < inp.ns) {
        rc = jmm_set_steps(m->eng, m->maxStep.data(), nullptr);
        if (rc) die_on(rc, "jmm_set_steps");
    }
    printMCP(m);
    jmm_counters c{};
    jmm_get_counters(m->eng, 0, &c);
    m->E = c.U;
    m->l = c.L;
    if (!inp.isRestart) {
        m->cf = std::fopen("config.dat.mcs", "w");
        m->tf = std::fopen("thermo.dat.mcs", "w");
        if (m->tf)
            std::fprintf(m->tf, "Step    Econf           Econf2          L       L2      LEconf          "
                                "rho             rho2            Virial         Virial2         EconfVir        "
                                "HV              HV2 \n");
    } else {
        if (!restart(m)) {
            freeMCS(m);
            return nullptr;
        }
        jmm_get_counters(m->eng, 0, &c);
        m->E = c.U;
        m->l = c.L;
        m->cf = std::fopen("config.dat.mcs", "a");
        m->tf = std::fopen("thermo.dat.mcs", "a");
    }
    // histograms: the first sample is the configuration the run starts from
    // (after a restart, the restored one)
    if ((inp.rhopi && inp.rhonb && inp.rbw > 0.0) || (inp.gpi && inp.gns > 0 && inp.gnb && inp.gbw > 0.0)) {
        jmm_hist_spec hs{};
        hs.rhonb = (uint32_t)std::max(1ul, inp.rhonb);
        hs.rbw = inp.rbw > 0.0 ? inp.rbw : 1.0;
        const bool g = inp.gns > 0 && inp.gnb && inp.gbw > 0.0 && inp.gsw > 0.0;
        hs.gns = g ? (uint32_t)inp.gns : 0;
        hs.gsw = g ? inp.gsw : 0.0;
        hs.gnb = g ? (uint32_t)inp.gnb : 0;
        hs.gbw = g ? inp.gbw : 0.0;
        rc = jmm_hist_configure(m->eng, &hs, (uint32_t)std::max(1ul, inp.histEvery));
        if (rc) {
            die_on(rc, "jmm_hist_configure");
        } else {
            m->hist = true;
            const size_t nc = m->specs.size();
            m->rhoAcc.assign(hs.rhonb, 0);
            m->gAcc.assign((size_t)hs.gns * hs.gnb, 0);
            m->tmpR.assign(nc * hs.rhonb, 0);
            m->tmpG.assign(std::max<size_t>(1, nc * hs.gns * hs.gnb), 0);
            const char *mode = inp.isRestart ? "a" : "w";
            m->rhof = std::fopen("rho.dat.mcs", mode);
            for (uint32_t k = 0; k < hs.gns; ++k) {
                char name[64];
                std::snprintf(name, sizeof name, "g%u.dat.mcs", k);
                m->gf.push_back(std::fopen(name, mode));
            }
            // the configuration at the start (a restart continues the
            // histogram blocks of the uninterrupted run: no extra sample)
            if (!inp.isRestart) jmm_hist_sample(m->eng);
        }
    }
    // step-0 virial (src/Main.cpp:66-68 computes it with fad): W from the
    // observables; HV = sigma dW/dsigma for LJ, the reference's only
    // potential with a hypervirial (zero otherwise until the first sweep)
    {
        std::vector<double> obs(6 * m->specs.size());
        if (jmm_observables(m->eng, obs.data(), nullptr) == 0) {
            m->Vir = obs[1];
            m->HV = m->specs[0].potential == JMM_POT_LJ ? m->specs[0].pot_p[1] * obs[5] : 0.0;
        }
    }
    return m;
}

void freeMCS(struct MCState *m)
{
    if (!m) return;
    if (m->cf) std::fclose(m->cf);
    if (m->tf) std::fclose(m->tf);
    if (m->rhof) std::fclose(m->rhof);
    for (FILE *f : m->gf)
        if (f) std::fclose(f);
    jmm_destroy(m->eng);
    delete m;
}

void printStep(struct MCState *m)
{
    std::printf("Step: %lu...\n", m->sn);
    std::fflush(stdout);
}

int printMCP(struct MCState *m)
{
    const MCInput &i = m->inp;
    if (i.isRestart) std::printf("*** THIS IS A RESTART RUN ***\n");
    std::printf("Printing Monte Carlo parameters...\nEnsemble: %s\n", i.ensembleStr);
    if (m->specs[0].ensemble == JMM_NPT) std::printf("N: %lu\nP: %.5G\nT: %.5G\n", i.N, i.P, i.T);
    else std::printf("N: %lu\nL: %.5G\nT: %.5G\n", i.N, i.L, i.T);
    std::printf("Potential: %s\nPotential cut-off: %.5G\n", i.potStr, m->specs[0].cutoff);
    if (i.nbn > 0) std::printf("Number of neighbors with which each particle can interact: %d\n", i.nbn);
    else std::printf("No neighbor number limit.\n");
    std::printf("numSteps: %lu\nmaxStep: %.5G\nmax vol change: %.5G\n", i.ns, i.maxStep, i.maxdl);
    std::printf("Boundary: %s\nVolume move: %s\nForm: %s\nChains: %zu\n\nSeed: %lu\n\n",
                i.boundary.c_str(), i.volMove.c_str(), i.form.c_str(), m->specs.size(), i.seed);
    std::fflush(stdout);
    return 0;
}

int fad(struct MCState *m, unsigned long *, double *)
{
    // the step-0 full energy (src/Main.cpp:66-68) is computed in FP64 at setup
    return collective<int>([m] {
        jmm_counters c{};
        int rc = jmm_get_counters(m->eng, 0, &c);
        m->E = c.U;
        return rc;
    });
}

unsigned long incrementStep(struct MCState *m)
{
    return collective<unsigned long>([m]() -> unsigned long {
        if (m->sn >= m->numSteps) return 0;
        m->lastStep = m->sn;
        m->sn += m->trialUnit ? 1 : m->numTrialTypes;
        if (m->sn / 10000 != m->lastStep / 10000) printStep(m);
        return m->sn;
    });
}

int Step(struct MCState *m)
{
    return collective<int>([m] {
        // TRIAL: a sweep is queued once its numTrialTypes steps are counted
        if (!m->trialUnit || ++m->trialInSweep == m->numTrialTypes) {
            m->trialInSweep = 0;
            m->pending++;
        }
        if (m->pending >= m->cap) {
            if (int rc = flush(m)) return rc;
        }
        if (isECheck(m)) ECheck(m);
        return 0;
    });
}

int isCoordPrint(struct MCState *m) { return crossed(m, m->inp.cpi); }
int isThermoPrint(struct MCState *m) { return crossed(m, m->inp.tpi); }
int isRhoPrint(struct MCState *m) { return crossed(m, m->inp.rhopi); }
int isGPrint(struct MCState *m) { return crossed(m, m->inp.gpi); }
int isECheck(struct MCState *m) { return crossed(m, m->inp.eci); }
int isMaxDisAdjust(struct MCState *m) { return crossed(m, m->inp.mdai); }
int isMaxDVAdjust(struct MCState *m) { return crossed(m, m->inp.mvai); }

int updateThermo(struct MCState *m)
{
    if (int rc = flush(m)) return rc;
    const double rho = (double)m->N / m->l;
    m->rhoSysA += rho;
    m->rhoSysSA += rho * rho;
    m->lA += m->l;
    m->lSA += m->l * m->l;
    m->EA += m->E;
    m->ESA += m->E * m->E;
    m->lEA += m->l * m->E;
    m->VirA += m->Vir;
    m->VirSA += m->Vir * m->Vir;
    m->EVirA += m->E * m->Vir;
    m->HVA += m->HV;
    m->HVSA += m->HV * m->HV;
    m->nSamples++;
    return 0;
}

int printThermo(struct MCState *m)
{
    if (int rc = flush(m)) return rc;
    const double ss = m->nSamples ? (double)m->nSamples : 1.0;
    if (m->tf) {
        std::fprintf(m->tf, "%lu\t%.8G\t%.8G\t%.8G\t%.8G\t%.8G\t%.8G\t%.8G\t%.8G\t%.8G\t%.8G\t%.8G\t%.8G\n",
                     m->sn, m->EA / ss, m->ESA / ss, m->lA / ss, m->lSA / ss, m->lEA / ss, m->rhoSysA / ss,
                     m->rhoSysSA / ss, m->VirA / ss, m->VirSA / ss, m->EVirA / ss, m->HVA / ss, m->HVSA / ss);
        std::fflush(m->tf);
    }
    std::printf("%lu  %.8G  %.8G  %.8G  %.8G\n", m->sn, m->E, m->l, m->Vir, m->HV);
    std::fflush(stdout);
    m->rhoSysA = m->rhoSysSA = m->lA = m->lSA = m->EA = m->ESA = m->lEA = 0;
    m->VirA = m->VirSA = m->EVirA = m->HVA = m->HVSA = 0;
    m->nSamples = 0;
    return 0;
}

int printCoords(struct MCState *m)
{
    if (int rc = flush(m)) return rc;
    std::vector<double> x(m->N);
    double L = 0;
    int rc = jmm_get_state(m->eng, 0, x.data(), &L, nullptr, nullptr);
    if (rc) return rc;
    if (!m->cf) return 0;
    std::fprintf(m->cf, "%lu\nStep no.: %lu  Box length: %.5f\n", m->N, m->sn, L);
    for (unsigned long i = 0; i < m->N; ++i) std::fprintf(m->cf, "%lu  0.0  0.0  %.8G\n", i, x[i]);
    std::fflush(m->cf);
    // the exact state of every chain at this frame, for a bit-exact restart
    if (write_checkpoint(m)) std::printf("WARNING: could not write %s\n", CKPT);
    return 0;
}

// Runs the queued sweeps and moves the engine's counts (chain 0, the
// reference's single chain) into both accumulators.
static int pull_hist(MCState *m)
{
    if (int rc = flush(m)) return rc;
    uint64_t ns = 0;
    int rc = jmm_hist_read(m->eng, m->tmpR.data(), m->tmpG.data(), &ns, 1);
    if (rc) return rc;
    for (size_t b = 0; b < m->rhoAcc.size(); ++b) m->rhoAcc[b] += m->tmpR[b];
    for (size_t b = 0; b < m->gAcc.size(); ++b) m->gAcc[b] += m->tmpG[b];
    m->nsRho += ns;
    m->nsG += ns;
    return 0;
}

// src/jmmMCState.cpp:1021-1037: one row "step rho_0 ... rho_{rhonb-1}" with
// rho_b = counts / samples / rbw.
int printRho(struct MCState *m)
{
    if (!m->hist || !m->rhof) return 0;
    if (int rc = pull_hist(m)) return rc;
    std::fprintf(m->rhof, "%lu", m->sn);
    const double ns = m->nsRho ? (double)m->nsRho : 1.0;
    for (size_t b = 0; b < m->rhoAcc.size(); ++b) {
        std::fprintf(m->rhof, " %.8G", (double)m->rhoAcc[b] / ns / m->inp.rbw);
        m->rhoAcc[b] = 0;
    }
    std::fprintf(m->rhof, "\n");
    std::fflush(m->rhof);
    m->nsRho = 0;
    return 0;
}

// src/jmmMCState.cpp:1042-1062: per segment s a row "step g_0 ... g_{gnb-1}"
// in g<s>.dat.mcs with g_b = counts / samples / gsw / gbw.
int printG(struct MCState *m)
{
    return collective<int>([m] {
        if (!m->hist || m->gf.empty()) return 0;
        if (int rc = pull_hist(m)) return rc;
        const size_t gnb = m->inp.gnb;
        const double ns = m->nsG ? (double)m->nsG : 1.0;
        for (size_t k = 0; k < m->gf.size(); ++k) {
            FILE *f = m->gf[k];
            if (!f) continue;
            std::fprintf(f, "%lu", m->sn);
            for (size_t b = 0; b < gnb; ++b) {
                std::fprintf(f, " %.8G", (double)m->gAcc[k * gnb + b] / ns / m->inp.gsw / m->inp.gbw);
                m->gAcc[k * gnb + b] = 0;
            }
            std::fprintf(f, "\n");
            std::fflush(f);
        }
        m->nsG = 0;
        return 0;
    });
}

int ECheck(struct MCState *m)
{
    if (int rc = flush(m)) return rc;
    std::vector<double> Uf(m->specs.size());
    int rc = jmm_full_energy(m->eng, Uf.data());
    if (rc) return rc;
    int result = 0;
    for (size_t k = 0; k < m->specs.size(); ++k) {
        jmm_counters c{};
        jmm_get_counters(m->eng, (int)k, &c);
        // absolute tolerance of src/jmmMCState.cpp:1998
        if (std::fabs(Uf[k] - c.U) > 0.0001) {
            std::printf("\n### !!! $$$ Energy discrepancy!!! ######### !!!!!!!!!!! $$$$$$$$$$ !!!!!!!!\n"
                        "Chain %zu: E = %.8G, ETest = %.8G, Resetting results to that of full calculation (E = %.8G)\n\n",
                        k, c.U, Uf[k], Uf[k]);
            jmm_perturb_energy(m->eng, (int)k, Uf[k] - c.U);
            result = 1;
        } else if (k == 0) {
            std::printf("####### Energy was just verified ########  E = %.8G ######## ETest = %.8G #######\n", c.U, Uf[k]);
        }
        if (k == 0) m->E = Uf[0];
    }
    return result;
}

// Swendsen's rule (Physics Procedia 15 (2011) 81): step' = step *
// ln(a P_ideal + b) / ln(a P + b), a = 0.672924, b = 0.0644284, P the
// cumulative acceptance ratio.  The reference writes the denominator as
// ln(a (P + b)) (src/jmmMCState.cpp:2105, :2128): ADJUST REFERENCE uses its
// form, the default ADJUST PUBLISHED the published one.
static double swendsen(double step, double ratio, bool ref)
{
    const double a = 0.672924, b = 0.0644284;
    return step * std::log(a * 0.5 + b) / (ref ? std::log(a * (ratio + b)) : std::log(a * ratio + b));
}

// src/jmmMCState.cpp:2100-2115: clamp [0.002, 0.5].
int maxDisAdjust(struct MCState *m)
{
    if (int rc = flush(m)) return rc;
    std::vector<jmm_counters> c(m->specs.size());
    if (int rc = jmm_get_counters(m->eng, -1, c.data())) return rc;
    for (size_t k = 0; k < c.size(); ++k) {
        if (!c[k].trials) continue;
        const double ratio = (double)c[k].acc_d / (double)c[k].trials;
        m->maxStep[k] = std::min(0.5, std::max(0.002, swendsen(m->maxStep[k], ratio, m->refAdjust)));
    }
    if (int rc = jmm_set_steps(m->eng, m->maxStep.data(), nullptr)) return rc;
    std::printf("Step: %lu  Updating max Step...dAcc: %lu,%lu   new maxStep: %.5G\n", m->sn,
                (unsigned long)c[0].acc_d, (unsigned long)(c[0].trials - c[0].acc_d), m->maxStep[0]);
    return 0;
}

// src/jmmMCState.cpp:2120-2139: only after new box-length trials; clamp to
// [0.002 N, 0.1 N] for the linear move, or with ADJUST REFERENCE the
// reference's rule, which sets 0.5 N when the new value exceeds 0.1 N
// (:2132-2133); [0.002, 0.1] in ln L for the log move (which the reference
// does not have).
int maxDVAdjust(struct MCState *m)
{
    if (int rc = flush(m)) return rc;
    std::vector<jmm_counters> c(m->specs.size());
    if (int rc = jmm_get_counters(m->eng, -1, c.data())) return rc;
    if (c[0].vol_trials <= m->vTrialsAtAdjust) return 0;
    m->vTrialsAtAdjust = c[0].vol_trials;
    const bool lg = m->inp.volMove == "LOG";
    const double lo = lg ? 0.002 : 0.002 * m->N, hi = lg ? 0.1 : 0.10 * m->N;
    for (size_t k = 0; k < c.size(); ++k) {
        if (!c[k].vol_trials) continue;
        const double ratio = (double)c[k].acc_v / (double)c[k].vol_trials;
        const double v = swendsen(m->maxDv[k], ratio, m->refAdjust);
        m->maxDv[k] = v < lo ? lo : v > hi ? ((m->refAdjust && !lg) ? 0.5 * m->N : hi) : v;
    }
    if (int rc = jmm_set_steps(m->eng, nullptr, m->maxDv.data())) return rc;
    std::printf("Step: %lu  Updating max vol.ch....vAcc: %lu,%lu   new maxStep: %.5G\n", m->sn,
                (unsigned long)c[0].acc_v, (unsigned long)(c[0].vol_trials - c[0].acc_v), m->maxDv[0]);
    return 0;
}

// src/jmmMCState.cpp:2396: Newton iteration on the box length.  With
// h = 0.1, E(l +- h) from the configuration scaled to l +- h, the step is
// dl = -(P - N T / l + dE/dl) / (d2E/dl2), clamped to 0.1 N and to the
// bracket of earlier lengths; converged when |dl| < 0.0025 N, at most 20
// steps.  Every chain relaxes at once; messages report chain 0.
int relaxVolume(struct MCState *m)
{
    return collective<int>([m]() -> int {
        if (int rc = flush(m)) return rc;
        const size_t n = m->specs.size();
        std::vector<jmm_counters> c(n);
        if (int rc = jmm_get_counters(m->eng, -1, c.data())) return rc;
        std::vector<double> l(n), E(n), lo(n, 0.0), hi(n, 1e10), Lu(n), Ld(n), Eu(n), Ed(n), Ln(n);
        std::vector<char> done(n, 0);
        for (size_t k = 0; k < n; ++k) {
            l[k] = c[k].L;
            E[k] = c[k].U;
        }
        std::printf("Relaxing the Volume...Starting Length,Energy: %.5G,%.5G\n", l[0], E[0]);
        const double h = 0.1, N = (double)m->N;
        for (int count = 0; count < 20; ++count) {
            std::printf("Relaxation step %d...", count);
            for (size_t k = 0; k < n; ++k) {
                Lu[k] = l[k] + h;
                Ld[k] = l[k] - h;
            }
            if (int rc = jmm_scaled_energy(m->eng, Lu.data(), Eu.data())) return rc;
            if (int rc = jmm_scaled_energy(m->eng, Ld.data(), Ed.data())) return rc;
            bool all = true;
            for (size_t k = 0; k < n; ++k) {
                Ln[k] = l[k];
                if (done[k]) continue;
                const double first = (Eu[k] - Ed[k]) / (2 * h);
                const double second = (Eu[k] - 2.0 * E[k] + Ed[k]) / (h * h);
                double dl = -(m->inp.P - (N / l[k]) * m->inp.T + first) / second;
                const double relaxMax = 0.10 * N;
                if (std::fabs(dl) > relaxMax) dl = dl < 0 ? -relaxMax : relaxMax;
                if (l[k] + dl > hi[k]) dl = 0.5 * (hi[k] - l[k]);
                else if (l[k] + dl < lo[k]) dl = 0.5 * (lo[k] - l[k]);
                if (dl > 0.0) lo[k] = l[k];
                else hi[k] = l[k];
                if (std::fabs(dl) < 0.0025 * N) done[k] = 1;
                else all = false;
                Ln[k] = l[k] + dl;
            }
            if (int rc = jmm_set_lengths(m->eng, Ln.data())) {
                std::printf("\nWARNING: relaxation stopped: %s\n", jmm_last_error());
                return rc;
            }
            if (int rc = jmm_get_counters(m->eng, -1, c.data())) return rc;
            for (size_t k = 0; k < n; ++k) {
                l[k] = c[k].L;
                E[k] = c[k].U;
            }
            m->E = E[0];
            m->l = l[0];
            if (all) {
                std::printf("Relaxation converged. Length,energy: %.10G,%.10G\n", l[0], E[0]);
                return 0;
            }
            std::printf("length,energy: %.5g,%.5G\n", l[0], E[0]);
        }
        std::printf("WARNING: Relaxation failed to converge after 20 steps. Length,energy: %.5G,%.5G\n",
                    l[0], E[0]);
        return 1;
    });
}

int isRelaxStep(struct MCState *m)
{
    // src/Main.cpp:173-176: every 10000 trials below step 1e6 when RELAX is
    // set; with STEPUNIT TRIAL exactly the reference's getStepNum() % 10000 == 0
    return m->inp.relaxFlag > 0 && m->sn < 1000000 && crossed(m, 10000);
}

unsigned long getStepNum(struct MCState *m) { return m->sn; }
int getRelaxFlag(struct MCState *m) { return m->inp.relaxFlag; }
bool getRestartFlag(struct MCState *m) { return m->inp.isRestart; }

int printE(struct MCState *m)
{
    flush(m);
    jmm_counters c{};
    jmm_get_counters(m->eng, 0, &c);
    std::printf("\nE = %.8G\n", c.U);
    return 0;
}

int printAcc(struct MCState *m)
{
    flush(m);
    jmm_counters c{};
    jmm_get_counters(m->eng, 0, &c);
    std::printf("Accepted/Rejected: Displacements VolumeChanges\n              "
                "        %lu/%lu          %lu/%lu\n",
                (unsigned long)c.acc_d, (unsigned long)(c.trials - c.acc_d), (unsigned long)c.acc_v,
                (unsigned long)(c.vol_trials - c.acc_v));
    return 0;
}

struct jmm_engine *getEngine(struct MCState *m) { return m->eng; }

double getEnergy(struct MCState *m, int chain)
{
    flush(m);
    jmm_counters c{};
    jmm_get_counters(m->eng, chain, &c);
    return c.U;
}

double getLength(struct MCState *m, int chain)
{
    flush(m);
    jmm_counters c{};
    jmm_get_counters(m->eng, chain, &c);
    return c.L;
}
