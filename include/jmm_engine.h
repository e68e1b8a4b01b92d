/*
 * jmm_engine.h -- C ABI of the MI355X (gfx950) Metropolis engine for 1D
 * classical particles: many independent chains of single-particle
 * displacement trials plus box-length trials, run by hand-written HIP kernels.
 *
 * Every entry point returns 0 on success or a negative JMM_E* code; the text
 * of the last error on the calling thread is available from jmm_last_error().
 * No entry point throws or exits.  An engine handle is not thread-safe: call
 * it from one host thread.  All device work is queued on the engine's own HIP
 * stream; host output arrays are written only by the calls that document it.
 *
 * Citations are to the reference repository (mmansell7/jmmOneDMC); each
 * entry point names the reference interface it replaces.
 */
#ifndef JMM_ENGINE_H
#define JMM_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JMM_ABI_VERSION 1

/* status codes */
#define JMM_OK 0
#define JMM_EINVAL (-1)   /* bad argument or chain specification */
#define JMM_EHIP (-2)     /* HIP runtime failure */
#define JMM_ENOMEM (-3)   /* device or host allocation failed */
#define JMM_EFULL (-4)    /* sample buffer full: read samples first */
#define JMM_ESTATE (-5)   /* call not valid in the engine's current state */

/* potentials (reference POT keyword, src/readInput.cpp:104-131, plus the
 * configurations' hard-rod/square-well and Yukawa potentials) */
#define JMM_POT_LJ 0       /* 4 eps((s/d)^12-(s/d)^6); p = {eps, sigma, exclusion}; inclusive
                              cutoff; cutoff = INFINITY is the reference's "LJ"; a trial
                              that would bring two particles closer than the exclusion
                              distance (0: 0.5 sigma, where e^(-phi/T) < 1e-3500 at T <= 2)
                              is rejected without evaluating it */
#define JMM_POT_HARMONIC 1 /* (d-1)^2, 1e11 for d <= 0, strict cutoff (reference HARMONIC) */
#define JMM_POT_HRSW 2     /* hard rod p0, well depth p1 out to p2 (strict) */
#define JMM_POT_YUKAWA 3   /* p0 exp(-p1 d)/d, hard core p2, inclusive cutoff */

#define JMM_WALLS 0        /* hard walls at +-L/2, reference behaviour */
#define JMM_PERIODIC 1     /* periodic line of length L */
#define JMM_NLT 0          /* constant length (reference "NLT", i.e. NVT) */
#define JMM_NPT 1          /* constant pressure: one box-length trial per sweep */
#define JMM_VOL_LINEAR 0   /* L' = L + U(-dv, dv), Jacobian N ln s (reference fav) */
#define JMM_VOL_LOG 1      /* ln L' = ln L + U(-dv, dv), Jacobian (N+1) ln s */

/* execution forms (all three run the same Markov chain) */
#define JMM_FORM_PLAIN_FP64 0   /* one chain per wavefront, trials one at a time,
                                   energy change = wavefront reduction over neighbours */
#define JMM_FORM_BATCHED_FP64 1 /* persistent work-queue kernel, one lane per block
                                   of particles, chains packed into workgroups */
#define JMM_FORM_BATCHED_FP32 2 /* as BATCHED_FP64 with FP32 pair energies summed
                                   in FP64; positions, cutoff decisions and
                                   accumulators FP64 */

/* One chain: the hot-path fields of the reference's MCInput
 * (src/jmmMCState.h:66-101) plus the engine's additive fields. */
typedef struct jmm_chain_spec {
    int32_t potential;  /* JMM_POT_* (reference POT) */
    int32_t boundary;   /* JMM_WALLS / JMM_PERIODIC */
    int32_t ensemble;   /* JMM_NLT / JMM_NPT (reference ENSEMBLE) */
    int32_t vol_move;   /* JMM_VOL_LINEAR / JMM_VOL_LOG */
    int32_t nbn;        /* reference NBN: partners farther than nbn in index
                           order are ignored; -1 = no limit */
    int32_t shifted;    /* 1: phi(d) - phi(cutoff) inside the cutoff */
    uint64_t N;         /* reference N */
    double T;           /* reference T (k_B = 1) */
    double P;           /* reference P */
    double L0;          /* initial length; lattice start x_i = ((i+1/2)/N - 1/2) L0 */
    double max_step;    /* reference MAXSTEP: displacement uniform in +-max_step */
    double max_dv;      /* reference MAXDV (linear) or the ln L half-width (log) */
    double cutoff;      /* reference POT cutoff; INFINITY allowed with walls */
    double pot_p[4];    /* potential parameters, see JMM_POT_* */
    uint64_t seed;      /* reference SEED: Philox4x32-10 key */
    uint32_t chain_id;  /* Philox counter word: distinct streams per seed */
    uint32_t eci;       /* reference ENGCHECK, in sweeps: FP64 recompute interval, 0 = never */
    uint32_t bcap;      /* largest number of particle blocks, at most 1024 (0 = 256).  A
                           power of two gives power-of-two block counts; a multiple
                           of 64 that is not one gives block counts that are
                           multiples of 64 (whole waves) when at least 64; such a
                           chain is never split over workgroups */
    uint32_t reserved;
} jmm_chain_spec;

typedef struct jmm_engine_opts {
    int32_t form;             /* JMM_FORM_* */
    int32_t device;           /* HIP device ordinal */
    uint32_t sample_capacity; /* per-chain sweeps held on device between reads (0 = 1024) */
    uint32_t block_sweeps;    /* sweeps per kernel launch (0 = 100) */
    uint32_t wg_threads;      /* batched forms: lanes per workgroup, rounded up to 256, 512 or
                                 1024 and to the largest chain's block count (0 = 256) */
    uint32_t use_graph;       /* 1: replay sweep blocks from a captured hipGraph */
    uint32_t log_accepts;     /* 1: record one byte per displacement trial (tests) */
    uint32_t lds_budget;      /* batched forms: LDS bytes per workgroup (0 = 64 KiB) */
    uint32_t seg_lds;         /* batched forms: a chain whose LDS image exceeds this many bytes
                                 (0 = the device's LDS per workgroup) is split into segments of
                                 consecutive blocks, one workgroup each, neighbours
                                 synchronising every (block length - window) phases; a
                                 chain that cannot be split runs from HBM */
    uint32_t seg_lanes;       /* split chains: lanes per block, 1 or 2 (two lanes share each
                                 trial, one per side; 0 = 1); results are the same bits
                                 either way */
    uint32_t sample_virial;   /* 1: after every sweep each chain's virial W = sum -d phi'(d)
                                 and hypervirial sum -d W'(d) are computed (FP64) and kept
                                 with the samples (jmm_read_samples_virial) */
    uint32_t reserved;
} jmm_engine_opts;

/* One per chain per sweep (the per-sweep form of the reference's per-trial
 * updateThermo, src/jmmMCState.cpp:1941-1961). */
typedef struct jmm_sample {
    double U;       /* configurational energy after the sweep */
    double L;       /* box length after the sweep */
    uint32_t acc_d; /* accepted displacement trials in the sweep */
    uint32_t acc_v; /* accepted box-length trials in the sweep (0/1) */
} jmm_sample;

/* Cumulative per-chain counters (reference dAcc/vAcc, src/jmmMCState.cpp:66-69). */
typedef struct jmm_counters {
    uint64_t sweeps;
    uint64_t trials;        /* displacement trials */
    uint64_t acc_d;         /* accepted displacements */
    uint64_t wall_rejects;  /* rejected at the walls */
    uint64_t core_rejects;  /* rejected by the hard-core rule (crossings included) */
    uint64_t pair_terms;    /* pair evaluations in displacement trials (n_old + n_new) */
    uint64_t vol_trials;
    uint64_t acc_v;
    uint64_t recomputes;    /* FP64 full-energy recomputes */
    double max_drift;       /* largest |U_running - U_recomputed| seen */
    double U;               /* current running energy */
    double L;               /* current box length */
} jmm_counters;

/* Density and pair-separation histograms (reference RHONB/RBW, GNS/GSW,
 * GNB/GBW; src/jmmMCState.h:60-61, :128-132). */
typedef struct jmm_hist_spec {
    uint32_t rhonb;  /* rho(x) bins, width rbw, centred on x = 0 */
    uint32_t gns;    /* g(x) segments of width gsw, centred on x = 0 (0 = no g(x)) */
    uint32_t gnb;    /* g(x) separation bins of width gbw */
    uint32_t reserved;
    double rbw, gsw, gbw;
} jmm_hist_spec;

/* Where a chain runs (see jmm_chain_layout). */
enum {
    JMM_MODE_LDS = 0,    /* whole chain in one workgroup's LDS */
    JMM_MODE_HBM = 1,    /* positions read and written in HBM */
    JMM_MODE_SPLIT = 2   /* split over several cooperating workgroups */
};

typedef struct jmm_layout {
    uint32_t B;          /* particle blocks (one lane each) */
    uint32_t w;          /* partner window: neighbours a trial may reach per side */
    uint32_t mode;       /* JMM_MODE_* */
    uint32_t segments;   /* workgroups the chain spans (1 unless split) */
    uint64_t lds_bytes;  /* LDS image of the chain, or of one segment when split */
    uint32_t lanes;      /* lanes per block (2 when split chains share trials over two lanes) */
    uint32_t reserved;
} jmm_layout;

typedef struct jmm_engine jmm_engine;

/* Library identity; no reference counterpart. */
int jmm_abi_version(void);
const char *jmm_last_error(void);

/* The trial schedule of one chain: B blocks of particles, M_min = smallest
 * index separation of trials that run at once, w = largest index distance of
 * an interacting partner (floor(rc / g_min) with a 1e-3 guard, g_min the
 * hard core, which every state keeps between neighbours; M_min = w + 1,
 * at least 2; NBN caps both).  Pure host function; no reference
 * counterpart (the reference draws trial types with replacement,
 * src/jmmMCState.cpp:1762). */
int jmm_schedule(const jmm_chain_spec *spec, uint64_t *B, uint64_t *M_min, uint64_t *w);

/* Replaces setupMCS (src/jmmMCState.cpp:261-790: parameter copy :267-290,
 * potential select :291-399, ensemble :401-425, lattice :549-562, RNG
 * :779-781) and the first-step fad energy (src/Main.cpp:66-68), for n_chains
 * chains at once.  Uploads the lattice start and computes each chain's FP64
 * energy on the GPU.  *out receives the handle. */
int jmm_create(const jmm_chain_spec *specs, int n_chains, const jmm_engine_opts *opts,
               jmm_engine **out);

/* Replaces freeMCS (src/jmmMCState.cpp:792-842). */
int jmm_destroy(jmm_engine *e);

/* Replaces the chain loop of src/Main.cpp:114-180 (incrementStep
 * src/jmmMCState.cpp:1734-1754 + Step :1758-1811): n_sweeps sweeps on every
 * chain, each sweep = N displacement trials (qad2, :1160-1464) in the
 * engine's block order, then one box-length trial for NPT chains (fav,
 * :2161-2293), then the ECheck-style FP64 recompute (:1965-2095) when due.
 * Asynchronous; samples accumulate on the device. */
int jmm_run_sweeps(jmm_engine *e, uint32_t n_sweeps);

/* Waits for queued work; no reference counterpart. */
int jmm_synchronize(jmm_engine *e);

/* Replaces ECheck's recompute (src/jmmMCState.cpp:1973-1993) as a query:
 * U_out[c] = FP64 energy of chain c's current configuration.  Synchronous;
 * does not modify the chains. */
int jmm_full_energy(jmm_engine *e, double *U_out);

/* Replaces updateThermo/printThermo's data path (src/jmmMCState.cpp:
 * 1896-1961): copies every chain's samples since the last read into
 * out[c * max_sweeps + s] and empties the device buffer.  *n_out = sweeps
 * copied per chain (all chains have the same count).  The samples are put
 * in this layout on the device (the first read allocates a staging buffer of
 * sample_capacity x n_chains x 24 bytes, kept until destroy) and arrive in
 * one copy.  Synchronous. */
int jmm_read_samples(jmm_engine *e, jmm_sample *out, uint32_t max_sweeps, uint32_t *n_out);

/* jmm_read_samples plus, for an engine created with sample_virial, the
 * virial and hypervirial after every sweep: virial[(c * max_sweeps + s) * 2]
 * = W, [... + 1] = HV (the reference's per-trial Vir/HV sums of
 * updateThermo, src/jmmMCState.cpp:1941-1961, per sweep).  virial may be
 * NULL; a non-NULL virial on an engine without sample_virial is
 * JMM_ESTATE.  Drains the buffer as jmm_read_samples.  Synchronous. */
int jmm_read_samples_virial(jmm_engine *e, jmm_sample *out, double *virial, uint32_t max_sweeps,
                            uint32_t *n_out);

/* jmm_read_samples without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the transposition
 * is staged in a fixed 16 MiB device array that the library holds per GPU
 * until the process ends, instead of the buffer jmm_read_samples allocates
 * on first use; no hipMalloc, hipFree or allocating library call is made.
 * At most (16 MiB - 4 bytes per chain, rounded up to 256) / 24 bytes of
 * chain-sweeps per call (about 699,000): a larger read is refused with
 * JMM_EINVAL and drains nothing.  Same output, layout and draining as
 * jmm_read_samples.  All engines on one GPU share the staging array: calls
 * that overlap on one GPU race.  Synchronous. */
int jmm_read_samples_noalloc(jmm_engine *e, jmm_sample *out, uint32_t max_sweeps, uint32_t *n_out);

/* Replaces printCoords (src/jmmMCState.cpp:1007-1017) and the restart read
 * (:572-765): chain c's positions (sorted), length, running energy and sweep
 * counter.  Any output pointer may be NULL.  Synchronous. */
int jmm_get_state(jmm_engine *e, int chain, double *x, double *L, double *U, uint64_t *sweep);

/* Bulk form of jmm_get_state for every chain: x[sum N] with chain c at
 * offset sum_{c'<c} N_c' (chain order of creation), L[n_chains], U[n_chains].
 * Any output may be NULL.  Synchronous. */
int jmm_get_states(jmm_engine *e, double *x, double *L, double *U);

/* Restart: installs positions (must be sorted and respect the hard core),
 * length and sweep counter; the running energy is recomputed in FP64.
 * Synchronous. */
int jmm_set_state(jmm_engine *e, int chain, const double *x, double L, uint64_t sweep);

/* Bulk form of jmm_set_state for every chain (restart of a whole engine, or
 * production runs started from one equilibrated configuration): x as in
 * jmm_get_states, L[n_chains]; U[n_chains] installs the running energies
 * as given (a checkpoint's, for a bit-exact continuation) or, when NULL,
 * recomputes them in FP64; sweep[n_chains] (nullable) sets the sweep
 * counters, which key the random numbers.  Every chain is validated before
 * any is changed.  Synchronous.  Replaces the reference's restart read
 * (src/jmmMCState.cpp:572-765) for many chains at once. */
int jmm_set_states(jmm_engine *e, const double *x, const double *L, const double *U, const uint64_t *sweep);

/* Step sizes between sweeps, for step-size adaptation (the reference's
 * maxDisAdjust/maxDVAdjust, src/jmmMCState.cpp:2100-2139, rescale maxStep and
 * maxdl in place): max_step[c] and max_dv[c] per chain, NULL leaves that kind
 * unchanged.  max_step may not exceed the chain's value at creation (the
 * periodic length bound, L > 2 (rc + max_step), is built for it).
 * Synchronous. */
int jmm_set_steps(jmm_engine *e, const double *max_step, const double *max_dv);

/* FP64 energy of each chain's configuration rescaled to length L_try[c]
 * (x -> x L_try/L), the state unchanged (calculateEnergyOfTrialVolumeChange,
 * src/jmmMCState.cpp:2783-2827, used by relaxVolume :2396).
 * U_out[n_chains].  Synchronous. */
int jmm_scaled_energy(jmm_engine *e, const double *L_try, double *U_out);

/* Rescales every chain to length L_new[c] (moveVolume,
 * src/jmmMCState.cpp:2831-2916): x -> x L_new/L, energy recomputed in FP64.
 * Refused if a pair would fall inside the hard core, a particle outside the
 * walls, or a periodic length below 2 (cutoff + max_step).  Synchronous. */
int jmm_set_lengths(jmm_engine *e, const double *L_new);

/* Fault injection for the drift guard: adds delta to chain c's running
 * energy (the next recompute must detect and repair it).  Synchronous. */
int jmm_perturb_energy(jmm_engine *e, int chain, double delta);

/* Replaces printAcc/printE (src/jmmMCState.cpp:2143-2156). chain = -1
 * fills out[0..n_chains) for every chain in one copy. Synchronous. */
int jmm_get_counters(jmm_engine *e, int chain, jmm_counters *out);

/* Accept log (opts.log_accepts): out[c][s][i] one byte per displacement
 * trial for the sweeps since the last read, c in chain order, particle index
 * order; same read/drain rules as jmm_read_samples.  out must hold
 * sum_c N_c * max_sweeps bytes; chain c starts at offset
 * max_sweeps * sum_{c'<c} N_c'. */
int jmm_read_accept_log(jmm_engine *e, uint8_t *out, uint32_t max_sweeps, uint32_t *n_out);

/* Replaces the histogram bookkeeping of fgrho/ugrho/qagrho
 * (src/jmmMCState.cpp:1069-1149, :2297-2384): every `every` sweeps (counted
 * from creation) each chain's current configuration is binned: rho bin
 * floor(x/rbw + rhonb/2) per particle, and for every pair closer than
 * gnb*gbw the separation bin floor(d/gbw) once in the segment
 * floor(x/gsw + gns/2) of each member (periodic chains: positions folded into
 * [-L/2, L/2), each pair at its separation below L/2).  The reference samples
 * after every trial; the engine samples configurations between sweeps.
 * every = 0 switches sampling off.  Clears all counts.  Synchronous. */
int jmm_hist_configure(jmm_engine *e, const jmm_hist_spec *spec, uint32_t every);

/* One extra sample of the current configurations now (the reference's
 * step-0 fgrho/ugrho, src/Main.cpp:66-68 via fad).  Asynchronous. */
int jmm_hist_sample(jmm_engine *e);

/* Replaces printRho/printG's accumulators (src/jmmMCState.cpp:1021-1060):
 * rho[c*rhonb + b] and g[(c*gns + s)*gnb + b] receive the counts summed over
 * the *n_samples samples since the last reset (the reference prints
 * rhoA/ns/rbw and gA/ns/gsw/gbw).  Either array may be NULL.  reset = 1
 * clears the counts.  Synchronous. */
int jmm_hist_read(jmm_engine *e, uint64_t *rho, uint64_t *g, uint64_t *n_samples, int reset);

/* Placement of chain `chain` (external index): block count, window, memory
 * mode and the number of workgroups it spans.  A diagnostic with no
 * reference counterpart (the reference's chains are host loops).
 * Synchronous only in that it reads host-side tables. */
int jmm_chain_layout(jmm_engine *e, int chain, jmm_layout *out);

/* Kernel timing of the sweep kernel, from HIP events recorded on the
 * engine's stream around every sweep-block launch since the last reset:
 * number of launches, total milliseconds, total sweeps.  The engine keeps
 * at most 64 event pairs in flight (older ones are read into the totals as
 * the queue advances), so long runs hold a bounded number of events.
 * Synchronous. */
int jmm_kernel_time(jmm_engine *e, uint64_t *launches, double *ms, uint64_t *sweeps, int reset);

/* Observables of every chain's current configuration, for the
 * differentiation path (no reference counterpart; the reference keeps the
 * virial totals Vir/Vir6/Vir12, src/jmmMCState.cpp:86-96):
 * obs[c*6 + 0..5] = U, W = -sum d phi'(d), dU/dtheta0, dU/dtheta1,
 * dW/dtheta0, dW/dtheta1 (theta = (eps, sigma) for LJ, (A, kappa) for
 * Yukawa).  grad (nullable) receives dU/dx for all chains, chain c at offset
 * sum_{c'<c} N_c'.  FP32 pair arithmetic for the FP32 form, FP64 otherwise;
 * reductions in FP64.  Synchronous. */
int jmm_observables(jmm_engine *e, double *obs, double *grad);

/* jmm_observables into caller-owned DEVICE arrays (obs_dev[n*6], grad_dev
 * nullable [sum N]), same layout and arithmetic, for a consumer on the GPU
 * (the autograd path, fluctuation-gradient sums): the kernel runs on the
 * engine's stream after the sweeps launched before it and is not waited for
 * (jmm_synchronize waits).  No reference counterpart. */
int jmm_observables_device(jmm_engine *e, double *obs_dev, double *grad_dev);

/* The backward kernel on caller-owned DEVICE memory (for the autograd
 * binding): n_configs configurations of N particles of one potential.
 * x_dev[n_configs*N] (sorted per configuration), L_dev[n_configs];
 * obs_dev[n_configs*6] as in jmm_observables; grad_dev (nullable)
 * [n_configs*N] = dU/dx.  precision: 32 or 64.  stream: hipStream_t or NULL.
 * Every gap must be at least the potential's hard-core distance (0.5 sigma
 * for LJ, p2 for Yukawa, p0 for hard rods), as in any state a chain visits:
 * partners are searched within floor((cutoff + max_step)/hard core) indices.
 * Asynchronous on that stream.  No reference counterpart. */
int jmm_pair_observables(const jmm_chain_spec *spec, int n_configs, const double *x_dev,
                         const double *L_dev, double *obs_dev, double *grad_dev, int precision,
                         void *stream);

#ifdef __cplusplus
}
#endif
#endif
