/*
 * oracle.h -- CPU restatement of jmmOneDMC's Metropolis chain loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in the product links it, and the product has no CPU
 * fallback that routes through it.
 *
 * What it restates (reference file:line, relative to the reference repo):
 *   pair potentials          src/pot.cpp:19-141 (phiLJcut, phiLJinfcutoff,
 *                            phiHarmoniccut) + the hard-rod/square-well and
 *                            Yukawa potentials the configurations add
 *   displacement trial       src/jmmMCState.cpp:1160-1464 (qad2)
 *   volume trial             src/jmmMCState.cpp:2161-2293 (fav),
 *                            src/jmmMCState.cpp:1648-1730 (qavLJ)
 *   full energy              src/jmmMCState.cpp:853-1003 (fad),
 *                            src/jmmMCState.cpp:1965-2095 (ECheck)
 *   lattice start            src/jmmMCState.cpp:559-562
 *   thermo samples           src/jmmMCState.cpp:1941-1961 (updateThermo)
 *
 * The sweep order, the Philox keying and the hard-core rule are the engine's
 * (they replace the reference's with-replacement trial selection and GSL
 * taus2 stream); they are restated here so the oracle replays the engine move
 * for move.
 *
 * Parity pinning, per function (tests/test_oracle.py, tests/test_exact_1d.py):
 *   orc_philox4x32_10  Random123 known-answer vectors + rocRAND's
 *                      philox4x32_10 engine on the host
 *   orc_phi            closed forms (LJ zero at sigma, minimum -eps at
 *                      2^(1/6) sigma with zero virial, value at the cutoff,
 *                      inclusive/strict cutoffs, harmonic penalty) and
 *                      central finite differences for the virial/gradients
 *   orc_trial          brute-force O(N^2) recompute before/after on tiny N;
 *                      hand-chosen (x, md, u) accept triples; its pair-term
 *                      counter against n_old + n_new counted over every
 *                      partner and periodic image
 *   orc_full_energy    brute-force all-pairs sum (orc_brute_energy, which has
 *                      no window logic) on random and lattice inputs
 *   orc_volume_trial   hand-chosen (L, dl, dU, u) acceptance cases and the
 *                      LJ scaling identity for an infinite cutoff
 *   orc_sweep          exact 1D NPT averages (Takahashi gas of independent
 *                      gaps, integrated by quadrature) and the exact
 *                      averages of two and three particles between walls
 *                      (configurational integrals) within 3 standard
 *                      errors; incremental energy == full recompute
 *   orc_observables    central finite differences in theta and x
 *   orc_det_exp        libm exp to 2 ulp over the move range
 *   orc_hist_x         hand-binned three- and four-particle cases (walls
 *                      and periodic, including the pair across the end) and
 *                      count identities (sum of rho = particles in range)
 *   orc_u32, orc_u53   end points and exact values; the trial stream
 *                      (particle pair per Philox block) recomputed from
 *                      orc_philox4x32_10 and replayed through orc_trial_with
 *   orc_e32_pair       float32 evaluations of the LJ and Yukawa pair
 *                      energies over hard core to cutoff lie within a
 *                      quarter of the bound; the FP32 box-length forms (LJ
 *                      scaling identity, Yukawa difference) within it; the
 *                      bound far below a dropped pair term
 *   orc_sweeps_follow  replaying the oracle's own decisions records no flip
 *                      and ends in the same state; a flipped displacement
 *                      or box-length decision planted in the log is recorded
 *                      with a margin beyond its bound
 *   orc_virials_x      three LJ particles (all pairs inside the cutoff) by
 *                      the closed forms 48/d^12 - 24/d^6, 576/d^12 - 144/d^6
 *   orc_sweeps_virial  the same samples and state as orc_sweeps, and after
 *                      every sweep the virials orc_virials_x gives
 *   schedule           the engine's (B, M_min, w) for LJ at two exclusion
 *                      distances, the hard rod + square well and Yukawa,
 *                      with rc / g_min on both sides of an integer; the
 *                      window holds every interacting partner: single
 *                      trials with steps up to four times max_step against
 *                      brute-force energies, windowed against all-partner
 *                      sweeps
 */
#ifndef JMM_ORACLE_H
#define JMM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_POT_LJ = 0, ORC_POT_HARMONIC = 1, ORC_POT_HRSW = 2, ORC_POT_YUKAWA = 3 };
enum { ORC_WALLS = 0, ORC_PERIODIC = 1 };
enum { ORC_NLT = 0, ORC_NPT = 1 };
enum { ORC_VOL_LINEAR = 0, ORC_VOL_LOG = 1 };

/* One chain's configuration (the hot-path fields of the reference's MCInput,
 * src/jmmMCState.h:66-101, plus the engine's additive fields). */
typedef struct orc_spec {
    int32_t potential, boundary, ensemble, vol_move;
    int32_t nbn;       /* -1: no neighbour-number limit (reference NBN) */
    int32_t shifted;   /* 1: subtract phi(rc) inside the cutoff */
    uint64_t N;
    double T, P, L0;   /* L0: initial box length */
    double max_step;   /* displacement half-width (reference MAXSTEP) */
    double max_dv;     /* volume half-width: dL (linear) or d ln L (log) */
    double cutoff;     /* INFINITY allowed for walls */
    double pot_p[4];
    uint64_t seed;
    uint32_t chain_id; /* Philox stream id */
    uint32_t eci;      /* full-energy recompute interval in sweeps, 0 = never */
    uint32_t bcap;     /* maximum number of blocks (a power of two; otherwise a
                          multiple of 64, and block counts are multiples of 64) */
    uint32_t pad;
} orc_spec;

typedef struct orc_sample {
    double U, L;
    uint32_t acc_d, acc_v;
} orc_sample;

typedef struct orc_chain orc_chain;

/* RNG */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_u53(uint32_t a, uint32_t b);
double orc_u32(uint32_t w);
void orc_hist_x(const orc_spec *s, const double *x, double L, uint32_t rhonb, double rbw,
                uint32_t gns, double gsw, uint32_t gnb, double gbw, uint64_t *rho, uint64_t *g);
double orc_det_exp(double x);

/* potentials: out[0]=E, out[1]=W=-d*dphi/dd, out[2]=dE/dtheta0, out[3]=dE/dtheta1,
 * out[4]=dW/dtheta0, out[5]=dW/dtheta1, out[6]=HV=-d*dW/dd, out[7]=1 if d is inside the
 * hard core (infinite energy), out[8]=1 if inside the cutoff. */
void orc_phi(const orc_spec *s, double d, double L, double out[9]);
/* The reference's own 9-component LJ output layout (src/pot.cpp:90-100). */
void orc_phi_ljcut_ref(double d, double cutoff, int virial, double out[9]);
void orc_phi_harmonic_ref(double d, double cutoff, double L, double out[2]);

/* schedule */
void orc_schedule(const orc_spec *s, uint64_t *B, uint64_t *M_min, uint64_t *w);

/* chain lifecycle */
orc_chain *orc_chain_new(const orc_spec *s);
void orc_chain_free(orc_chain *c);
int orc_chain_set_steps(orc_chain *c, double max_step, double max_dv);
void orc_chain_get(const orc_chain *c, double *x, double *L, double *U, uint64_t *sweep);
void orc_chain_set(orc_chain *c, const double *x, double L, uint64_t sweep);
uint64_t orc_chain_counter(const orc_chain *c, int which);

/* moves */
int orc_trial(orc_chain *c, uint64_t i, uint64_t sweep, double *dE_out);
int orc_trial_with(orc_chain *c, uint64_t i, double md, double u, double *dE_out);
int orc_volume_trial(orc_chain *c, uint64_t sweep, double *dU_out);
int orc_volume_trial_with(orc_chain *c, double dv, double u, double *dU_out);
int orc_accept(double dE, double beta, double u);
int orc_vol_accept(double dU, double dL, double ln_s, double P, double beta,
                   double jac, double u);

/* sweeps: writes n samples; accept_log (nullable) receives one byte per
 * displacement trial, indexed [sweep][particle], n*N bytes */
void orc_sweeps(orc_chain *c, uint64_t n, orc_sample *samples, uint8_t *accept_log);
/* as orc_sweeps, and after every sweep the virial and hypervirial of the
 * configuration into virials[2*it], virials[2*it+1] (orc_virials_x) */
void orc_sweeps_virial(orc_chain *c, uint64_t n, orc_sample *samples, double *virials);
/* the reference's cost model: every trial sums over all N-1 partners
 * (src/jmmMCState.cpp:1213-1348); same chain, same results up to rounding */
void orc_sweeps_allpairs(orc_chain *c, uint64_t n, orc_sample *samples);

/* A decision of a replayed chain that differs from the oracle's own. */
typedef struct orc_flip {
    uint64_t sweep;   /* sweep counter of the chain at the trial */
    uint64_t index;   /* particle index; UINT64_MAX for the box-length trial */
    double margin;    /* |ln u + beta dE| (displacement) or |ln u - arg| (box length),
                         from the oracle's double-precision energy change; INFINITY if
                         the oracle rejected at the walls, hard core or length bound */
    double bound;     /* single-precision error bound of that decision (same units) */
    int32_t taken;    /* the replayed form's decision, which the oracle followed */
    int32_t gated;    /* 1: the oracle's rejection was not an energy decision */
} orc_flip;

/* Single-precision error bound of one pair-energy evaluation at separation d
 * inside the cutoff (see the error model in oracle.c). */
double orc_e32_pair(const orc_spec *s, double d);

/* Replays n sweeps taking every decision from accept_log ([sweep][particle],
 * n*N bytes) and vol_acc (n bytes, NPT): records the decisions that differ
 * from the oracle's own (the first max_flips) and returns their number. */
uint64_t orc_sweeps_follow(orc_chain *c, uint64_t n, const uint8_t *accept_log, const uint8_t *vol_acc,
                           orc_sample *samples, orc_flip *flips, uint64_t max_flips);

/* energies */
double orc_full_energy(const orc_chain *c);
double orc_brute_energy(const orc_chain *c);
double orc_full_energy_x(const orc_spec *s, const double *x, double L);
double orc_brute_energy_x(const orc_spec *s, const double *x, double L);
/* U, W, dU/dtheta[2], dW/dtheta[2] into obs[6]; dU/dx into grad (nullable) */
void orc_observables_x(const orc_spec *s, const double *x, double L, double *obs, double *grad);

/* W = sum -d phi'(d), HV = sum -d dW/dd over the pairs inside the cutoff */
void orc_virials_x(const orc_spec *s, const double *x, double L, double out[2]);

/* CPU baseline: run n_sweeps on every chain, one chain per thread */
double orc_run_batch(const orc_spec *specs, int n_chains, uint64_t n_sweeps, int n_threads,
                     int allpairs, double *U_out);

#ifdef __cplusplus
}
#endif
#endif
