/*
 * oracle.c -- CPU restatement of jmmOneDMC's Metropolis chain loop (FP64).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP engine,
 * loaded by tests/, smoke() and bench.py's cpu_baseline leg, never by the
 * product.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 *
 * Floating-point contract: every operation below is a plain IEEE double
 * operation with contraction disabled, so the engine's FP64 kernels, which
 * follow the same operation sequence under -ffp-contract=off, produce the
 * same bits for positions, per-pair energies and per-trial energy changes.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ------------------------------------------------------------------ RNG */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the
 * Weyl key schedule.  Replaces the reference's GSL taus2 stream
 * (src/jmmMCState.cpp:779-781). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * c0;
        uint64_t p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in [0, 1): 27 high bits of a, 26 high bits of b. */
double orc_u53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* Uniform in (0, 1) from one word: (w + 1/2) 2^-32, exact in FP64. */
double orc_u32(uint32_t w)
{
    return ((double)w + 0.5) * (1.0 / 4294967296.0);
}

/* Deterministic exp for state updates that must match bit for bit on CPU and
 * GPU (the log-length move's scale factor).  Method: Cody-Waite reduction
 * x = k ln2 + r, |r| <= ln2/2, two-constant ln2 split, degree-13 Taylor
 * polynomial in Horner form (truncation < 5e-18 relative), exact scaling by
 * 2^k.  Only IEEE add/mul/div and floor are used. */
double orc_det_exp(double x)
{
    if (x > 709.0) return INFINITY;
    if (x < -708.0) return 0.0;
    const double inv_ln2 = 1.4426950408889634;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double kd = floor(x * inv_ln2 + 0.5);
    double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.0 / 6227020800.0;            /* 1/13! */
    p = p * r + 1.0 / 479001600.0;            /* 1/12! */
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int64_t k = (int64_t)kd;
    union { uint64_t u; double d; } sc;
    sc.u = (uint64_t)(k + 1023) << 52;
    return p * sc.d;
}

/* ----------------------------------------------------------- potentials */

/* The reference's truncated LJ with its 9-component output
 * (src/pot.cpp:19-101): E, W, E12, W12, E6, W6, HV, HV12, HV6 with an
 * inclusive cutoff (src/pot.cpp:53) and virial terms only when the flag is
 * set (src/pot.cpp:59).  Restated from the closed form, reference layout. */
void orc_phi_ljcut_ref(double d, double cutoff, int virial, double out[9])
{
    for (int k = 0; k < 9; ++k) out[k] = 0.0;
    double d3 = d * d * d;
    double i6 = 1.0 / (d3 * d3);
    double i12 = i6 * i6;
    if (d <= cutoff) {
        out[2] = 4.0 * i12;
        out[4] = 4.0 * i6;
        out[0] = out[2] - out[4];
        if (virial) {
            out[3] = 48.0 * i12;
            out[5] = 24.0 * i6;
            out[1] = out[3] - out[5];
            out[7] = 576.0 * i12;
            out[8] = 144.0 * i6;
            out[6] = out[7] - out[8];
        }
    }
}

/* The reference's harmonic bond (src/pot.cpp:110-134): 1e11 when d <= 0, a
 * strict cutoff, (d-1)^2 inside; its "virial" is (2/L) d (d-1).  Only the two
 * components it writes are produced. */
void orc_phi_harmonic_ref(double d, double cutoff, double L, double out[2])
{
    if (d <= 0.0) {
        out[0] = 10e10;
        out[1] = 10e10;
    } else if (d < cutoff) {
        double m = d - 1.0;
        out[0] = m * m;
        out[1] = (2.0 / L) * d * m;
    } else {
        out[0] = 0.0;
        out[1] = 0.0;
    }
}

/* Unshifted pair terms of the engine's potential family.  W = -d dphi/dd. */
static void phi_raw(const orc_spec *s, double d, double L, double o[9])
{
    for (int k = 0; k < 9; ++k) o[k] = 0.0;
    const double *p = s->pot_p;
    double rc = s->cutoff;
    switch (s->potential) {
    case ORC_POT_LJ: {
        /* 4 eps ((sig/d)^12 - (sig/d)^6); src/pot.cpp:47-56 with eps, sig */
        if (d < 0.5 * p[1]) o[7] = 1.0;
        if (!(d <= rc)) return;
        double sig3 = p[1] * p[1] * p[1];
        double sig6 = sig3 * sig3;
        double d3 = d * d * d;
        double i6 = (1.0 / (d3 * d3)) * sig6;
        double i12 = i6 * i6;
        double e4 = 4.0 * p[0];
        double a12 = e4 * i12, a6 = e4 * i6;
        o[0] = a12 - a6;
        o[1] = 12.0 * a12 - 6.0 * a6;
        o[6] = 144.0 * a12 - 36.0 * a6;
        o[2] = 4.0 * i12 - 4.0 * i6;           /* dE/deps */
        o[3] = o[1] / p[1];                    /* dE/dsig */
        o[4] = 12.0 * (4.0 * i12) - 6.0 * (4.0 * i6); /* dW/deps */
        o[5] = o[6] / p[1];                    /* dW/dsig */
        o[8] = 1.0;
        return;
    }
    case ORC_POT_HARMONIC:
        if (d <= 0.0) { o[0] = 10e10; o[1] = 10e10; o[7] = 1.0; return; }
        if (d < rc) {
            double m = d - 1.0;
            o[0] = m * m;
            o[1] = -2.0 * d * m;
            o[6] = 4.0 * d * d - 2.0 * d;  /* HV = -d dW/dd */
            o[8] = 1.0;
        }
        return;
    case ORC_POT_HRSW:
        /* hard rod of length p0, well of depth p1 out to p2 (strict) */
        if (d < p[0]) { o[0] = INFINITY; o[7] = 1.0; return; }
        if (d < p[2]) { o[0] = -p[1]; o[2] = -1.0; o[8] = 1.0; }
        return;
    case ORC_POT_YUKAWA: {
        /* A exp(-kappa d)/d for hard core p2 <= d <= rc */
        if (d < p[2]) { o[0] = INFINITY; o[7] = 1.0; return; }
        if (!(d <= rc)) return;
        double e = exp(-p[1] * d);
        double ed = e / d;
        double E = p[0] * ed;
        o[0] = E;
        o[1] = E * (p[1] * d + 1.0);
        o[2] = ed;                             /* dE/dA */
        o[3] = -p[0] * e;                      /* dE/dkappa */
        o[4] = ed * (p[1] * d + 1.0);          /* dW/dA */
        o[5] = -p[1] * d * d * E;              /* dW/dkappa */
        o[6] = E * ((p[1] * d + 1.0) * (p[1] * d + 1.0) - p[1] * d); /* HV = -d dW/dd */
        o[8] = 1.0;
        return;
    }
    }
}

static double shift_terms(const orc_spec *s, double o[9])
{
    /* values at the cutoff subtracted by the shifted form */
    if (!s->shifted || !isfinite(s->cutoff)) { for (int k = 0; k < 9; ++k) o[k] = 0; return 0.0; }
    phi_raw(s, s->cutoff, 0.0, o);
    return o[0];
}

void orc_phi(const orc_spec *s, double d, double L, double out[9])
{
    phi_raw(s, d, L, out);
    if (s->shifted && out[8] != 0.0) {
        double c[9];
        shift_terms(s, c);
        out[0] -= c[0];
        out[2] -= c[2];
        out[3] -= c[3];
    }
}

/* Energy only, the hot-path form (same operations as orc_phi's o[0]). */
static inline double phi_e(const orc_spec *s, double d, double shift)
{
    const double *p = s->pot_p;
    switch (s->potential) {
    case ORC_POT_LJ: {
        if (!(d <= s->cutoff)) return 0.0;
        double sig3 = p[1] * p[1] * p[1];
        double sig6 = sig3 * sig3;
        double d3 = d * d * d;
        double i6 = (1.0 / (d3 * d3)) * sig6;
        double i12 = i6 * i6;
        double e4 = 4.0 * p[0];
        return (e4 * i12 - e4 * i6) - shift;
    }
    case ORC_POT_HARMONIC:
        if (d <= 0.0) return 10e10;
        if (d < s->cutoff) { double m = d - 1.0; return m * m; }
        return 0.0;
    case ORC_POT_HRSW:
        if (d < p[0]) return INFINITY;
        return d < p[2] ? -p[1] : 0.0;
    case ORC_POT_YUKAWA:
        if (d < p[2]) return INFINITY;
        if (!(d <= s->cutoff)) return 0.0;
        return (p[0] * (exp(-p[1] * d) / d)) - shift;
    }
    return 0.0;
}

/* The cutoff test of phi_e: LJ and Yukawa d <= rc, harmonic d < rc, the
 * square well d < its edge.  The pair-term counter adds one for each of a
 * partner's old and new separations inside it (n_old + n_new). */
static inline uint64_t n_inside(const orc_spec *s, double dold, double dnew)
{
    double edge = s->potential == ORC_POT_HRSW ? s->pot_p[2] : s->cutoff;
    int strict = s->potential == ORC_POT_HRSW || s->potential == ORC_POT_HARMONIC;
    uint64_t n = 0;
    n += strict ? (dold < edge) : (dold <= edge);
    n += strict ? (dnew < edge) : (dnew <= edge);
    return n;
}

/* ------------------------------------------------------------- schedule */

static double gmin_of(const orc_spec *s)
{
    switch (s->potential) {
    case ORC_POT_LJ: return s->pot_p[2] > 0.0 ? s->pot_p[2] : 0.5 * s->pot_p[1];
    case ORC_POT_HRSW: return s->pot_p[0];
    case ORC_POT_YUKAWA: return s->pot_p[2];
    default: return DBL_TRUE_MIN; /* harmonic: reject gap <= 0 */
    }
}

/* B blocks (a power of two, or with a bcap that is not a power of two a
 * multiple of 64 when at least 64), M_min = minimum index separation of
 * trials that may run at once, w = largest index distance of any partner a
 * trial can interact with.  Every gap between neighbours stays at least
 * g_min (the hard-core test rejects a displacement or length move that
 * would break it), so partner k is at least k g_min from the old position
 * and from any new position that passes that test: w = floor(rc / g_min)
 * with a relative guard of 1e-3 (partner w + 1 lies beyond rc by at least
 * 1e-3 g_min), M_min = w + 1 (at least 2). */
void orc_schedule(const orc_spec *s, uint64_t *B, uint64_t *M_min, uint64_t *w)
{
    uint64_t N = s->N;
    uint64_t big = N > 0 ? N : 1;
    uint64_t wc = big, mc = big + 1;
    double g = gmin_of(s);
    if (isfinite(s->cutoff) && g > 1e-300) {
        double a = floor(s->cutoff / g + 1e-3);
        if (a < (double)big) wc = (uint64_t)a;
        if (a + 1.0 < (double)(big + 1)) mc = (uint64_t)a + 1;
    }
    if (s->nbn >= 0) {
        if ((uint64_t)s->nbn < wc) wc = (uint64_t)s->nbn;
        if ((uint64_t)s->nbn + 1 < mc) mc = (uint64_t)s->nbn + 1;
    }
    if (mc < 2) mc = 2; /* the hard-core test reads both nearest partners */
    if (N >= 1 && wc > N - 1) wc = N - 1;
    uint64_t nb = 1;
    if (mc >= 1 && mc <= N) {
        uint64_t q = N / mc;
        uint64_t cap = s->bcap ? s->bcap : 256;
        if (q > cap) q = cap;
        if ((cap & (cap - 1)) != 0 && q >= 64) nb = q / 64 * 64; /* whole waves */
        else while (nb * 2 <= q) nb *= 2;
    }
    *B = nb;
    *M_min = mc;
    *w = wc;
}

/* ---------------------------------------------------------------- chain */

struct orc_chain {
    orc_spec s;
    double *x;
    double L, U, beta, shift, gmin;
    double step_cap; /* max_step at creation: the periodic length bound,
                        L > 2 (rc + max_step), is built for it */
    uint64_t sweep, B, M_min, w;
    uint64_t cnt[8]; /* trials, acc_d, wall_rej, hc_rej, pair_terms, vol_trials, acc_v, recomputes */
    double max_drift;
};

enum { C_TRIALS, C_ACC_D, C_WALL, C_HC, C_TERMS, C_VTRIALS, C_ACC_V, C_RECOMP };

static inline int nb_pos(const orc_chain *c, int64_t j, double *xj)
{
    int64_t N = (int64_t)c->s.N;
    if (j >= 0 && j < N) { *xj = c->x[j]; return 1; }
    if (c->s.boundary == ORC_WALLS) return 0;
    if (j < 0) *xj = c->x[j + N] - c->L;
    else *xj = c->x[j - N] + c->L;
    return 1;
}

orc_chain *orc_chain_new(const orc_spec *s)
{
    orc_chain *c = (orc_chain *)calloc(1, sizeof(orc_chain));
    c->s = *s;
    c->x = (double *)malloc(sizeof(double) * (s->N ? s->N : 1));
    c->L = s->L0;
    c->beta = 1.0 / s->T;
    double sh[9];
    c->shift = shift_terms(s, sh);
    c->gmin = gmin_of(s);
    c->step_cap = s->max_step;
    orc_schedule(s, &c->B, &c->M_min, &c->w);
    /* lattice start, src/jmmMCState.cpp:559-562 */
    for (uint64_t i = 0; i < s->N; ++i)
        c->x[i] = ((i + 0.5) / s->N - 0.5) * c->L;
    c->U = orc_full_energy(c);
    return c;
}

/* New step sizes between sweeps (the reference's maxDisAdjust/maxDVAdjust,
 * src/jmmMCState.cpp:2100-2139, change maxStep/maxdl in place); the block
 * schedule stays the one built at creation, so max_step may not exceed the
 * creation value.  Returns 0 on success, -1 if it does. */
int orc_chain_set_steps(orc_chain *c, double max_step, double max_dv)
{
    if (!(max_step >= 0.0) || max_step > c->step_cap || !(max_dv >= 0.0)) return -1;
    c->s.max_step = max_step;
    c->s.max_dv = max_dv;
    return 0;
}

void orc_chain_free(orc_chain *c)
{
    if (!c) return;
    free(c->x);
    free(c);
}

void orc_chain_get(const orc_chain *c, double *x, double *L, double *U, uint64_t *sweep)
{
    if (x) memcpy(x, c->x, sizeof(double) * c->s.N);
    if (L) *L = c->L;
    if (U) *U = c->U;
    if (sweep) *sweep = c->sweep;
}

void orc_chain_set(orc_chain *c, const double *x, double L, uint64_t sweep)
{
    memcpy(c->x, x, sizeof(double) * c->s.N);
    c->L = L;
    c->sweep = sweep;
    c->U = orc_full_energy(c);
}

uint64_t orc_chain_counter(const orc_chain *c, int which)
{
    if (which == 100) return c->B;
    if (which == 101) return c->M_min;
    if (which == 102) return c->w;
    if (which < 0 || which >= 8) return 0;
    return c->cnt[which];
}

/* ---------------------------------------------------------------- moves */

/* Metropolis predicate of qad2 (src/jmmMCState.cpp:1367-1377): accept when
 * dE <= 0 or when the Boltzmann factor strictly exceeds the uniform.  The
 * factor is exp(-dE*beta) with beta = 1/T precomputed (the reference divides
 * by T); the two differ only in the last bit, i.e. only at exact ties. */
int orc_accept(double dE, double beta, double u)
{
    if (dE <= 0.0) return 1;
    return exp(-dE * beta) > u;
}

/* Acceptance of a box-length trial (src/jmmMCState.cpp:2249-2255, :1665-1672):
 * bf = exp(-beta (dU + P dL) + jac ln s); accept when bf >= 1 or bf > u. */
int orc_vol_accept(double dU, double dL, double ln_s, double P, double beta,
                   double jac, double u)
{
    double bf = exp(-beta * (dU + P * dL) + jac * ln_s);
    return bf >= 1.0 || bf > u;
}

/* --------------------------------------------- single-precision error model
 *
 * The FP32 form evaluates pair energies in single precision from separations
 * formed and tested against the cutoff in double precision, and sums them
 * in double precision.  Its decision can differ from the double-precision one
 * only where the margin of the decision lies within the error of its energy
 * change.  Bound of one single-precision pair evaluation at a separation d
 * inside the cutoff, u = 2^-24:
 *     e32(d) = 64 u (|phi(d)| + |W(d)| + |shift|),
 * the |W| = |d phi'(d)| term for the rounding of d itself, the |phi| term for
 * the evaluation (about ten roundings, the hardware reciprocal and
 * exponential within two ulp each), 64 a safety factor over both; the hard
 * rod + square well is exact (well counts).  A displacement trial's bound on
 * |ln u + beta dE| is beta (sum of e32 over the old and new separations
 * inside the cutoff + 8 u sum |dphi|) + 8 u (1 + beta |dE|), the last term for
 * the single-precision exponential of the Metropolis test; a box-length
 * trial's is beta (sum of e32 over the pairs' unscaled and scaled
 * separations + 8 u sum |dphi|) + 1e-12 (1 + |arg|), its acceptance being
 * evaluated in double precision.  Pinned in tests/test_oracle.py against
 * float32 evaluations of every potential. */

#define ORC_U32 (1.0 / 16777216.0)

static inline int inside_of(const orc_spec *s, double d)
{
    if (s->potential == ORC_POT_HRSW) return d < s->pot_p[2];
    if (s->potential == ORC_POT_HARMONIC) return d < s->cutoff;
    return d <= s->cutoff;
}

double orc_e32_pair(const orc_spec *s, double d)
{
    if (s->potential == ORC_POT_HRSW) return 0.0;
    double o[9], c[9];
    phi_raw(s, d, 0.0, o);
    if (o[7] != 0.0) return INFINITY;
    const double shift = shift_terms(s, c);
    return 64.0 * ORC_U32 * (fabs(o[0]) + fabs(o[1]) + fabs(shift));
}

/* The bound on |ln u + beta dE| of displacement trial i by md (state before
 * the move), walking the partners as trial_core does. */
static double trial_bound32(const orc_chain *c, uint64_t i, double md, double dE)
{
    const orc_spec *s = &c->s;
    const double xi = c->x[i], xn = xi + md, rc = s->cutoff;
    double E = 0.0, T = 0.0;
    for (int side = 0; side < 2; ++side) {
        for (uint64_t k = 1; k <= c->w; ++k) {
            double xj;
            if (!nb_pos(c, side ? (int64_t)i + (int64_t)k : (int64_t)i - (int64_t)k, &xj)) break;
            double dold = side ? xj - xi : xi - xj, dnew = side ? xj - xn : xn - xj;
            if (dold > rc && dnew > rc) break;
            if (inside_of(s, dold)) E += orc_e32_pair(s, dold);
            if (inside_of(s, dnew)) E += orc_e32_pair(s, dnew);
            T += fabs(phi_e(s, dnew, c->shift) - phi_e(s, dold, c->shift));
        }
    }
    return c->beta * (E + 8.0 * ORC_U32 * T) + 8.0 * ORC_U32 * (1.0 + c->beta * fabs(dE));
}

/* What a trial decided and how close it was (see orc_flip). */
typedef struct decision {
    int own;        /* the oracle's decision */
    int gated;      /* ended at the walls, the hard core or an inadmissible length */
    double margin;  /* distance of the decision from its threshold, ln units */
    double bound;   /* single-precision error bound of the same (only where computed) */
} decision;

/* One displacement trial of particle i by md with acceptance uniform u.
 * Restates qad2 (src/jmmMCState.cpp:1160-1464): wall rejection before any
 * energy work (:1188-1193), energy change over the partners (:1210-1348),
 * Metropolis test (:1367-1377), running totals on acceptance (:1384-1394).
 * Partners are read from the sorted neighbour window instead of the O(N^2)
 * pair tables; pairs beyond the cutoff or the NBN limit contribute zero in
 * both, so the sums agree up to the order of addition.
 *
 * force < 0: the oracle's own decision.  force = 0 or 1: that decision is
 * taken instead (replaying another form's chain); *dec (nullable) receives
 * the oracle's own decision, its margin |ln u + beta dE|, and, where the
 * forced decision differs from it, the single-precision error bound. */
static int trial_core(orc_chain *c, uint64_t i, double md, double u, int force, double *dE_out,
                      decision *dec)
{
    const orc_spec *s = &c->s;
    double xi = c->x[i];
    double xn = xi + md;
    double rc = s->cutoff;
    c->cnt[C_TRIALS]++;
    if (dE_out) *dE_out = 0.0;
    if (dec) { dec->own = 0; dec->gated = 1; dec->margin = INFINITY; dec->bound = 0.0; }
    if (s->boundary == ORC_WALLS && fabs(xn) > c->L / 2.0) { c->cnt[C_WALL]++; return 0; }
    if (s->N >= 2) {
        double xl, xr;
        if (nb_pos(c, (int64_t)i - 1, &xl) && (xn - xl) < c->gmin) { c->cnt[C_HC]++; return 0; }
        if (nb_pos(c, (int64_t)i + 1, &xr) && (xr - xn) < c->gmin) { c->cnt[C_HC]++; return 0; }
    }
    /* left and right partner sums are kept apart and added at the end */
    double dEL = 0.0, dER = 0.0;
    uint64_t terms = 0;
    for (uint64_t k = 1; k <= c->w; ++k) {
        double xj;
        if (!nb_pos(c, (int64_t)i - (int64_t)k, &xj)) break;
        double dold = xi - xj, dnew = xn - xj;
        if (dold > rc && dnew > rc) break;
        dEL += phi_e(s, dnew, c->shift) - phi_e(s, dold, c->shift);
        terms += n_inside(s, dold, dnew);
    }
    for (uint64_t k = 1; k <= c->w; ++k) {
        double xj;
        if (!nb_pos(c, (int64_t)i + (int64_t)k, &xj)) break;
        double dold = xj - xi, dnew = xj - xn;
        if (dold > rc && dnew > rc) break;
        dER += phi_e(s, dnew, c->shift) - phi_e(s, dold, c->shift);
        terms += n_inside(s, dold, dnew);
    }
    double dE = dEL + dER;
    c->cnt[C_TERMS] += terms;
    if (dE_out) *dE_out = dE;
    const int mine = orc_accept(dE, c->beta, u);
    if (dec) {
        dec->own = mine;
        dec->gated = 0;
        dec->margin = fabs(log(u) + c->beta * dE);
        if (force >= 0 && force != mine) dec->bound = trial_bound32(c, i, md, dE);
    }
    if (force < 0 ? mine : force) {
        c->x[i] = xn;
        c->U += dE;
        c->cnt[C_ACC_D]++;
        return 1;
    }
    return 0;
}

int orc_trial_with(orc_chain *c, uint64_t i, double md, double u, double *dE_out)
{
    return trial_core(c, i, md, u, -1, dE_out, NULL);
}

/* Position t of particle i in its block (blocks of q+1 particles first,
 * then blocks of q; q = N/B, N%B blocks of q+1). */
static uint64_t block_pos(uint64_t N, uint64_t B, uint64_t i)
{
    uint64_t q = N / B, r = N % B, big = r * (q + 1);
    return i < big ? i % (q + 1) : (i - big) % q;
}

/* The uniforms of displacement trial i (block position t) in a sweep: one
 * Philox4x32-10 block per pair of consecutive particles of a block,
 * counter (first particle of the pair, sweep, chain_id, 0), key = seed; the
 * pair's first particle (t even) takes words 0 (displacement) and 1
 * (acceptance), the second words 2 and 3. */
static inline void draw_trial(const orc_chain *c, uint64_t i, uint64_t t, uint64_t sweep,
                              double *a, double *b)
{
    uint32_t ctr[4] = { (uint32_t)(i - (t & 1u)), (uint32_t)sweep, c->s.chain_id, 0u };
    uint32_t key[2] = { (uint32_t)c->s.seed, (uint32_t)(c->s.seed >> 32) };
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    const int p = (int)(t & 1u) * 2;
    *a = orc_u32(r[p]);
    *b = orc_u32(r[p + 1]);
}

/* The uniforms of the box-length trial: counter (0xFFFFFFFF, sweep,
 * chain_id, 0), 53-bit uniforms from words (0,1) and (2,3). */
static inline void draw_volume(const orc_chain *c, uint64_t sweep, double *a, double *b)
{
    uint32_t ctr[4] = { 0xFFFFFFFFu, (uint32_t)sweep, c->s.chain_id, 0u };
    uint32_t key[2] = { (uint32_t)c->s.seed, (uint32_t)(c->s.seed >> 32) };
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    *a = orc_u53(r[0], r[1]);
    *b = orc_u53(r[2], r[3]);
}

int orc_trial(orc_chain *c, uint64_t i, uint64_t sweep, double *dE_out)
{
    double a, u;
    draw_trial(c, i, block_pos(c->s.N, c->B, i), sweep, &a, &u);
    double md = ((a - 0.5) * 2.0) * c->s.max_step; /* src/jmmMCState.cpp:1182 */
    return orc_trial_with(c, i, md, u, dE_out);
}

/* Energy of the scaled configuration minus the current one, pair by pair.
 * Returns 0 when a scaled pair falls inside the hard core. */
static int scaled_dU(const orc_chain *c, double sc, double Lp, double *dU)
{
    const orc_spec *s = &c->s;
    int64_t N = (int64_t)s->N;
    double acc = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        double xi = c->x[i], xis = xi * sc;
        for (uint64_t k = 1; k <= c->w; ++k) {
            int64_t j = i + (int64_t)k;
            double xj, xjs;
            if (j < N) { xj = c->x[j]; xjs = xj * sc; }
            else if (s->boundary == ORC_PERIODIC) { xj = c->x[j - N] + c->L; xjs = c->x[j - N] * sc + Lp; }
            else break;
            double d0 = xj - xi, d1 = xjs - xis;
            if (k == 1 && d1 < c->gmin) return 0;
            if (d0 > s->cutoff && d1 > s->cutoff) break;
            acc += phi_e(s, d1, c->shift) - phi_e(s, d0, c->shift);
        }
    }
    *dU = acc;
    return 1;
}

/* The proposal of a box-length trial: scale factor, new length, ln s, the
 * Jacobian factor and the length change; returns 0 when the new length is
 * not admissible (non-positive, or a periodic length at or below
 * 2 (cutoff + max_step)). */
static int vol_proposal(const orc_chain *c, double dv, double *sc_o, double *Lp_o, double *lns_o,
                        double *jac_o, double *dL_o)
{
    const orc_spec *s = &c->s;
    double sc, Lp, lns, jac, dL;
    if (s->vol_move == ORC_VOL_LINEAR) {
        Lp = c->L + dv;
        sc = Lp / c->L;
        lns = log(sc);
        jac = (double)s->N;
        dL = dv;
    } else {
        sc = orc_det_exp(dv);
        Lp = c->L * sc;
        lns = dv;
        jac = (double)s->N + 1.0;
        dL = Lp - c->L;
    }
    *sc_o = sc; *Lp_o = Lp; *lns_o = lns; *jac_o = jac; *dL_o = dL;
    if (!(Lp > 0.0)) return 0;
    if (s->boundary == ORC_PERIODIC && !(Lp > 2.0 * (s->cutoff + c->step_cap))) return 0;
    return 1;
}

/* The bound on |ln u - arg| of a box-length trial scaling by sc to Lp
 * (state before the move), over the pairs scaled_dU walks. */
static double vol_bound32(const orc_chain *c, double sc, double Lp, double arg)
{
    const orc_spec *s = &c->s;
    int64_t N = (int64_t)s->N;
    double E = 0.0, T = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        double xi = c->x[i], xis = xi * sc;
        for (uint64_t k = 1; k <= c->w; ++k) {
            int64_t j = i + (int64_t)k;
            double xj, xjs;
            if (j < N) { xj = c->x[j]; xjs = xj * sc; }
            else if (s->boundary == ORC_PERIODIC) { xj = c->x[j - N] + c->L; xjs = c->x[j - N] * sc + Lp; }
            else break;
            double d0 = xj - xi, d1 = xjs - xis;
            if (d0 > s->cutoff && d1 > s->cutoff) break;
            if (inside_of(s, d0)) E += orc_e32_pair(s, d0);
            if (inside_of(s, d1)) E += orc_e32_pair(s, d1);
            T += fabs(phi_e(s, d1, c->shift) - phi_e(s, d0, c->shift));
        }
    }
    return c->beta * (E + 8.0 * ORC_U32 * T) + 1e-12 * (1.0 + fabs(arg));
}

/* Box-length trial.  Restates fav (src/jmmMCState.cpp:2161-2293): scale all
 * coordinates by s = L'/L (:2175-2180), full energy at the scaled positions
 * (:2193-2235), acceptance with the N ln s Jacobian (:2249-2255).  The
 * log-length form draws ln s uniformly and uses an (N+1) ln s Jacobian.
 * force < 0: the oracle's own decision; force 0/1: that decision is taken
 * instead; *dec as in trial_core, the margin |ln u - arg| with
 * arg = -beta (dU + P dL) + jac ln s (gated: the new length is not
 * admissible or a scaled nearest pair overlaps). */
static int vol_core(orc_chain *c, double dv, double u, int force, double *dU_out, decision *dec)
{
    const orc_spec *s = &c->s;
    double sc, Lp, lns, jac, dL;
    c->cnt[C_VTRIALS]++;
    if (dU_out) *dU_out = 0.0;
    if (dec) { dec->own = 0; dec->gated = 1; dec->margin = INFINITY; dec->bound = 0.0; }
    if (!vol_proposal(c, dv, &sc, &Lp, &lns, &jac, &dL)) return 0;
    double dU;
    if (!scaled_dU(c, sc, Lp, &dU)) return 0;
    if (dU_out) *dU_out = dU;
    const int mine = orc_vol_accept(dU, dL, lns, s->P, c->beta, jac, u);
    if (dec) {
        const double arg = -c->beta * (dU + s->P * dL) + jac * lns;
        dec->own = mine;
        dec->gated = 0;
        dec->margin = fabs(arg - log(u));
        if (force >= 0 && force != mine) dec->bound = vol_bound32(c, sc, Lp, arg);
    }
    if (!(force < 0 ? mine : force)) return 0;
    for (uint64_t i = 0; i < s->N; ++i) c->x[i] = c->x[i] * sc;
    c->L = Lp;
    c->U += dU;
    c->cnt[C_ACC_V]++;
    return 1;
}

int orc_volume_trial_with(orc_chain *c, double dv, double u, double *dU_out)
{
    return vol_core(c, dv, u, -1, dU_out, NULL);
}

int orc_volume_trial(orc_chain *c, uint64_t sweep, double *dU_out)
{
    double a, u;
    draw_volume(c, sweep, &a, &u);
    double dv = ((a - 0.5) * 2.0) * c->s.max_dv; /* src/jmmMCState.cpp:2172 */
    return orc_volume_trial_with(c, dv, u, dU_out);
}

/* --------------------------------------------------------------- energy */

/* Full energy from the sorted windows: every pair once, from its left
 * member.  Plays the role of fad's initial energy (src/jmmMCState.cpp:
 * 903-946) and ECheck's recompute (src/jmmMCState.cpp:1973-1993). */
double orc_full_energy_x(const orc_spec *s, const double *x, double L)
{
    uint64_t B, M, w;
    orc_schedule(s, &B, &M, &w);
    double sh[9];
    double shift = shift_terms(s, sh);
    int64_t N = (int64_t)s->N;
    double U = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        for (uint64_t k = 1; k <= w; ++k) {
            int64_t j = i + (int64_t)k;
            double xj;
            if (j < N) xj = x[j];
            else if (s->boundary == ORC_PERIODIC) xj = x[j - N] + L;
            else break;
            double d = xj - x[i];
            if (d > s->cutoff) break;
            U += phi_e(s, d, shift);
        }
    }
    return U;
}

double orc_full_energy(const orc_chain *c)
{
    return orc_full_energy_x(&c->s, c->x, c->L);
}

/* All N(N-1)/2 pairs with no window logic: walls use r_j - r_i
 * (src/jmmMCState.cpp:767-769), periodic adds the wrapped image; the NBN
 * limit is the index distance (src/jmmMCState.cpp:1979). */
double orc_brute_energy_x(const orc_spec *s, const double *x, double L)
{
    double sh[9];
    double shift = shift_terms(s, sh);
    int64_t N = (int64_t)s->N;
    double U = 0.0;
    for (int64_t i = 0; i < N; ++i)
        for (int64_t j = i + 1; j < N; ++j) {
            int64_t k = j - i;
            if (s->nbn < 0 || k <= s->nbn) U += phi_e(s, x[j] - x[i], shift);
            if (s->boundary == ORC_PERIODIC) {
                int64_t k2 = N - k;
                if (s->nbn < 0 || k2 <= s->nbn) U += phi_e(s, x[i] + L - x[j], shift);
            }
        }
    return U;
}

double orc_brute_energy(const orc_chain *c)
{
    return orc_brute_energy_x(&c->s, c->x, c->L);
}

/* U, W, dU/dtheta, dW/dtheta and dU/dx for one configuration, FP64. */
void orc_observables_x(const orc_spec *s, const double *x, double L, double *obs, double *grad)
{
    uint64_t B, M, w;
    orc_schedule(s, &B, &M, &w);
    int64_t N = (int64_t)s->N;
    for (int k = 0; k < 6; ++k) obs[k] = 0.0;
    if (grad) for (int64_t i = 0; i < N; ++i) grad[i] = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        for (uint64_t k = 1; k <= w; ++k) {
            int64_t j = i + (int64_t)k, jj = j;
            double xj;
            if (j < N) xj = x[j];
            else if (s->boundary == ORC_PERIODIC) { jj = j - N; xj = x[jj] + L; }
            else break;
            double d = xj - x[i];
            if (d > s->cutoff) break;
            double o[9];
            orc_phi(s, d, L, o);
            obs[0] += o[0];
            obs[1] += o[1];
            obs[2] += o[2];
            obs[3] += o[3];
            obs[4] += o[4];
            obs[5] += o[5];
            if (grad) {
                double f = o[1] / d; /* -dphi/dd */
                grad[i] += f;        /* d = x_j - x_i */
                grad[jj] -= f;
            }
        }
    }
}

/* Virial W = sum -d phi'(d) and hypervirial HV = sum -d W'(d) over the
 * pairs inside the cutoff (the reference's Vir and HV totals,
 * src/jmmMCState.cpp:86-96, accumulated by updateThermo :1941-1961). */
void orc_virials_x(const orc_spec *s, const double *x, double L, double out[2])
{
    uint64_t B, M, w;
    orc_schedule(s, &B, &M, &w);
    int64_t N = (int64_t)s->N;
    out[0] = out[1] = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        for (uint64_t k = 1; k <= w; ++k) {
            int64_t j = i + (int64_t)k;
            double xj;
            if (j < N) xj = x[j];
            else if (s->boundary == ORC_PERIODIC) xj = x[j - N] + L;
            else break;
            double d = xj - x[i];
            if (d > s->cutoff) break;
            if (!inside_of(s, d)) continue;
            double o[9];
            phi_raw(s, d, L, o);
            out[0] += o[1];
            out[1] += o[6];
        }
    }
}

/* ----------------------------------------------------------- histograms */

static double fold_x(double x, double L, int per)
{
    return per ? x - L * floor((x + L / 2.0) / L) : x;
}

/* One sample of the density and pair-separation histograms, all pairs
 * (fgrho, src/jmmMCState.cpp:1069-1128): rho bin floor(x/rbw + rhonb/2);
 * pair bin floor(d/gbw) for d < gnb*gbw counted in the segment
 * floor(x/gsw + gns/2) of each member.  Periodic: positions folded into
 * [-L/2, L/2) and each pair at its separation below L/2, the separation
 * across the end formed as (x_j + L) - x_i.  Counts are added to rho/g. */
void orc_hist_x(const orc_spec *s, const double *x, double L, uint32_t rhonb, double rbw,
                uint32_t gns, double gsw, uint32_t gnb, double gbw, uint64_t *rho, uint64_t *g)
{
    const int per = s->boundary == ORC_PERIODIC;
    const int64_t N = (int64_t)s->N;
    const double gmax = gnb * gbw;
    for (int64_t i = 0; i < N; ++i) {
        const double fi = fold_x(x[i], L, per);
        const int64_t rb = (int64_t)floor(fi / rbw + rhonb / 2.0);
        if (rb >= 0 && rb < (int64_t)rhonb) rho[rb]++;
        if (!gns || !gnb) continue;
        const int64_t si = (int64_t)floor(fi / gsw + gns / 2.0);
        for (int64_t j = 0; j < N; ++j) {
            if (j == i) continue;
            double d;
            if (j > i) d = x[j] - x[i];
            else if (per) d = (x[j] + L) - x[i];
            else continue;
            if (!(d < gmax) || (per && !(d < L / 2.0))) continue;
            const int64_t gb = (int64_t)floor(d / gbw);
            if (gb < 0 || gb >= (int64_t)gnb) continue;
            const int64_t sj = (int64_t)floor(fold_x(x[j], L, per) / gsw + gns / 2.0);
            if (si >= 0 && si < (int64_t)gns) g[si * gnb + gb]++;
            if (sj >= 0 && sj < (int64_t)gns) g[sj * gnb + gb]++;
        }
    }
}

/* --------------------------------------------------------------- sweeps */

static void block_of(uint64_t N, uint64_t B, uint64_t b, uint64_t *start, uint64_t *size)
{
    uint64_t q = N / B, r = N % B;
    *size = q + (b < r ? 1 : 0);
    *start = b * q + (b < r ? b : r);
}

static void sweep_common(orc_chain *c, uint64_t n, orc_sample *samples, uint8_t *log, int allpairs,
                         double *virials);

void orc_sweeps(orc_chain *c, uint64_t n, orc_sample *samples, uint8_t *accept_log)
{
    sweep_common(c, n, samples, accept_log, 0, NULL);
}

void orc_sweeps_virial(orc_chain *c, uint64_t n, orc_sample *samples, double *virials)
{
    sweep_common(c, n, samples, NULL, 0, virials);
}

/* All-partner energy change: the reference's O(N) loop over every other
 * particle per trial (src/jmmMCState.cpp:1213-1348). */
static int trial_allpairs(orc_chain *c, uint64_t i, double md, double u)
{
    const orc_spec *s = &c->s;
    int64_t N = (int64_t)s->N;
    double xi = c->x[i], xn = xi + md;
    c->cnt[C_TRIALS]++;
    if (s->boundary == ORC_WALLS && fabs(xn) > c->L / 2.0) { c->cnt[C_WALL]++; return 0; }
    if (N >= 2) {
        double xl, xr;
        if (nb_pos(c, (int64_t)i - 1, &xl) && (xn - xl) < c->gmin) { c->cnt[C_HC]++; return 0; }
        if (nb_pos(c, (int64_t)i + 1, &xr) && (xr - xn) < c->gmin) { c->cnt[C_HC]++; return 0; }
    }
    double dE = 0.0;
    for (int64_t j = 0; j < N; ++j) {
        if (j == (int64_t)i) continue;
        int64_t kr = j > (int64_t)i ? j - (int64_t)i : j - (int64_t)i + N; /* right index distance */
        double xjr = j > (int64_t)i ? c->x[j] : c->x[j] + c->L;
        if (s->boundary == ORC_WALLS) {
            int64_t k = j > (int64_t)i ? j - (int64_t)i : (int64_t)i - j;
            if (s->nbn >= 0 && k > s->nbn) continue;
            double d0 = j > (int64_t)i ? c->x[j] - xi : xi - c->x[j];
            double d1 = j > (int64_t)i ? c->x[j] - xn : xn - c->x[j];
            dE += phi_e(s, d1, c->shift) - phi_e(s, d0, c->shift);
        } else {
            if (s->nbn < 0 || kr <= s->nbn)
                dE += phi_e(s, xjr - xn, c->shift) - phi_e(s, xjr - xi, c->shift);
            int64_t kl = N - kr;
            /* the left image: x_j itself for j < i, x_j - L for j > i (the
             * window's images; xjr - L would round differently) */
            double xjl = j > (int64_t)i ? c->x[j] - c->L : c->x[j];
            if (s->nbn < 0 || kl <= s->nbn)
                dE += phi_e(s, xn - xjl, c->shift) - phi_e(s, xi - xjl, c->shift);
        }
    }
    if (orc_accept(dE, c->beta, u)) {
        c->x[i] = xn;
        c->U += dE;
        c->cnt[C_ACC_D]++;
        return 1;
    }
    return 0;
}

static void sweep_common(orc_chain *c, uint64_t n, orc_sample *samples, uint8_t *log, int allpairs,
                         double *virials)
{
    const orc_spec *s = &c->s;
    uint64_t N = s->N, B = c->B;
    uint64_t smax = (N + B - 1) / B;
    for (uint64_t it = 0; it < n; ++it) {
        uint64_t sweep = c->sweep;
        uint64_t acc0 = c->cnt[C_ACC_D], accv0 = c->cnt[C_ACC_V];
        for (uint64_t t = 0; t < smax; ++t) {
            for (uint64_t b = 0; b < B; ++b) {
                uint64_t start, size;
                block_of(N, B, b, &start, &size);
                if (t >= size) continue;
                uint64_t i = start + t;
                double a, u;
                draw_trial(c, i, t, sweep, &a, &u);
                double md = ((a - 0.5) * 2.0) * s->max_step;
                int acc = allpairs ? trial_allpairs(c, i, md, u) : trial_core(c, i, md, u, -1, NULL, NULL);
                if (log) log[it * N + i] = (uint8_t)acc;
            }
        }
        if (s->ensemble == ORC_NPT) orc_volume_trial(c, sweep, NULL);
        if (s->eci && (sweep + 1) % s->eci == 0) {
            double Uf = orc_full_energy(c);
            double dr = fabs(Uf - c->U);
            if (dr > c->max_drift) c->max_drift = dr;
            c->U = Uf;
            c->cnt[C_RECOMP]++;
        }
        if (samples) {
            samples[it].U = c->U;
            samples[it].L = c->L;
            samples[it].acc_d = (uint32_t)(c->cnt[C_ACC_D] - acc0);
            samples[it].acc_v = (uint32_t)(c->cnt[C_ACC_V] - accv0);
        }
        if (virials) orc_virials_x(s, c->x, c->L, virials + 2 * it);
        c->sweep++;
    }
}

void orc_sweeps_allpairs(orc_chain *c, uint64_t n, orc_sample *samples)
{
    sweep_common(c, n, samples, NULL, 1, NULL);
}

/* Replay of another form's chain (the FP32 form's): the same sweeps as
 * orc_sweeps, taking every decision from accept_log[it*N + i] (displacement
 * trials) and vol_acc[it] (box-length trials, NPT) instead of the oracle's
 * own.  Where the two differ the oracle records an orc_flip with its margin
 * and the single-precision error bound of that trial (the first max_flips
 * of them), then continues from the replayed form's decision, so both chains
 * stay in the same state and every later trial is compared again.  Returns
 * the number of flips. */
uint64_t orc_sweeps_follow(orc_chain *c, uint64_t n, const uint8_t *accept_log, const uint8_t *vol_acc,
                           orc_sample *samples, orc_flip *flips, uint64_t max_flips)
{
    const orc_spec *s = &c->s;
    uint64_t N = s->N, B = c->B;
    uint64_t smax = (N + B - 1) / B;
    uint64_t nflip = 0;
    for (uint64_t it = 0; it < n; ++it) {
        uint64_t sweep = c->sweep;
        uint64_t acc0 = c->cnt[C_ACC_D], accv0 = c->cnt[C_ACC_V];
        for (uint64_t t = 0; t < smax; ++t) {
            for (uint64_t b = 0; b < B; ++b) {
                uint64_t start, size;
                block_of(N, B, b, &start, &size);
                if (t >= size) continue;
                uint64_t i = start + t;
                double a, u;
                draw_trial(c, i, t, sweep, &a, &u);
                double md = ((a - 0.5) * 2.0) * s->max_step;
                const int want = accept_log[it * N + i] ? 1 : 0;
                decision d;
                trial_core(c, i, md, u, want, NULL, &d);
                if (d.gated ? want : d.own != want) {
                    if (nflip < max_flips) {
                        orc_flip *f = &flips[nflip];
                        f->sweep = sweep;
                        f->index = i;
                        f->margin = d.gated ? INFINITY : d.margin;
                        f->bound = d.bound;
                        f->taken = want;
                        f->gated = d.gated;
                    }
                    ++nflip;
                }
            }
        }
        if (s->ensemble == ORC_NPT) {
            double a, u;
            draw_volume(c, sweep, &a, &u);
            double dv = ((a - 0.5) * 2.0) * s->max_dv;
            const int want = vol_acc[it] ? 1 : 0;
            decision d;
            const int took = vol_core(c, dv, u, want, NULL, &d);
            if ((d.gated ? want : d.own != want) || took != want) {
                if (nflip < max_flips) {
                    orc_flip *f = &flips[nflip];
                    f->sweep = sweep;
                    f->index = UINT64_MAX;
                    f->margin = d.gated ? INFINITY : d.margin;
                    f->bound = d.bound;
                    f->taken = want;
                    f->gated = d.gated;
                }
                ++nflip;
            }
        }
        if (s->eci && (sweep + 1) % s->eci == 0) {
            double Uf = orc_full_energy(c);
            double dr = fabs(Uf - c->U);
            if (dr > c->max_drift) c->max_drift = dr;
            c->U = Uf;
            c->cnt[C_RECOMP]++;
        }
        if (samples) {
            samples[it].U = c->U;
            samples[it].L = c->L;
            samples[it].acc_d = (uint32_t)(c->cnt[C_ACC_D] - acc0);
            samples[it].acc_v = (uint32_t)(c->cnt[C_ACC_V] - accv0);
        }
        c->sweep++;
    }
    return nflip;
}

/* ------------------------------------------------------------- baseline */

double orc_run_batch(const orc_spec *specs, int n_chains, uint64_t n_sweeps, int n_threads,
                     int allpairs, double *U_out)
{
    if (n_threads < 1) n_threads = 1;
    orc_chain **cs = (orc_chain **)calloc((size_t)n_chains, sizeof(orc_chain *));
    for (int k = 0; k < n_chains; ++k) cs[k] = orc_chain_new(&specs[k]);
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads)
    for (int k = 0; k < n_chains; ++k) sweep_common(cs[k], n_sweeps, NULL, NULL, allpairs, NULL);
    double dt = omp_get_wtime() - t0;
    for (int k = 0; k < n_chains; ++k) {
        if (U_out) U_out[k] = cs[k]->U;
        orc_chain_free(cs[k]);
    }
    free(cs);
    return dt;
}
