// kernels.hip -- gfx950 kernels of the Metropolis engine.
//
// Compiled with -ffp-contract=off: the FP64 paths round once per operation in
// the oracle's order, which makes positions and per-trial energy changes
// reproducible bit for bit (only sums over lanes differ in their order).
//
//   k_sweep_batched  persistent, work-queue driven: one workgroup takes a
//                    bundle of chains, keeps their sorted positions in LDS,
//                    one lane per block of particles, phases separated by
//                    workgroup barriers.  FP64 or FP32 pair arithmetic.
//   k_sweep_plain    one wavefront per chain; trials one at a time, the
//                    energy change a wavefront reduction over the neighbour
//                    window (lanes 0-31 left partners, 32-63 right).
//   k_observables    U, virial, dU/dtheta, dW/dtheta and dU/dx per
//                    configuration (the differentiation path).
//   k_begin          resets a work queue and advances the sample row.
#include "kernels.h"

// Occupancy target of the batched kernels (waves per SIMD); bounds VGPRs.
#ifndef JMM_WAVES_PER_EU
#define JMM_WAVES_PER_EU 2
#endif

namespace jmm {

__device__ __forceinline__ bool nbpos(const double *X, int64_t j, int64_t N, bool per, double L,
                                      double &xj)
{
    if (j >= 0 && j < N) {
        xj = X[j];
        return true;
    }
    if (!per) return false;
    xj = (j < 0) ? X[j + N] - L : X[j - N] + L;
    return true;
}

// A partner's contribution to the pair-term counter: 1 for each of the old
// and new separations inside the cutoff (the counter is n_old + n_new).
template <int KIND>
__device__ __forceinline__ uint32_t n_inside(const Pot &q, double o, double n)
{
    return (inside64<KIND>(q, o) ? 1u : 0u) + (inside64<KIND>(q, n) ? 1u : 0u);
}

// The same from the flags a walk step has: `in` (the partner is walked),
// io and inn (inside at the old / new position).  For LJ and Yukawa a
// walked partner is inside at one position at least (the walk's test is
// io || inn), so two selects suffice; with a strict cutoff a partner at
// exactly rc is walked but inside at neither.
template <int KIND>
__device__ __forceinline__ uint32_t n_walked(bool in, bool io, bool inn)
{
    if (KIND == POT_LJ || KIND == POT_YUKAWA) return in ? ((io && inn) ? 2u : 1u) : 0u;
    return in ? (io ? 1u : 0u) + (inn ? 1u : 0u) : 0u;
}

// Energy change for moving the particle at xi (index i) to xn.  Partners:
// left side nearest first and right side nearest first, each stopping at the
// first partner beyond the cutoff from both positions or at index distance
// w; the two side sums are added at the end.  Restates the
// partner sums of qad2 (src/jmmMCState.cpp:1213-1348), whose far pairs
// contribute exactly zero.
template <int KIND>
__device__ __forceinline__ double delta_e64(const Pot &q, const double *X, int64_t i, int64_t N,
                                            bool per, double L, uint32_t w, double xi, double xn,
                                            uint32_t &terms)
{
    double dEL = 0.0, dER = 0.0;
    for (uint32_t k = 1; k <= w; ++k) {
        double xj;
        if (!nbpos(X, i - (int64_t)k, N, per, L, xj)) break;
        const double dold = xi - xj, dnew = xn - xj;
        if (dold > q.rc && dnew > q.rc) break;
        dEL += phi64<KIND>(q, dnew) - phi64<KIND>(q, dold);
        terms += n_inside<KIND>(q, dold, dnew);
    }
    for (uint32_t k = 1; k <= w; ++k) {
        double xj;
        if (!nbpos(X, i + (int64_t)k, N, per, L, xj)) break;
        const double dold = xj - xi, dnew = xj - xn;
        if (dold > q.rc && dnew > q.rc) break;
        dER += phi64<KIND>(q, dnew) - phi64<KIND>(q, dold);
        terms += n_inside<KIND>(q, dold, dnew);
    }
    return dEL + dER;
}

// FP32 form: separations formed in FP64 (positions are FP64), cutoff decided
// in FP64, pair energies in FP32, each side's terms summed in FP64 (exact
// sums of the FP32 terms; an FP64 add issues at the FP32 rate on gfx950).
template <int KIND>
__device__ __forceinline__ double delta_e32(const Pot &q, const double *X, int64_t i, int64_t N,
                                            bool per, double L, uint32_t w, double xi, double xn,
                                            uint32_t &terms)
{
    double left = 0.0, right = 0.0;
    for (uint32_t k = 1; k <= w; ++k) {
        double xj;
        if (!nbpos(X, i - (int64_t)k, N, per, L, xj)) break;
        const double dold = xi - xj, dnew = xn - xj;
        if (dold > q.rc && dnew > q.rc) break;
        const bool io = inside64<KIND>(q, dold), in = inside64<KIND>(q, dnew);
        left += (double)(phi32<KIND>(q, (float)dnew, in) - phi32<KIND>(q, (float)dold, io));
        terms += (io ? 1u : 0u) + (in ? 1u : 0u);
    }
    for (uint32_t k = 1; k <= w; ++k) {
        double xj;
        if (!nbpos(X, i + (int64_t)k, N, per, L, xj)) break;
        const double dold = xj - xi, dnew = xj - xn;
        if (dold > q.rc && dnew > q.rc) break;
        const bool io = inside64<KIND>(q, dold), in = inside64<KIND>(q, dnew);
        right += (double)(phi32<KIND>(q, (float)dnew, in) - phi32<KIND>(q, (float)dold, io));
        terms += (io ? 1u : 0u) + (in ? 1u : 0u);
    }
    return left + right;
}

// Box-length trial energy over the pairs whose left member is in [p0, p1),
// for a scaling of every coordinate by sc.  Sets bad when a scaled nearest
// pair falls inside the hard core.  The energy change pair by pair,
// restating fav's trial energy (src/jmmMCState.cpp:2175-2235) as a
// difference per pair, each partner walk ending at the first partner beyond
// the cutoff at both separations: FP64 form phi64(s d) - phi64(d); FP32 form
// box_term32 of the two separations (summed in FP64).  FP32 form: u0
// receives the same pairs' energy at the unscaled separations (FP32 pair
// energies summed in FP64); for the square well the return value is the
// scaled configuration's energy instead (-eps times its pairs inside the
// well, each walk ending at the first scaled partner outside), from which
// the caller subtracts the running energy (exact, as in vol_lane).
template <typename Real, int KIND>
__device__ __forceinline__ double scaled_delta(const Pot &q, const BoxK &bk, double gmin, const double *X,
                                               int64_t N, bool per, double L, double Lp, double sc,
                                               uint32_t w, uint32_t p0, uint32_t p1, uint32_t stride,
                                               uint32_t &bad, double &u0)
{
    double part = 0.0;
    int wc = 0;
    u0 = 0.0;
    for (uint32_t p = p0; p < p1 && !bad; p += stride) {
        const double xi = X[p], xis = xi * sc;
        for (uint32_t k = 1; k <= w; ++k) {
            const int64_t j = (int64_t)p + k;
            double xj, xjs;
            if (j < N) {
                xj = X[j];
                xjs = xj * sc;
            } else if (per) {
                xj = X[j - N] + L;
                xjs = X[j - N] * sc + Lp;
            } else {
                break;
            }
            const double d0 = xj - xi, d1 = xjs - xis;
            if (k == 1 && d1 < gmin) {
                bad = 1;
                break;
            }
            if (sizeof(Real) == 8) {
                if (d0 > q.rc && d1 > q.rc) break;
                part += phi64<KIND>(q, d1) - phi64<KIND>(q, d0);
            } else {
                const bool in0 = inside64<KIND>(q, d0), in1 = inside64<KIND>(q, d1);
                if (!(in0 || in1)) break;
                if (KIND == POT_HRSW) {
                    if (!in1) break;
                    ++wc;
                    continue;
                }
                float e0;
                part += (double)box_term32<KIND>(q, bk, (float)d0, (float)d1, in0, in1, e0);
                u0 += (double)e0;
            }
        }
    }
    if (sizeof(Real) == 8 || KIND != POT_HRSW) return part;
    return -q.p1 * (double)wc;
}

// FP64 full energy of the pairs whose left member is in [p0, p1) (ECheck's
// recompute, src/jmmMCState.cpp:1973-1993).
template <int KIND>
__device__ __forceinline__ double energy_part(const Pot &q, const double *X, int64_t N, bool per,
                                              double L, uint32_t w, uint32_t p0, uint32_t p1,
                                              uint32_t stride)
{
    double U = 0.0;
    for (uint32_t p = p0; p < p1; p += stride) {
        const double xi = X[p];
        for (uint32_t k = 1; k <= w; ++k) {
            const int64_t j = (int64_t)p + k;
            double xj;
            if (j < N) xj = X[j];
            else if (per) xj = X[j - N] + L;
            else break;
            const double d = xj - xi;
            if (d > q.rc) break;
            U += phi64<KIND>(q, d);
        }
    }
    return U;
}

// Virial and hypervirial (FP64) of the pairs whose left member is in
// [p0, p1) (the reference's Vir and HV totals, src/jmmMCState.cpp:86-96).
template <int KIND>
__device__ __forceinline__ void virial_part(const Pot &q, const double *X, int64_t N, bool per, double L,
                                            uint32_t w, uint32_t p0, uint32_t p1, uint32_t stride,
                                            double &W, double &H)
{
    W = 0.0;
    H = 0.0;
    for (uint32_t p = p0; p < p1; p += stride) {
        const double xi = X[p];
        for (uint32_t k = 1; k <= w; ++k) {
            const int64_t j = (int64_t)p + k;
            double xj;
            if (j < N) xj = X[j];
            else if (per) xj = X[j - N] + L;
            else break;
            const double d = xj - xi;
            if (d > q.rc) break;
            if (!inside64<KIND>(q, d)) continue;
            double wp, hp;
            virial64<KIND>(q, d, wp, hp);
            W += wp;
            H += hp;
        }
    }
}

// Acceptance of a box-length trial (src/jmmMCState.cpp:2249-2255).
__device__ __forceinline__ bool vol_accept(double dU, double dL, double lns, double P, double beta,
                                           double jac, double u)
{
    const double bf = exp(-beta * (dU + P * dL) + jac * lns);
    return bf >= 1.0 || bf > u;
}

struct VolProp {
    double sc, Lp, lns, jac, dL, u;
    BoxK bk;  // FP32 form: the scaling constants of the pair differences
    bool valid;
};

// Out of line: its FP64 log/exp/division temporaries would otherwise raise
// the register count of the whole sweep loop.
__device__ __noinline__ VolProp vol_propose(const DevChain &ch, double L, uint64_t sweep)
{
    uint32_t r[4];
    philox4x32_10(0xFFFFFFFFu, (uint32_t)sweep, ch.chain_id, 0u, (uint32_t)ch.seed,
                  (uint32_t)(ch.seed >> 32), r);
    const double a = u53(r[0], r[1]);
    VolProp v;
    v.u = u53(r[2], r[3]);
    const double dv = ((a - 0.5) * 2.0) * ch.max_dv;
    if (!(ch.flags & F_VOLLOG)) {
        v.Lp = L + dv;
        v.sc = v.Lp / L;
        v.lns = log(v.sc);
        v.jac = (double)ch.N;
        v.dL = dv;
    } else {
        v.sc = det_exp(dv);
        v.Lp = L * v.sc;
        v.lns = dv;
        v.jac = (double)ch.N + 1.0;
        v.dL = v.Lp - L;
    }
    v.valid = v.Lp > 0.0 && (!(ch.flags & F_PERIODIC) || v.Lp > ch.Lmin);
    // s^-12 - 1 and s^-6 - 1 from expm1 of ln s: no cancellation
    const double a6 = ch.pot.e4 * ch.pot.sig6;
    v.bk.c12 = (float)(a6 * ch.pot.sig6 * expm1(-12.0 * v.lns));
    v.bk.c6 = (float)(a6 * expm1(-6.0 * v.lns));
    return v;
}


// Sum over lane segments of size seg (seg <= T) whose first lane is base:
// a power of two up to 64 with base a multiple of seg, or a multiple of 64
// with base a multiple of 64 (whole waves).  Every lane of a segment
// receives the same bits: the butterfly is symmetric in its operands, and
// the cross-wave stage reads the segment's wave totals in a fixed order.
// rd/ri: flat LDS scratch of (T/64)*ND doubles and (T/64)*NI words.
template <int T, int ND, int NI>
__device__ __forceinline__ void seg_allreduce(double (&v)[ND], uint32_t (&k)[NI], uint32_t seg,
                                              uint32_t base, double *rd, uint32_t *ri)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
        for (int m = 0; m < ND; ++m) {
            const double p = __shfl_xor(v[m], off);
            if ((uint32_t)off < seg) v[m] += p;
        }
#pragma unroll
        for (int m = 0; m < NI; ++m) {
            const uint32_t p = __shfl_xor(k[m], off);
            if ((uint32_t)off < seg) k[m] += p;
        }
    }
    if (T > 64) {
        const uint32_t tid = threadIdx.x, wv = tid >> 6;
        if ((tid & 63) == 0) {
#pragma unroll
            for (int m = 0; m < ND; ++m) rd[wv * ND + m] = v[m];
#pragma unroll
            for (int m = 0; m < NI; ++m) ri[wv * NI + m] = k[m];
        }
        __syncthreads();
        if (seg > 64) {
            const uint32_t w0 = base >> 6, nw = seg >> 6;
#pragma unroll
            for (int m = 0; m < ND; ++m) {
                double acc = 0.0;
                for (uint32_t q = 0; q < nw; ++q) acc += rd[(w0 + q) * ND + m];
                v[m] = acc;
            }
#pragma unroll
            for (int m = 0; m < NI; ++m) {
                uint32_t acc = 0;
                for (uint32_t q = 0; q < nw; ++q) acc += ri[(w0 + q) * NI + m];
                k[m] = acc;
            }
        }
        __syncthreads();
    }
}

// Full 64-lane sum, identical in every lane.
template <typename V>
__device__ __forceinline__ V wave_sum(V v)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off);
    return v;
}

// ------------------------------------------------------------ batched form

template <int T>
struct BatchedScratch {
    double rd[(T / 64) * 4];
    uint32_t ri[(T / 64) * 4];
    uint32_t item;
};

template <typename Real, int KIND, int T>
__device__ __forceinline__ void run_bundle_gx(const SweepArgs &a, const DevBundle &bd, BatchedScratch<T> &sc,
                              uint32_t soff)
{
    const uint32_t tid = threadIdx.x;
    // lane -> slot: the largest slot whose first lane is <= tid
    uint32_t lo = 0, hi = bd.nslots - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (a.slots[bd.slot0 + mid].lane_off <= tid) lo = mid;
        else hi = mid - 1;
    }
    const DevSlot me = a.slots[bd.slot0 + lo];
    const bool act = tid >= me.lane_off && tid < me.lane_off + me.B;
    const uint32_t ic = me.chain;

    double *X = a.x + bd.x_off;
    __syncthreads();

    const DevChain &ch = a.chains[ic];
    const Pot q = ch.pot;
    const int64_t N = ch.N;
    const uint32_t w = ch.w;
    const bool per = (ch.flags & F_PERIODIC) != 0;
    const bool npt = (ch.flags & F_NPT) != 0;  // uniform within a bundle
    const double beta = ch.beta, gmin = ch.gmin, max_step = ch.max_step;
    const float betaf = ch.betaf;
    const uint32_t key0 = (uint32_t)ch.seed, key1 = (uint32_t)(ch.seed >> 32);
    const uint64_t log_off = ch.log_off;
    const uint32_t eci = ch.eci;
    const uint32_t chain_id = ch.chain_id;
    const double P = ch.P;

    const DevState st = a.states[ic];
    double L = st.L, U = st.U, mdrift = st.max_drift;
    uint64_t sweep = st.sweep;

    const uint32_t B = me.B;
    const uint32_t r = tid - me.lane_off;
    const uint32_t qb = (uint32_t)N / B, rem = (uint32_t)N % B;
    const uint32_t size = act ? qb + (r < rem ? 1u : 0u) : 0u;
    const uint32_t start = act ? r * qb + (r < rem ? r : rem) : 0u;
    const uint32_t seg = act ? B : 1u;
    const bool leader = act && r == 0;

    uint32_t n_tr = 0, n_wall = 0, n_core = 0, n_terms = 0;
    uint32_t acc_total = 0, n_vt = 0, n_av = 0, n_rc = 0;

    for (uint32_t sw = 0; sw < a.n_sweeps; ++sw, ++sweep) {
        double dUs = 0.0;
        uint32_t acc = 0;
        uint32_t w_md = 0, w_u = 0, w_next_md = 0, w_next_u = 0;
        for (uint32_t t = 0; t < bd.smax; ++t) {
            if (t < size) {
                const uint32_t i = start + t;
                // one Philox block per pair of block positions (t, t+1):
                // computed at even t, its second half kept for odd t (t is
                // the same on every lane, so the branch is uniform)
                if ((t & 1u) == 0) {
                    uint32_t rr[4];
                    philox4x32_10(i, (uint32_t)sweep, chain_id, 0u, key0, key1, rr);
                    w_md = rr[0];
                    w_u = rr[1];
                    w_next_md = rr[2];
                    w_next_u = rr[3];
                } else {
                    w_md = w_next_md;
                    w_u = w_next_u;
                }
                const double a1 = u32o(w_md);
                const double u = u32o(w_u);
                const double md = ((a1 - 0.5) * 2.0) * max_step;
                const double xi = X[i];
                const double xn = xi + md;
                ++n_tr;
                bool ok = true;
                if (!per && fabs(xn) > L / 2.0) {
                    ok = false;
                    ++n_wall;
                } else if (N >= 2) {
                    double xl, xr;
                    if ((nbpos(X, (int64_t)i - 1, N, per, L, xl) && (xn - xl) < gmin) ||
                        (nbpos(X, (int64_t)i + 1, N, per, L, xr) && (xr - xn) < gmin)) {
                        ok = false;
                        ++n_core;
                    }
                }
                bool accd = false;
                if (ok) {
                    double dE;
                    if (sizeof(Real) == 8)
                        dE = delta_e64<KIND>(q, X, i, N, per, L, w, xi, xn, n_terms);
                    else
                        dE = delta_e32<KIND>(q, X, i, N, per, L, w, xi, xn, n_terms);
                    if (dE <= 0.0) accd = true;
                    else if (sizeof(Real) == 8) accd = exp(-dE * beta) > u;
                    else accd = (double)fexp(-betaf * (float)dE) > u;
                    if (accd) {
                        X[i] = xn;
                        dUs += dE;
                        ++acc;
                    }
                }
                if (a.acclog) a.acclog[log_off + (uint64_t)(soff + sw) * N + i] = accd ? 1 : 0;
            }
            __syncthreads();
        }

        // box-length trial (fav, src/jmmMCState.cpp:2161-2293)
        VolProp vp{};
        uint32_t bad = 0;
        double vpart = 0.0, upart = 0.0;
        if (npt) {
            vp = vol_propose(ch, L, sweep);
            if (vp.valid && act)
                vpart = scaled_delta<Real, KIND>(q, vp.bk, gmin, X, N, per, L, vp.Lp, vp.sc, w, start,
                                                 start + size, 1, bad, upart);
        }
        {
            double v[3] = {dUs, vpart, upart};
            uint32_t k2[2] = {acc, bad};
            seg_allreduce<T, 3, 2>(v, k2, seg, me.lane_off, sc.rd, sc.ri);
            U = U + v[0];
            vpart = v[1];
            acc = k2[0];
            bad = k2[1];
            // FP32 form: the box-length pass summed the configuration's FP32
            // energy over every pair; it replaces the running sum, so FP32
            // trial errors do not accumulate between recomputes.  Square
            // well: the pass gave the scaled energy, the running energy is
            // exact
            if (sizeof(Real) == 4 && KIND == POT_HRSW) vpart = vpart - U;
            else if (sizeof(Real) == 4 && npt && vp.valid && !bad) U = v[2];
        }
        acc_total += acc;
        uint32_t accv = 0;
        if (npt) {
            ++n_vt;
            if (vp.valid && !bad && vol_accept(vpart, vp.dL, vp.lns, P, beta, vp.jac, vp.u)) {
                for (uint32_t p = start; p < start + size; ++p) X[p] = X[p] * vp.sc;
                L = vp.Lp;
                U = U + vpart;
                accv = 1;
                ++n_av;
            }
            __syncthreads();
        }

        // drift guard (ECheck, src/jmmMCState.cpp:1965-2095), always FP64
        const int need = act && eci && ((sweep + 1) % eci == 0);
        if (__syncthreads_or(need)) {
            double v[1] = {need ? energy_part<KIND>(q, X, N, per, L, w, start, start + size, 1)
                                : 0.0};
            uint32_t k1[1] = {0};
            seg_allreduce<T, 1, 1>(v, k1, seg, me.lane_off, sc.rd, sc.ri);
            if (need) {
                const double dr = fabs(v[0] - U);
                if (dr > mdrift) mdrift = dr;
                U = v[0];
                ++n_rc;
            }
        }

        double Ws = 0.0, Hs = 0.0;
        if (a.sW) {
            double v[2] = {0.0, 0.0};
            if (act) virial_part<KIND>(q, X, N, per, L, w, start, start + size, 1, v[0], v[1]);
            uint32_t k1[1] = {0};
            seg_allreduce<T, 2, 1>(v, k1, seg, me.lane_off, sc.rd, sc.ri);
            Ws = v[0];
            Hs = v[1];
        }
        if (leader) {
            const uint64_t row = (uint64_t)(soff + sw) * a.n_chains + ic;
            a.sU[row] = U;
            a.sL[row] = L;
            a.sA[row] = make_uint2(acc, accv);
            if (a.sW) {
                a.sW[row] = Ws;
                a.sH[row] = Hs;
            }
        }
    }

    {
        double vz[1] = {0.0};
        uint32_t k4[4] = {n_tr, n_wall, n_core, n_terms};
        seg_allreduce<T, 1, 4>(vz, k4, seg, me.lane_off, sc.rd, sc.ri);
        if (leader) {
            DevState &o = a.states[ic];
            o.L = L;
            o.U = U;
            o.sweep = sweep;
            o.cnt[0] += k4[0];
            o.cnt[1] += acc_total;
            o.cnt[2] += k4[1];
            o.cnt[3] += k4[2];
            o.cnt[4] += k4[3];
            o.cnt[5] += n_vt;
            o.cnt[6] += n_av;
            o.cnt[7] += n_rc;
            o.max_drift = mdrift;
        }
    }
    __syncthreads();
}

// Per-lane view of a chain held in LDS.  Region layout:
//   [ghost-left: wg][block 0 .. block B-1 at stride Sp][ghost-right: wg]
// Sp is odd and at least the longest block, so the lanes of one chain, which
// touch the same offset of consecutive blocks in a phase, hit distinct LDS
// banks.  Periodic chains keep images of the particles near the ends:
// ghost-left[m] = x_{N-wg+m} - L, ghost-right[m] = x_m + L (the same FP64
// operation the oracle applies on the fly), so a neighbour walk never
// branches on the wrap.  wg = max(w, 1).
struct LdsChain {
    double *R;
    uint32_t N, B, w, wg, Sp, qb, rem;
    bool per;
};

__device__ __forceinline__ uint32_t lds_addr(const LdsChain &c, uint32_t i)
{
    const uint32_t big = c.rem * (c.qb + 1);
    uint32_t b, t;
    if (i < big) {
        b = i / (c.qb + 1);
        t = i - b * (c.qb + 1);
    } else {
        b = c.rem + (i - big) / c.qb;
        t = i - (b * c.qb + c.rem);
    }
    return c.wg + b * c.Sp + t;
}

// Images of the first and last wg particles; lanes r = 0..B-1 share the work.
__device__ __forceinline__ void fill_ghosts(const LdsChain &c, uint32_t r, double L)
{
    if (!c.per) return;
    for (uint32_t m = r; m < c.wg && m < c.N; m += c.B) {
        c.R[c.wg + c.B * c.Sp + m] = c.R[lds_addr(c, m)] + L;
        c.R[m] = c.R[lds_addr(c, c.N - c.wg + m)] - L;
    }
}

// The FP32 walks' keep-going test for one partner: inside the cutoff at the
// old or the new position (the FP64 flags).  For LJ and Yukawa (inside is
// d <= rc) this is the FP64 form's rule, not beyond the cutoff from both
// positions.  For the strict cutoffs of the square well and the spring the
// two differ only for a partner at exactly rc: it adds nothing to the sums
// or the pair-term count, and the FP64 walk's next partner lies beyond rc
// and adds nothing either, so both walks give the same result.
template <int KIND>
__device__ __forceinline__ bool walk_in(const Pot &, double, double, bool io, bool in)
{
    return io || in;
}

// One displacement trial of particle i (block position t of a block of
// length size, LDS address my) by md.  Same decisions as orc_trial_with: wall
// test, hard core against the nearest partners (FP64), then the partner sums
// of qad2 (src/jmmMCState.cpp:1213-1348) walked outward on both sides at once
// in a wave-uniform loop; each side stops at its first partner beyond the
// cutoff from both positions.  FP64 form: the oracle's operations exactly.
// FP32 form: separations formed in FP64, cutoff decided in FP64 (the FP64
// set-up and recompute energies count the same pairs), pair energies in FP32
// without branches, each side's FP32 terms summed in FP64 (on gfx950 an FP64
// add issues at the FP32 rate: exact sums of the FP32 terms for fewer
// instructions than a compensated FP32 sum with its freeze selects); the
// loop-exit test runs on the separations in FP64 as in the FP64 form.  Partner k's address is read
// on every lane of the wave, live or not: for k <= w it lies in the chain's
// image (ghosts or halos are wg >= w wide, blocks longer than w), and a lane
// that is not live drops the value.  Returns the decision; dE receives the
// energy change.
//
// PAIR = false: one lane walks both sides.  PAIR = true: two adjacent lanes
// share the trial, side 0 (even lane) walking the left partners and side 1
// the right; each forms its side's sum exactly as the one-lane walk does
// (same terms, same order; once a side's walk ends its terms are zero and
// adding them is exact), then the two sums are exchanged and added.  Addition commutes, so
// both lanes hold the one-lane result bit for bit.  Both lanes of a pair
// must be active together.
//
// W2 (FP32, one lane): two partners per side per loop iteration, four
// independent pair evaluations in flight instead of two and half the
// loop-exit tests; the terms are added in the same order, so the result is
// the one-step walk's bit for bit.
template <typename Real, int KIND, bool PAIR = false, bool W2 = false>
__device__ __forceinline__ bool trial_lds(const LdsChain &c, const Pot &q, double gmin,
                                          double beta, float betaf, uint32_t i, uint32_t t,
                                          uint32_t size, uint32_t my, uint32_t gapL,
                                          uint32_t gapR, double L, double xi, double xn, double md,
                                          double u, double &dE, uint32_t &terms, uint32_t &n_wall,
                                          uint32_t &n_core, uint32_t side = 0)
{
    const double *X = c.R;
    // rejections by the walls and the hard core are flags, not early exits:
    // the wavefront walks the window anyway for its other lanes
    const bool wall = !c.per && fabs(xn) > L / 2.0;
    const bool hasL = c.N >= 2 && (c.per || i > 0);
    const bool hasR = c.N >= 2 && (c.per || i + 1 < c.N);
    // the nearest slots on both sides (ghost or halo slots at a chain's
    // ends: in the image; their values count only where hasL / hasR)
    const double xl = X[my - 1 - (t == 0 ? gapL : 0)];
    const double xr = X[my + 1 + (t + 1 >= size ? gapR : 0)];
    const bool core = !wall && ((hasL && (xn - xl) < gmin) || (hasR && (xr - xn) < gmin));
    n_wall += wall ? 1u : 0u;
    n_core += core ? 1u : 0u;
    const bool go = !wall && !core;
    const uint32_t kL = c.per ? c.w : min(c.w, i);
    const uint32_t kR = c.per ? c.w : min(c.w, c.N - 1 - i);
    if (PAIR) {
        // one side per lane: partner k at my - k (left) or my + k (right),
        // separations oriented as in the two-sided walk
        const bool left = side == 0;
        const uint32_t kS = left ? kL : kR;
        bool live = go && kS >= 1;
        if (sizeof(Real) == 8) {
            double dS = 0.0;
            uint32_t k = 1;
            if (__any(live)) {
#pragma unroll 1
            do {
                const uint32_t jl = my - k - (k > t ? gapL : 0), jr = my + k + (t + k >= size ? gapR : 0);
                const double xj = X[left ? jl : jr];
                const double o = left ? xi - xj : xj - xi, nn = left ? xn - xj : xj - xn;
                const bool in = live && !(o > q.rc && nn > q.rc);
                const double tt = phi64<KIND>(q, nn) - phi64<KIND>(q, o);
                dS += in ? tt : 0.0;
                terms += in ? n_inside<KIND>(q, o, nn) : 0u;
                live = in && k < kS;
                ++k;
            } while (__any(live));
            }
            dE = dS + __shfl_xor(dS, 1);
            const double bf = exp(-dE * beta);
            return go && (dE <= 0.0 || bf > u);
        }
        double sS = 0.0;  // the side's FP32 pair terms summed in FP64
        int wS = 0;  // hard rod + square well: the well count, as the one-lane walk
        uint32_t k = 1;
        if (__any(live)) {
#pragma unroll 1
        do {
            const uint32_t jl = my - k - (k > t ? gapL : 0), jr = my + k + (t + k >= size ? gapR : 0);
            const double xj = X[left ? jl : jr];
            const double o = left ? xi - xj : xj - xi, nn = left ? xn - xj : xj - xn;
            const bool io = inside64<KIND>(q, o), inn = inside64<KIND>(q, nn);
            const bool in = live && walk_in<KIND>(q, o, nn, io, inn);
            if (KIND == POT_HRSW) {
                wS += in ? (int)inn - (int)io : 0;
            } else {
                // terms of partners off the walk are zero: exact no-op adds
                sS += (double)(phi32<KIND>(q, (float)nn, in && inn) - phi32<KIND>(q, (float)o, in && io));
            }
            terms += n_walked<KIND>(in, io, inn);
            live = in && k < kS;
            ++k;
        } while (__any(live));
        }
        const double mine = KIND == POT_HRSW ? -q.p1 * (double)wS : sS;
        dE = mine + __shfl_xor(mine, 1);
        const float bf = fexp(-betaf * (float)dE);
        return go && (dE <= 0.0 || (double)bf > u);
    }
    bool liveL = go && kL >= 1, liveR = go && kR >= 1;
    if (sizeof(Real) == 8) {
        double dEL = 0.0, dER = 0.0;
        uint32_t k = 1;
        if (__any(liveL || liveR)) {
#pragma unroll 1
        do {
            const double xjL = X[my - k - (k > t ? gapL : 0)];
            const double xjR = X[my + k + (t + k >= size ? gapR : 0)];
            const double oL = xi - xjL, nL = xn - xjL;
            const double oR = xjR - xi, nR = xjR - xn;
            const bool inL = liveL && !(oL > q.rc && nL > q.rc);
            const bool inR = liveR && !(oR > q.rc && nR > q.rc);
            const double tL = phi64<KIND>(q, nL) - phi64<KIND>(q, oL);
            const double tR = phi64<KIND>(q, nR) - phi64<KIND>(q, oR);
            dEL += inL ? tL : 0.0;
            dER += inR ? tR : 0.0;
            terms += n_walked<KIND>(inL, inside64<KIND>(q, oL), inside64<KIND>(q, nL)) +
                     n_walked<KIND>(inR, inside64<KIND>(q, oR), inside64<KIND>(q, nR));
            liveL = inL && k < kL;
            liveR = inR && k < kR;
            ++k;
        } while (__any(liveL || liveR));
        }
        dE = dEL + dER;
        const double bf = exp(-dE * beta);
        return go && (dE <= 0.0 || bf > u);
    }
    // each side's FP32 pair terms summed in FP64 (exact; the terms of
    // partners off a side's walk are zero, exact no-op adds)
    double sL = 0.0, sR = 0.0;
    // hard rod + square well: every pair term is -eps or 0, so each side's
    // change is -eps times the change in the number of partners inside the
    // well, counted exactly
    int wL = 0, wR = 0;
    uint32_t k = 1;
    if (W2) {
    if (__any(liveL || liveR)) {
#pragma unroll 1
    do {
        // partners k and k + 1 on both sides: a side still walking after
        // step k has k < its limit <= w, so partner k + 1 is in the image;
        // the clamp keeps the other lanes' reads there too
        const uint32_t k2 = min(k + 1, c.w);
        const double xjL = X[my - k - (k > t ? gapL : 0)];
        const double xjR = X[my + k + (t + k >= size ? gapR : 0)];
        const double xjL2 = X[my - k2 - (k2 > t ? gapL : 0)];
        const double xjR2 = X[my + k2 + (t + k2 >= size ? gapR : 0)];
        const double oL = xi - xjL, nL = xn - xjL;
        const double oR = xjR - xi, nR = xjR - xn;
        const bool ioL = inside64<KIND>(q, oL), inL64 = inside64<KIND>(q, nL);
        const bool ioR = inside64<KIND>(q, oR), inR64 = inside64<KIND>(q, nR);
        const bool inL = liveL && walk_in<KIND>(q, oL, nL, ioL, inL64);
        const bool inR = liveR && walk_in<KIND>(q, oR, nR, ioR, inR64);
        const float tL = phi32<KIND>(q, (float)nL, inL && inL64) - phi32<KIND>(q, (float)oL, inL && ioL);
        const float tR = phi32<KIND>(q, (float)nR, inR && inR64) - phi32<KIND>(q, (float)oR, inR && ioR);
        const bool liveL2 = inL && k < kL, liveR2 = inR && k < kR;
        const double oL2 = xi - xjL2, nL2 = xn - xjL2;
        const double oR2 = xjR2 - xi, nR2 = xjR2 - xn;
        const bool ioL2 = inside64<KIND>(q, oL2), inL642 = inside64<KIND>(q, nL2);
        const bool ioR2 = inside64<KIND>(q, oR2), inR642 = inside64<KIND>(q, nR2);
        const bool inL2 = liveL2 && walk_in<KIND>(q, oL2, nL2, ioL2, inL642);
        const bool inR2 = liveR2 && walk_in<KIND>(q, oR2, nR2, ioR2, inR642);
        const float tL2 = phi32<KIND>(q, (float)nL2, inL2 && inL642) - phi32<KIND>(q, (float)oL2, inL2 && ioL2);
        const float tR2 = phi32<KIND>(q, (float)nR2, inR2 && inR642) - phi32<KIND>(q, (float)oR2, inR2 && ioR2);
        if (KIND == POT_HRSW) {
            wL += (inL ? (int)inL64 - (int)ioL : 0) + (inL2 ? (int)inL642 - (int)ioL2 : 0);
            wR += (inR ? (int)inR64 - (int)ioR : 0) + (inR2 ? (int)inR642 - (int)ioR2 : 0);
        } else {
            sL += (double)tL;
            sR += (double)tR;
            sL += (double)tL2;
            sR += (double)tR2;
        }
        terms += n_walked<KIND>(inL, ioL, inL64) + n_walked<KIND>(inR, ioR, inR64) +
                 n_walked<KIND>(inL2, ioL2, inL642) + n_walked<KIND>(inR2, ioR2, inR642);
        liveL = inL2 && k + 1 < kL;
        liveR = inR2 && k + 1 < kR;
        k += 2;
    } while (__any(liveL || liveR));
    }
    } else {
    if (__any(liveL || liveR)) {
#pragma unroll 1
    do {
        const double xjL = X[my - k - (k > t ? gapL : 0)];
        const double xjR = X[my + k + (t + k >= size ? gapR : 0)];
        // separations are FP64 differences; the cutoff decisions are made on
        // them (the FP64 energies of set-up and recompute count the same
        // pairs, and the walk ends where the FP64 form's does), the pair
        // energies on their FP32 roundings
        const double oL = xi - xjL, nL = xn - xjL;
        const double oR = xjR - xi, nR = xjR - xn;
        const bool ioL = inside64<KIND>(q, oL), inL64 = inside64<KIND>(q, nL);
        const bool ioR = inside64<KIND>(q, oR), inR64 = inside64<KIND>(q, nR);
        const bool inL = liveL && walk_in<KIND>(q, oL, nL, ioL, inL64);
        const bool inR = liveR && walk_in<KIND>(q, oR, nR, ioR, inR64);
        const float tL = phi32<KIND>(q, (float)nL, inL && inL64) - phi32<KIND>(q, (float)oL, inL && ioL);
        const float tR = phi32<KIND>(q, (float)nR, inR && inR64) - phi32<KIND>(q, (float)oR, inR && ioR);
        if (KIND == POT_HRSW) {
            wL += inL ? (int)inL64 - (int)ioL : 0;
            wR += inR ? (int)inR64 - (int)ioR : 0;
        } else {
            sL += (double)tL;
            sR += (double)tR;
        }
        terms += n_walked<KIND>(inL, ioL, inL64) + n_walked<KIND>(inR, ioR, inR64);
        liveL = inL && k < kL;
        liveR = inR && k < kR;
        ++k;
    } while (__any(liveL || liveR));
    }
    }
    dE = KIND == POT_HRSW ? -q.p1 * (double)wL + -q.p1 * (double)wR : sL + sR;
    const float bf = fexp(-betaf * (float)dE);
    return go && (dE <= 0.0 || (double)bf > u);
}

// Box-length trial (fav, src/jmmMCState.cpp:2161-2293), one lane's share
// over the pairs (i, i+k) whose left member is in the lane's block, in the
// order t = 0.., k = 1.. : the energy change of scaling the configuration by
// vp.sc pair by pair, FP64 form phi64(s d) - phi64(d), FP32 form box_term32
// (one reciprocal per LJ pair through the scaling identity) summed in FP64;
// each particle's walk ends at its first partner beyond
// the cutoff at both the unscaled and the scaled separation (the FP64
// inside-the-cutoff flags decide in both forms, so both sum the same pairs).
// The walk over (particle, partner) is one
// flat loop with per-lane state, so lanes at different particles advance
// together without nested divergent loops; it adds the same terms in the
// same order as the nested walk (particle t's walk ends at its first partner
// beyond the cutoff from both scaled and unscaled positions, at k = kR, or
// at an overlap of nearest neighbours, which ends the lane's walk and sets
// `bad`).  The partner at k = 1 is the next particle, so its position and
// scaled position are kept rather than read again.  `last`: the lane holds
// the chain's last block of a periodic chain, whose partners past the end
// are the first particles shifted by L; wrap(m) returns particle m's
// unshifted position (scaled and shifted by Lp here).  Hard rod + square
// well, FP32 form: the lane returns the scaled configuration's energy, -eps
// times the number of its scaled pairs inside the well (each walk ends at
// the first scaled partner outside), and the caller subtracts the running
// energy, which is -eps times the unscaled count exactly (every trial and
// recompute changes it by whole pairs), so the energy change is the one the
// pairwise sum gives.  W2: two partners of
// a particle per iteration (for the common walk of one partner inside and
// the next beyond the cutoff, one iteration per particle), same terms in
// the same order.
template <typename Real, int KIND, bool W2 = false, typename Wrap>
__device__ __forceinline__ double vol_lane(const LdsChain &c, const Pot &q, double gmin,
                                           const VolProp &vp, bool on, uint32_t start,
                                           uint32_t size, uint32_t a0, uint32_t gapR, bool last,
                                           Wrap wrap, uint32_t &bad, double &u0)
{
    constexpr bool WELL = sizeof(Real) == 4 && KIND == POT_HRSW;
    const double *X = c.R;
    double part = 0.0;  // FP64 terms, or the FP32 form's FP32 terms summed in FP64
    int wc = 0;  // hard rod + square well, FP32: scaled pairs inside the well
    u0 = 0.0;    // FP32 form: the pairs' FP32 energy at the unscaled separations
    bool run = on && size > 0;
    uint32_t t = 0, k = 1;
    double xi = X[run ? a0 : 0u];
    double xis = xi * vp.sc;
    double x1 = 0.0, x1s = 0.0;
    if (W2) {
    if (__any(run)) {
#pragma unroll 1
    do {
        // partners k and k + 1 of particle t; the second counts only where
        // the first continued the walk (then k < kR <= w, so the clamp to w
        // changes no address that counts and keeps the others in the image)
        const uint32_t i = start + t;
        const uint32_t kR = c.per ? c.w : min(c.w, c.N - 1 - i);
        const bool valid = run && k <= kR;
        const uint32_t k2 = min(k + 1, c.w);
        const bool cross = t + k >= size, cross2 = t + k2 >= size;
        const double xj = X[valid ? a0 + t + k + (cross ? gapR : 0u) : 0u];
        const double xj2 = X[valid ? a0 + t + k2 + (cross2 ? gapR : 0u) : 0u];
        double xjs = xj * vp.sc, xjs2 = xj2 * vp.sc;
        if (valid && cross && last) xjs = wrap(t + k - size) * vp.sc + vp.Lp;
        if (valid && cross2 && last) xjs2 = wrap(t + k2 - size) * vp.sc + vp.Lp;
        const double d0 = xj - xi, d1 = xjs - xis;
        const double e0 = xj2 - xi, e1 = xjs2 - xis;
        const bool in0 = inside64<KIND>(q, d0), in1 = inside64<KIND>(q, d1);
        const bool jn0 = inside64<KIND>(q, e0), jn1 = inside64<KIND>(q, e1);
        const bool isbad = valid && k == 1 && d1 < gmin;
        const bool stop = valid && (sizeof(Real) == 8 ? d0 > q.rc && d1 > q.rc : WELL ? !in1 : !(in0 || in1));
        const bool add = valid && !isbad && !stop;
        const bool valid2 = add && k < kR;
        const bool stop2 = valid2 && (sizeof(Real) == 8 ? e0 > q.rc && e1 > q.rc : WELL ? !jn1 : !(jn0 || jn1));
        const bool add2 = valid2 && !stop2;
        if (sizeof(Real) == 8) {
            const double tt = phi64<KIND>(q, d1) - phi64<KIND>(q, d0);
            const double tt2 = phi64<KIND>(q, e1) - phi64<KIND>(q, e0);
            part = add ? part + tt : part;
            part = add2 ? part + tt2 : part;
        } else if (WELL) {
            wc += (add && in1 ? 1 : 0) + (add2 && jn1 ? 1 : 0);
        } else {
            // flags masked by add / add2: pairs off the walk give zeros
            float g0, g1;
            const float tt = box_term32<KIND>(q, vp.bk, (float)d0, (float)d1, add && in0, add && in1, g0);
            const float tt2 = box_term32<KIND>(q, vp.bk, (float)e0, (float)e1, add2 && jn0, add2 && jn1, g1);
            part += (double)tt;
            part += (double)tt2;
            u0 += (double)g0;
            u0 += (double)g1;
        }
        bad |= isbad ? 1u : 0u;
        x1 = k == 1 ? xj : x1;
        x1s = k == 1 ? xjs : x1s;
        // the same particle goes on at k + 2 only if partner k + 1 was
        // inside and more partners remain
        const bool nextp = !(add2 && k + 1 < kR);
        const uint32_t tn = nextp ? t + 1 : t;
        run = run && !isbad && tn < size;
        xi = nextp ? x1 : xi;
        xis = nextp ? x1s : xis;
        k = nextp ? 1u : k + 2;
        t = tn;
    } while (__any(run));
    }
    } else {
    if (__any(run)) {
#pragma unroll 1
    do {
        const uint32_t i = start + t;
        const uint32_t kR = c.per ? c.w : min(c.w, c.N - 1 - i);
        const bool valid = run && k <= kR;
        const bool cross = t + k >= size;
        const double xj = X[valid ? a0 + t + k + (cross ? gapR : 0u) : 0u];
        double xjs = xj * vp.sc;
        if (valid && cross && last) xjs = wrap(t + k - size) * vp.sc + vp.Lp;
        const double d0 = xj - xi, d1 = xjs - xis;
        const bool in0 = inside64<KIND>(q, d0), in1 = inside64<KIND>(q, d1);
        const bool isbad = valid && k == 1 && d1 < gmin;
        const bool stop = valid && (sizeof(Real) == 8 ? d0 > q.rc && d1 > q.rc : WELL ? !in1 : !(in0 || in1));
        const bool add = valid && !isbad && !stop;
        if (sizeof(Real) == 8) {
            const double tt = phi64<KIND>(q, d1) - phi64<KIND>(q, d0);
            part = add ? part + tt : part;
        } else if (WELL) {
            wc += add && in1 ? 1 : 0;
        } else {
            float g0;
            part += (double)box_term32<KIND>(q, vp.bk, (float)d0, (float)d1, add && in0, add && in1, g0);
            u0 += (double)g0;
        }
        bad |= isbad ? 1u : 0u;
        x1 = k == 1 ? xj : x1;
        x1s = k == 1 ? xjs : x1s;
        // next partner, or the next particle (lanes no longer running keep
        // stepping t; their reads are redirected and their terms dropped)
        const bool nextp = !valid || stop || k == kR;
        const uint32_t tn = nextp ? t + 1 : t;
        run = run && !isbad && tn < size;
        xi = nextp ? x1 : xi;
        xis = nextp ? x1s : xis;
        k = nextp ? 1u : k + 1;
        t = tn;
    } while (__any(run));
    }
    }
    return WELL ? -q.p1 * (double)wc : part;
}

// Virial and hypervirial (FP64) of the pairs whose left member is in the
// lane's block, read from the LDS image (the recompute's walk).
template <int KIND>
__device__ __forceinline__ void virial_lane(const LdsChain &c, const Pot &q, uint32_t start, uint32_t size,
                                            uint32_t a0, uint32_t gapR, double &W, double &H)
{
    for (uint32_t t = 0; t < size; ++t) {
        const uint32_t i = start + t;
        const double xi = c.R[a0 + t];
        const uint32_t kR = c.per ? c.w : min(c.w, c.N - 1 - i);
        for (uint32_t k = 1; k <= kR; ++k) {
            const double d = c.R[a0 + t + k + (t + k >= size ? gapR : 0)] - xi;
            if (d > q.rc) break;
            if (!inside64<KIND>(q, d)) continue;
            double wp, hp;
            virial64<KIND>(q, d, wp, hp);
            W += wp;
            H += hp;
        }
    }
}

template <typename Real, int KIND, int T>
__device__ __forceinline__ void run_bundle_lds(const SweepArgs &a, const DevBundle &bd, double *smem,
                                            BatchedScratch<T> &sc, uint32_t soff)
{
    const uint32_t tid = threadIdx.x;
    uint32_t lo = 0, hi = bd.nslots - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (a.slots[bd.slot0 + mid].lane_off <= tid) lo = mid;
        else hi = mid - 1;
    }
    const DevSlot me = a.slots[bd.slot0 + lo];
    const bool act = tid >= me.lane_off && tid < me.lane_off + me.B;
    const uint32_t ic = me.chain;
    const DevChain &ch = a.chains[ic];
    // Chains of a bundle share their shape (potential and its parameters,
    // boundary, ensemble, step sizes, window): read it once per bundle.
    const DevChain &sh = a.chains[__builtin_amdgcn_readfirstlane(a.slots[bd.slot0].chain)];
    const Pot q = sh.pot;
    const double gmin = sh.gmin;
    const double max_step = ch.max_step;  // per chain: step adaptation may change it
    // per-chain fields as values: read once, not reloaded after every store
    // of the loops below (the compiler cannot rule out aliasing)
    const double beta = ch.beta;
    const float betaf = ch.betaf;
    const uint32_t key0 = (uint32_t)ch.seed, key1 = (uint32_t)(ch.seed >> 32);
    const uint32_t chain_id = ch.chain_id;
    const uint64_t log_off = ch.log_off;
    const double P = ch.P;
    const uint32_t eci = ch.eci;

    LdsChain c;
    c.R = smem + me.lds_off;
    c.N = ch.N;
    c.B = me.B;
    c.w = sh.w;
    c.wg = sh.wg;
    c.Sp = ch.Sp;
    c.qb = ch.N / me.B;
    c.rem = ch.N % me.B;
    c.per = (sh.flags & F_PERIODIC) != 0;
    const bool npt = (sh.flags & F_NPT) != 0;
    // FP32: walks of one partner per step where the window is one partner
    const bool one = sizeof(Real) == 4 && c.w == 1;

    const uint32_t r = tid - me.lane_off;
    const uint32_t size = act ? c.qb + (r < c.rem ? 1u : 0u) : 0u;
    const uint32_t start = act ? r * c.qb + (r < c.rem ? r : c.rem) : 0u;
    const uint32_t a0 = c.wg + r * c.Sp;
    const uint32_t gapR = c.Sp - size;
    const uint32_t gapL = r == 0 ? 0u : c.Sp - (c.qb + (r - 1 < c.rem ? 1u : 0u));
    const uint32_t seg = act ? c.B : 1u;
    const bool leader = act && r == 0;

    const double *xg = a.x + ch.x_off;
    for (uint32_t t = 0; t < size; ++t) c.R[a0 + t] = xg[start + t];
    const DevState st = a.states[ic];
    double L = st.L, U = st.U, mdrift = st.max_drift;
    uint64_t sweep = st.sweep;
    __syncthreads();
    if (act) fill_ghosts(c, r, L);
    __syncthreads();

    uint32_t n_tr = 0, n_wall = 0, n_core = 0, n_terms = 0;
    uint32_t acc_total = 0, n_vt = 0, n_av = 0, n_rc = 0;

#pragma unroll 1
    for (uint32_t sw = 0; sw < a.n_sweeps; ++sw, ++sweep) {
        double dUs = 0.0;
        uint32_t acc = 0;
        uint32_t w_md = 0, w_u = 0, w_next_md = 0, w_next_u = 0;
#pragma unroll 1
        for (uint32_t t = 0; t < bd.smax; ++t) {
            if (t < size) {
                const uint32_t i = start + t;
                const uint32_t my = a0 + t;
                // one Philox block per pair of block positions (t, t+1):
                // computed at even t, its second half kept for odd t (t is
                // the same on every lane, so the branch is uniform)
                if ((t & 1u) == 0) {
                    uint32_t rr[4];
                    philox4x32_10(i, (uint32_t)sweep, chain_id, 0u, key0, key1, rr);
                    w_md = rr[0];
                    w_u = rr[1];
                    w_next_md = rr[2];
                    w_next_u = rr[3];
                } else {
                    w_md = w_next_md;
                    w_u = w_next_u;
                }
                const double a1 = u32o(w_md);
                const double u = u32o(w_u);
                const double md = ((a1 - 0.5) * 2.0) * max_step;
                const double xi = c.R[my];
                const double xn = xi + md;
                ++n_tr;
                double dE = 0.0;
                // a window of one partner per side (the hard core keeps the
                // second beyond the cutoff): the one-partner walk, which
                // evaluates no clamped duplicate (wave-uniform: a bundle
                // shares its window)
                const bool accd =
                    one ? trial_lds<Real, KIND, false, false>(c, q, gmin, beta, betaf, i, t, size, my, gapL, gapR,
                                                              L, xi, xn, md, u, dE, n_terms, n_wall, n_core)
                        : trial_lds<Real, KIND, false, true>(c, q, gmin, beta, betaf, i, t, size, my, gapL, gapR,
                                                             L, xi, xn, md, u, dE, n_terms, n_wall, n_core);
                const double xs = accd ? xn : xi;
                c.R[my] = xs;
                if (c.per) {
                    if (i < c.wg) c.R[c.wg + c.B * c.Sp + i] = xs + L;
                    if (i + c.wg >= c.N) c.R[i + c.wg - c.N] = xs - L;
                }
                dUs += accd ? dE : 0.0;
                acc += accd ? 1u : 0u;
                if (a.acclog) a.acclog[log_off + (uint64_t)(soff + sw) * c.N + i] = accd ? 1 : 0;
            }
            __syncthreads();
        }

        // box-length trial (fav, src/jmmMCState.cpp:2161-2293): energy change
        // pair by pair over the pairs whose left member is in this block
        VolProp vp{};
        uint32_t bad = 0;
        double vpart = 0.0, upart = 0.0;
        if (npt) {
            vp = vol_propose(ch, L, sweep);
            auto wrap = [&](uint32_t m) { return c.R[c.wg + m]; };
            const bool last = c.per && r + 1 == c.B, von = vp.valid && act;
            vpart = one ? vol_lane<Real, KIND, false>(c, q, gmin, vp, von, start, size, a0, gapR, last, wrap, bad,
                                                      upart)
                        : vol_lane<Real, KIND, sizeof(Real) == 4>(c, q, gmin, vp, von, start, size, a0, gapR, last,
                                                                  wrap, bad, upart);
        }
        {
            double v[3] = {dUs, vpart, upart};
            uint32_t k2[2] = {acc, bad};
            seg_allreduce<T, 3, 2>(v, k2, seg, me.lane_off, sc.rd, sc.ri);
            U = U + v[0];
            vpart = v[1];
            acc = k2[0];
            bad = k2[1];
            // FP32 form: the box-length pass summed the configuration's FP32
            // energy over every pair; it replaces the running sum, so FP32
            // trial errors do not accumulate between recomputes.  Square
            // well: the pass gave the scaled energy, the running energy is
            // exact
            if (sizeof(Real) == 4 && KIND == POT_HRSW) vpart = vpart - U;
            else if (sizeof(Real) == 4 && npt && vp.valid && !bad) U = v[2];
        }
        acc_total += acc;
        uint32_t accv = 0;
        if (npt) {
            ++n_vt;
            if (vp.valid && !bad && vol_accept(vpart, vp.dL, vp.lns, P, beta, vp.jac, vp.u)) {
                for (uint32_t t = 0; t < size; ++t) c.R[a0 + t] = c.R[a0 + t] * vp.sc;
                L = vp.Lp;
                U = U + vpart;
                accv = 1;
                ++n_av;
            }
            __syncthreads();
            // only the chain's own lanes: idle lanes of the workgroup carry no
            // energy partials and so no valid decision
            if (accv && act) fill_ghosts(c, r, L);
            __syncthreads();
        }

        // drift guard (ECheck, src/jmmMCState.cpp:1965-2095), always FP64
        const int need = act && eci && ((sweep + 1) % eci == 0);
        if (__syncthreads_or(need)) {
            double Up = 0.0;
            if (need) {
                for (uint32_t t = 0; t < size; ++t) {
                    const uint32_t i = start + t;
                    const uint32_t my = a0 + t;
                    const double xi = c.R[my];
                    const uint32_t kR = c.per ? c.w : min(c.w, c.N - 1 - i);
                    for (uint32_t k = 1; k <= kR; ++k) {
                        const double d = c.R[my + k + (t + k >= size ? gapR : 0)] - xi;
                        if (d > q.rc) break;
                        Up += phi64<KIND>(q, d);
                    }
                }
            }
            double v[1] = {Up};
            uint32_t k1[1] = {0};
            seg_allreduce<T, 1, 1>(v, k1, seg, me.lane_off, sc.rd, sc.ri);
            if (need) {
                const double dr = fabs(v[0] - U);
                if (dr > mdrift) mdrift = dr;
                U = v[0];
                ++n_rc;
            }
        }

        double Ws = 0.0, Hs = 0.0;
        if (a.sW) {
            double v[2] = {0.0, 0.0};
            if (act) virial_lane<KIND>(c, q, start, size, a0, gapR, v[0], v[1]);
            uint32_t k1[1] = {0};
            seg_allreduce<T, 2, 1>(v, k1, seg, me.lane_off, sc.rd, sc.ri);
            Ws = v[0];
            Hs = v[1];
        }
        if (leader) {
            const uint64_t row = (uint64_t)(soff + sw) * a.n_chains + ic;
            a.sU[row] = U;
            a.sL[row] = L;
            a.sA[row] = make_uint2(acc, accv);
            if (a.sW) {
                a.sW[row] = Ws;
                a.sH[row] = Hs;
            }
        }
    }

    {
        double vz[1] = {0.0};
        uint32_t k4[4] = {n_tr, n_wall, n_core, n_terms};
        seg_allreduce<T, 1, 4>(vz, k4, seg, me.lane_off, sc.rd, sc.ri);
        if (leader) {
            DevState &o = a.states[ic];
            o.L = L;
            o.U = U;
            o.sweep = sweep;
            o.cnt[0] += k4[0];
            o.cnt[1] += acc_total;
            o.cnt[2] += k4[1];
            o.cnt[3] += k4[2];
            o.cnt[4] += k4[3];
            o.cnt[5] += n_vt;
            o.cnt[6] += n_av;
            o.cnt[7] += n_rc;
            o.max_drift = mdrift;
        }
    }
    double *xw = a.x + ch.x_off;
    for (uint32_t t = 0; t < size; ++t) xw[start + t] = c.R[a0 + t];
    __syncthreads();
}

// Persistent kernel, one instantiation per (precision, workgroup size,
// memory mode): every workgroup pulls bundles from one queue (all potentials
// together, longest first) until it is empty; all waves of a workgroup leave
// together when the queue is drained.  The potential is uniform per bundle.
template <typename Real, int T, bool GX>
__global__ void __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(JMM_WAVES_PER_EU)))
k_sweep_batched(SweepArgs a)
{
    extern __shared__ double smem[];
    __shared__ BatchedScratch<T> sc;
    const uint32_t soff = a.ctrl->sample_off;
    for (;;) {
        if (threadIdx.x == 0) sc.item = atomicAdd(&a.ctrl->queue, 1u);
        __syncthreads();
        // the same value in every lane: keep the bundle in scalar registers
        const uint32_t item = __builtin_amdgcn_readfirstlane(sc.item);
        __syncthreads();
        if (item >= a.n_items) break;
        const DevBundle bd = a.bundles[item];
        if (GX) {
            switch (bd.kind) {
            case POT_LJ: run_bundle_gx<Real, POT_LJ, T>(a, bd, sc, soff); break;
            case POT_HARMONIC: run_bundle_gx<Real, POT_HARMONIC, T>(a, bd, sc, soff); break;
            case POT_HRSW: run_bundle_gx<Real, POT_HRSW, T>(a, bd, sc, soff); break;
            default: run_bundle_gx<Real, POT_YUKAWA, T>(a, bd, sc, soff); break;
            }
        } else {
            switch (bd.kind) {
            case POT_LJ: run_bundle_lds<Real, POT_LJ, T>(a, bd, smem, sc, soff); break;
            case POT_HARMONIC: run_bundle_lds<Real, POT_HARMONIC, T>(a, bd, smem, sc, soff); break;
            case POT_HRSW: run_bundle_lds<Real, POT_HRSW, T>(a, bd, smem, sc, soff); break;
            default: run_bundle_lds<Real, POT_YUKAWA, T>(a, bd, smem, sc, soff); break;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------- segmented chains

// Synchronisation between the workgroups (segments) of one chain.  A
// workgroup barrier brings every wave's boundary stores to the CU's L2;
// thread 0 then fences at agent scope (one write-back of that L2 per
// workgroup, where a fence in every wave cost one per wave) and advances the
// segment's counter.  A waiter polls its partners' counters with agent-scope
// loads and, once all have arrived, fences once before the workgroup reads
// what they published.  Every wait is bounded: after 2^22
// polls the error flag is set and all later waits return at once, so the
// grid drains in bounded time and the host reports the error.
__device__ __forceinline__ void seg_signal(unsigned long long *ctr, unsigned long long v)
{
    __threadfence();
    __hip_atomic_store(ctr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __noinline__ void seg_wait(const unsigned long long *ctr, unsigned long long target,
                                      uint32_t *err)
{
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    for (uint32_t it = 0; it < (1u << 22); ++it) {
        if (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) return;
        __builtin_amdgcn_s_sleep(2);
    }
    __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double seg_load(const double *p)
{
    const unsigned long long v =
        __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)v);
}

// Slot of a segment's partial sum: [segment][sweep parity][field][wave].
__device__ __forceinline__ uint64_t seg_pidx(uint32_t j, uint32_t par, uint32_t f, uint32_t wv)
{
    return (((uint64_t)j * 2 + par) * SEG_FIELDS + f) * SEG_WAVES + wv;
}

// One segment of a chain: blocks s*Bs .. s*Bs+Bs-1, one lane each, in LDS
// with halos of wg positions on either side that hold the neighbouring
// segments' boundary particles (periodic images at the chain's ends, as the
// single-workgroup layout's ghosts).  The trials, the random numbers, the
// per-lane sums and the order in which the sums are combined are those of
// run_bundle_lds, so a split chain produces the same bits as a whole one.
//
// Neighbour syncs every K = qb - wg phases (qb the smaller block size): in
// phase t the last block of a segment reads the right neighbour's first-block
// particles m <= t + wg - size, updated in phases before the previous sync,
// and the first block reads the left neighbour's last-block particles
// p >= size - wg + t, which the neighbour (at most K phases ahead) has not
// updated yet in this sweep.  So between two syncs every halo value read is
// the one the sequential schedule reads.
//
// LPB = 2 gives every block two lanes (adjacent; the even lane owns the
// block's state): they share each displacement trial, one walking the left
// partners and one the right (trial_lds with PAIR), which halves the walk a
// lane does and doubles the waves per CU when LDS limits the workgroups.
// Everything else (loads, stores, the box-length and recompute sums) is the
// even lane's, and the sums are combined per 64 blocks as with one lane per
// block, so LPB changes no result bit.  Sync points, the same in every
// segment of a chain: the initial publication (neighbours); phases
// K-1, 2K-1, ... and the last phase (neighbours); every sweep's partial sums
// (all segments); a box-length acceptance (neighbours); a recompute (all);
// the final counters (all).  Partial sums are double-buffered by sweep
// parity: a segment is at most one all-segment point ahead of another.
template <typename Real, int KIND, int TW, int LPB>
__device__ __forceinline__ void run_segment(const SegArgs &g, const DevSeg &sg, double *smem,
                            BatchedScratch<TW> &sc, uint32_t soff)
{
    const SweepArgs &a = g.a;
    const uint32_t tid = threadIdx.x;
    const uint32_t ic = sg.chain, S = sg.S, s = sg.s, Bs = sg.Bs;
    const uint32_t me = sg.first + s;
    const DevChain &ch = a.chains[ic];
    const Pot q = ch.pot;
    const double gmin = ch.gmin, max_step = ch.max_step;
    const double beta = ch.beta, P = ch.P;
    const float betaf = ch.betaf;
    const uint32_t key0 = (uint32_t)ch.seed, key1 = (uint32_t)(ch.seed >> 32);
    const uint32_t chain_id = ch.chain_id, eci = ch.eci;
    const uint64_t log_off = ch.log_off;
    const bool per = (ch.flags & F_PERIODIC) != 0, npt = (ch.flags & F_NPT) != 0;
    const uint32_t N = ch.N, w = ch.w, wg = ch.wg, Sp = ch.Sp, smax = ch.smax;
    const uint32_t qb = N / ch.B, rem = N % ch.B;
    const bool hasLn = per || s > 0, hasRn = per || s + 1 < S;
    const uint32_t ln = sg.first + (s + S - 1) % S, rn = sg.first + (s + 1) % S;
    const uint32_t W = g.wgmax;
    const uint32_t K = qb > wg ? qb - wg : 1u;
    const uint32_t nwv = Bs >= 64 ? Bs / 64 : 1u;  // 64-block totals per segment and field

    LdsChain c;
    c.R = smem;
    c.N = N;
    c.B = Bs;
    c.w = w;
    c.wg = wg;
    c.Sp = Sp;
    c.qb = qb;
    c.rem = rem;
    c.per = per;

    const uint32_t r = LPB == 2 ? tid >> 1 : tid;
    const uint32_t side = LPB == 2 ? (tid & 1u) : 0u;
    const bool own = side == 0;  // the lane that owns the block's state
    const bool act = r < Bs;
    const uint32_t b = s * Bs + r;
    const uint32_t size = act ? qb + (b < rem ? 1u : 0u) : 0u;
    const uint32_t start = act ? b * qb + (b < rem ? b : rem) : 0u;
    const uint32_t a0 = wg + r * Sp;
    const uint32_t gapR = Sp - size;
    const uint32_t gapL = (r == 0 || !act) ? 0u : Sp - (qb + (b - 1 < rem ? 1u : 0u));
    const uint32_t seg = act ? Bs * LPB : 1u;
    // the segment's last block
    const uint32_t bl = s * Bs + Bs - 1;
    const uint32_t size_last = qb + (bl < rem ? 1u : 0u);
    const uint32_t a0_last = wg + (Bs - 1) * Sp;

    const double *xg = a.x + ch.x_off;
    if (own)
        for (uint32_t t = 0; t < size; ++t) c.R[a0 + t] = xg[start + t];
    const DevState st = a.states[ic];
    double L = st.L, U = st.U, mdrift = st.max_drift;
    uint64_t sweep = st.sweep;
    unsigned long long sp = __hip_atomic_load(&g.ctr[me], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();

    auto publish_all = [&]() {
        for (uint32_t m = tid; m < wg; m += TW) {
            g.pub_head[(uint64_t)me * W + m] = c.R[wg + m];
            g.pub_tail[(uint64_t)me * W + m] = c.R[a0_last + size_last - wg + m];
        }
    };
    auto refresh = [&]() {
        for (uint32_t m = tid; m < wg; m += TW) {
            if (hasLn) {
                const double v = seg_load(&g.pub_tail[(uint64_t)ln * W + m]);
                c.R[m] = (per && s == 0) ? v - L : v;
            }
            if (hasRn) {
                const double v = seg_load(&g.pub_head[(uint64_t)rn * W + m]);
                c.R[wg + Bs * Sp + m] = (per && s + 1 == S) ? v + L : v;
            }
        }
        __syncthreads();
    };
    auto sync_nb = [&]() {
        ++sp;
        __syncthreads();
        if (tid == 0) {
            seg_signal(&g.ctr[me], sp);
            if (hasLn) seg_wait(&g.ctr[ln], sp, g.err);
            if (hasRn) seg_wait(&g.ctr[rn], sp, g.err);
            __threadfence();
        }
        __syncthreads();
    };
    auto sync_all = [&]() {
        ++sp;
        __syncthreads();
        if (tid == 0) {
            seg_signal(&g.ctr[me], sp);
            for (uint32_t j = 0; j < S; ++j)
                if (j != s) seg_wait(&g.ctr[sg.first + j], sp, g.err);
            __threadfence();
        }
        __syncthreads();
    };
    // this segment's totals per 64 blocks of slot m of the last nd-wide
    // seg_allreduce (thread 0): wave totals, or pairs of them with two lanes
    // per block (the top level of the 64-block tree)
    auto put = [&](uint32_t par, uint32_t f, uint32_t m, uint32_t nd) {
        for (uint32_t wv = 0; wv < nwv; ++wv)
            g.part[seg_pidx(me, par, f, wv)] =
                (LPB == 2 && Bs >= 64) ? sc.rd[(2 * wv) * nd + m] + sc.rd[(2 * wv + 1) * nd + m]
                                       : sc.rd[wv * nd + m];
    };
    // Field f summed over the chain's blocks in the order of seg_allreduce
    // over B = S*Bs lanes: pairwise trees within 64-lane groups, the group
    // totals then added in sequence from zero (no sequence when B <= 64).
    // Every segment reads the same values in the same order: the same bits.
    __shared__ double tot[SEG_FIELDS];
    auto gather = [&](uint32_t par, uint32_t f) {
        if (tid < 64) {
            double acc = 0.0;
            if (Bs >= 64) {
                if (tid == 0)
                    for (uint32_t j = 0; j < S; ++j)
                        for (uint32_t wv = 0; wv < nwv; ++wv)
                            acc += seg_load(&g.part[seg_pidx(sg.first + j, par, f, wv)]);
            } else {
                const uint32_t Bt = S * Bs;
                const uint32_t gs = Bt <= 64 ? S : 64u / Bs;  // segment totals per group
                const uint32_t ng = Bt <= 64 ? 1u : Bt / 64u;
                for (uint32_t k = 0; k < ng; ++k) {
                    double v = tid < gs ? seg_load(&g.part[seg_pidx(sg.first + k * gs + tid, par, f, 0)]) : 0.0;
#pragma unroll
                    for (uint32_t off = 1; off < 64; off <<= 1) {
                        const double p = __shfl_xor(v, (int)off);
                        if (off < gs) v += p;
                    }
                    acc = Bt <= 64 ? v : acc + v;
                }
            }
            if (tid == 0) tot[f] = acc;
        }
        __syncthreads();
    };
    // integer totals (exact in double), any order
    auto gather_int = [&](uint32_t par, uint32_t f, uint32_t wv) {
        if (tid == 0) {
            double v = 0.0;
            for (uint32_t j = 0; j < S; ++j) v += seg_load(&g.part[seg_pidx(sg.first + j, par, f, wv)]);
            tot[f] = v;
        }
        __syncthreads();
    };

    publish_all();
    sync_nb();
    refresh();

    uint32_t n_tr = 0, n_wall = 0, n_core = 0, n_terms = 0;
    uint32_t acc_total = 0, n_vt = 0, n_av = 0, n_rc = 0;

#pragma unroll 1
    for (uint32_t sw = 0; sw < a.n_sweeps; ++sw, ++sweep) {
        double dUs = 0.0;
        uint32_t acc = 0;
        uint32_t w_md = 0, w_u = 0, w_next_md = 0, w_next_u = 0;
#pragma unroll 1
        for (uint32_t t = 0; t < smax; ++t) {
            if (t < size) {
                const uint32_t i = start + t;
                const uint32_t my = a0 + t;
                if ((t & 1u) == 0) {
                    uint32_t rr[4];
                    philox4x32_10(i, (uint32_t)sweep, chain_id, 0u, key0, key1, rr);
                    w_md = rr[0];
                    w_u = rr[1];
                    w_next_md = rr[2];
                    w_next_u = rr[3];
                } else {
                    w_md = w_next_md;
                    w_u = w_next_u;
                }
                const double a1 = u32o(w_md);
                const double u = u32o(w_u);
                const double md = ((a1 - 0.5) * 2.0) * max_step;
                const double xi = c.R[my];
                const double xn = xi + md;
                ++n_tr;
                double dE = 0.0;
                // one lane per block, FP32: two partners per side per walk
                // step, as the whole-chain kernel (the same sums, bit for bit)
                const bool accd = trial_lds<Real, KIND, LPB == 2, LPB == 1 && sizeof(Real) == 4>(
                    c, q, gmin, beta, betaf, i, t, size, my, gapL, gapR, L, xi, xn, md, u, dE, n_terms, n_wall,
                    n_core, side);
                if (own) {
                    const double xs = accd ? xn : xi;
                    c.R[my] = xs;
                    dUs += accd ? dE : 0.0;
                    acc += accd ? 1u : 0u;
                    if (a.acclog) a.acclog[log_off + (uint64_t)(soff + sw) * N + i] = accd ? 1 : 0;
                    // boundary particles the neighbouring segments read
                    if (r == 0 && t < wg) g.pub_head[(uint64_t)me * W + t] = xs;
                    if (r + 1 == Bs && t + wg >= size) g.pub_tail[(uint64_t)me * W + (t + wg - size)] = xs;
                }
            }
            if ((t + 1) % K == 0 || t + 1 == smax) {
                sync_nb();
                refresh();
            } else {
                __syncthreads();
            }
        }

        // box-length trial: per-lane sums over the pairs whose left member
        // is in the lane's block, across segment boundaries through the halo
        VolProp vp{};
        uint32_t bad = 0;
        double vpart = 0.0, upart = 0.0;
        if (npt) {
            vp = vol_propose(ch, L, sweep);
            vpart = vol_lane<Real, KIND, sizeof(Real) == 4>(
                c, q, gmin, vp, vp.valid && act && own, start, size, a0, gapR,
                per && r + 1 == Bs && s + 1 == S,
                [&](uint32_t m) { return seg_load(&g.pub_head[(uint64_t)sg.first * W + m]); }, bad, upart);
        }
        const uint32_t par = (uint32_t)(sweep & 1u);
        {
            double v[3] = {dUs, vpart, upart};
            uint32_t k2[2] = {acc, bad};
            seg_allreduce<TW, 3, 2>(v, k2, seg, 0u, sc.rd, sc.ri);
            if (tid == 0) {
                put(par, 0, 0, 3);
                put(par, 1, 1, 3);
                put(par, 8, 2, 3);
                g.part[seg_pidx(me, par, 2, 0)] = (double)k2[0];
                g.part[seg_pidx(me, par, 3, 0)] = (double)k2[1];
            }
        }
        sync_all();
        gather(par, 0);
        gather(par, 1);
        gather(par, 8);
        gather_int(par, 2, 0);
        gather_int(par, 3, 0);
        U = U + tot[0];
        // FP32 form: the configuration's FP32 energy from the box-length
        // pass replaces the running sum (as in the whole-chain kernel);
        // square well: the pass gave the scaled energy
        if (sizeof(Real) == 4 && KIND != POT_HRSW && npt && vp.valid && tot[3] == 0.0) U = tot[8];
        acc = (uint32_t)tot[2];
        acc_total += acc;
        uint32_t accv = 0;
        if (npt) {
            ++n_vt;
            const double dUv = (sizeof(Real) == 4 && KIND == POT_HRSW) ? tot[1] - U : tot[1];
            if (vp.valid && tot[3] == 0.0 && vol_accept(dUv, vp.dL, vp.lns, P, beta, vp.jac, vp.u)) {
                if (own)
                    for (uint32_t t = 0; t < size; ++t) c.R[a0 + t] = c.R[a0 + t] * vp.sc;
                L = vp.Lp;
                U = U + dUv;
                accv = 1;
                ++n_av;
                __syncthreads();
                publish_all();
                sync_nb();
                refresh();
            }
        }

        // drift guard (ECheck, src/jmmMCState.cpp:1965-2095), always FP64
        if (eci && ((sweep + 1) % eci == 0)) {
            double Up = 0.0;
            if (act && own) {
                for (uint32_t t = 0; t < size; ++t) {
                    const uint32_t i = start + t;
                    const uint32_t my = a0 + t;
                    const double xi = c.R[my];
                    const uint32_t kR = per ? w : min(w, N - 1 - i);
                    for (uint32_t k = 1; k <= kR; ++k) {
                        const double d = c.R[my + k + (t + k >= size ? gapR : 0)] - xi;
                        if (d > q.rc) break;
                        Up += phi64<KIND>(q, d);
                    }
                }
            }
            double v[1] = {Up};
            uint32_t k1[1] = {0};
            seg_allreduce<TW, 1, 1>(v, k1, seg, 0u, sc.rd, sc.ri);
            if (tid == 0) put(par, 4, 0, 1);
            sync_all();
            gather(par, 4);
            const double dr = fabs(tot[4] - U);
            if (dr > mdrift) mdrift = dr;
            U = tot[4];
            ++n_rc;
        }

        double Ws = 0.0, Hs = 0.0;
        if (a.sW) {
            double v[2] = {0.0, 0.0};
            if (act && own) virial_lane<KIND>(c, q, start, size, a0, gapR, v[0], v[1]);
            uint32_t k1[1] = {0};
            seg_allreduce<TW, 2, 1>(v, k1, seg, 0u, sc.rd, sc.ri);
            if (tid == 0) {
                put(par, 6, 0, 2);
                put(par, 7, 1, 2);
            }
            sync_all();
            gather(par, 6);
            gather(par, 7);
            Ws = tot[6];
            Hs = tot[7];
        }
        if (s == 0 && tid == 0) {
            const uint64_t row = (uint64_t)(soff + sw) * a.n_chains + ic;
            a.sU[row] = U;
            a.sL[row] = L;
            a.sA[row] = make_uint2(acc, accv);
            if (a.sW) {
                a.sW[row] = Ws;
                a.sH[row] = Hs;
            }
        }
    }

    // counters: segment totals, summed by segment 0, which alone writes the
    // chain's state
    {
        double vz[1] = {0.0};
        uint32_t k4[4] = {own ? n_tr : 0u, own ? n_wall : 0u, own ? n_core : 0u, n_terms};
        seg_allreduce<TW, 1, 4>(vz, k4, seg, 0u, sc.rd, sc.ri);
        const uint32_t par = (uint32_t)(sweep & 1u);
        if (tid == 0)
            for (uint32_t m = 0; m < 4; ++m) g.part[seg_pidx(me, par, 5, m)] = (double)k4[m];
        sync_all();
        if (s == 0 && tid == 0) {
            uint64_t tt[4];
            for (uint32_t m = 0; m < 4; ++m) {
                double v = 0.0;
                for (uint32_t j = 0; j < S; ++j) v += seg_load(&g.part[seg_pidx(sg.first + j, par, 5, m)]);
                tt[m] = (uint64_t)v;
            }
            DevState &o = a.states[ic];
            o.L = L;
            o.U = U;
            o.sweep = sweep;
            o.cnt[0] += tt[0];
            o.cnt[1] += acc_total;
            o.cnt[2] += tt[1];
            o.cnt[3] += tt[2];
            o.cnt[4] += tt[3];
            o.cnt[5] += n_vt;
            o.cnt[6] += n_av;
            o.cnt[7] += n_rc;
            o.max_drift = mdrift;
        }
    }
    double *xw = a.x + ch.x_off;
    if (own)
        for (uint32_t t = 0; t < size; ++t) xw[start + t] = c.R[a0 + t];
}

template <typename Real, int TW, int LPB>
__global__ void __launch_bounds__(TW) __attribute__((amdgpu_waves_per_eu(1)))
k_sweep_seg(SegArgs g)
{
    extern __shared__ double smem[];
    __shared__ BatchedScratch<TW> sc;
    const uint32_t soff = g.a.ctrl->sample_off;
    const DevSeg sg = g.segs[blockIdx.x];
    switch (g.a.chains[sg.chain].kind) {
    case POT_LJ: run_segment<Real, POT_LJ, TW, LPB>(g, sg, smem, sc, soff); break;
    case POT_HARMONIC: run_segment<Real, POT_HARMONIC, TW, LPB>(g, sg, smem, sc, soff); break;
    case POT_HRSW: run_segment<Real, POT_HRSW, TW, LPB>(g, sg, smem, sc, soff); break;
    default: run_segment<Real, POT_YUKAWA, TW, LPB>(g, sg, smem, sc, soff); break;
    }
}

// --------------------------------------------------------------- plain form

template <int KIND, bool GX>
__device__ void run_plain(const SweepArgs &a, uint32_t ic, double *smem, uint32_t soff)
{
    const uint32_t lane = threadIdx.x;
    const DevChain &ch = a.chains[ic];
    const Pot q = ch.pot;
    const int64_t N = ch.N;
    const uint32_t w = ch.w, B = ch.B, smax = ch.smax;
    const bool per = (ch.flags & F_PERIODIC) != 0;
    const bool npt = (ch.flags & F_NPT) != 0;
    const double beta = ch.beta, gmin = ch.gmin, max_step = ch.max_step;
    const uint32_t key0 = (uint32_t)ch.seed, key1 = (uint32_t)(ch.seed >> 32);
    const uint64_t log_off = ch.log_off;
    const double P = ch.P;
    const uint32_t eci = ch.eci;
    const uint32_t chain_id = ch.chain_id;
    double *X;
    if (GX) {
        X = a.x + ch.x_off;
    } else {
        for (uint32_t k = lane; k < (uint32_t)N; k += 64) smem[k] = a.x[ch.x_off + k];
        X = smem;
    }
    __syncthreads();
    const DevState st = a.states[ic];
    double L = st.L, U = st.U, mdrift = st.max_drift;
    uint64_t sweep = st.sweep;
    const uint32_t qb = (uint32_t)N / B, rem = (uint32_t)N % B;
    const uint32_t side = lane >> 5, kk = (lane & 31) + 1;
    uint64_t n_tr = 0, n_wall = 0, n_core = 0, n_terms = 0, acc_total = 0;
    uint32_t n_vt = 0, n_av = 0, n_rc = 0;

    for (uint32_t sw = 0; sw < a.n_sweeps; ++sw, ++sweep) {
        double dUs = 0.0;
        uint32_t acc = 0;
        for (uint32_t t = 0; t < smax; ++t) {
            for (uint32_t b = 0; b < B; ++b) {
                const uint32_t size = qb + (b < rem ? 1u : 0u);
                if (t >= size) continue;
                const uint32_t i = b * qb + (b < rem ? b : rem) + t;
                uint32_t rr[4];
                philox4x32_10(i - (t & 1u), (uint32_t)sweep, chain_id, 0u, key0, key1, rr);
                const double a1 = u32o((t & 1u) ? rr[2] : rr[0]);
                const double u = u32o((t & 1u) ? rr[3] : rr[1]);
                const double md = ((a1 - 0.5) * 2.0) * max_step;
                const double xi = X[i];
                const double xn = xi + md;
                ++n_tr;
                bool ok = true;
                if (!per && fabs(xn) > L / 2.0) {
                    ok = false;
                    ++n_wall;
                } else if (N >= 2) {
                    double xl, xr;
                    if ((nbpos(X, (int64_t)i - 1, N, per, L, xl) && (xn - xl) < gmin) ||
                        (nbpos(X, (int64_t)i + 1, N, per, L, xr) && (xr - xn) < gmin)) {
                        ok = false;
                        ++n_core;
                    }
                }
                bool accd = false;
                if (ok) {
                    double term = 0.0;
                    uint32_t tm = 0;
                    for (uint32_t base = 0; base < w; base += 32) {
                        const uint32_t k = base + kk;
                        if (k <= w) {
                            const int64_t j = side ? (int64_t)i + k : (int64_t)i - k;
                            double xj;
                            if (nbpos(X, j, N, per, L, xj)) {
                                const double dold = side ? xj - xi : xi - xj;
                                const double dnew = side ? xj - xn : xn - xj;
                                if (!(dold > q.rc && dnew > q.rc)) {
                                    term += phi64<KIND>(q, dnew) - phi64<KIND>(q, dold);
                                    tm += n_inside<KIND>(q, dold, dnew);
                                }
                            }
                        }
                    }
                    const double dE = wave_sum(term);
                    n_terms += wave_sum(tm);
                    accd = dE <= 0.0 || exp(-dE * beta) > u;
                    if (accd) {
                        if (lane == 0) X[i] = xn;
                        dUs += dE;
                        ++acc;
                    }
                    __syncthreads();
                }
                if (a.acclog && lane == 0)
                    a.acclog[log_off + (uint64_t)(soff + sw) * N + i] = accd ? 1 : 0;
            }
        }
        U = U + dUs;
        acc_total += acc;

        uint32_t accv = 0;
        if (npt) {
            ++n_vt;
            const VolProp vp = vol_propose(ch, L, sweep);
            if (vp.valid) {
                uint32_t bad = 0;
                double u0;
                double part = scaled_delta<double, KIND>(q, vp.bk, gmin, X, N, per, L, vp.Lp, vp.sc, w,
                                                         lane, (uint32_t)N, 64, bad, u0);
                part = wave_sum(part);
                bad = wave_sum(bad);
                if (!bad && vol_accept(part, vp.dL, vp.lns, P, beta, vp.jac, vp.u)) {
                    __syncthreads();
                    for (uint32_t p = lane; p < (uint32_t)N; p += 64) X[p] = X[p] * vp.sc;
                    L = vp.Lp;
                    U = U + part;
                    accv = 1;
                    ++n_av;
                }
                __syncthreads();
            }
        }
        if (eci && ((sweep + 1) % eci == 0)) {
            const double Uf = wave_sum(energy_part<KIND>(q, X, N, per, L, w, lane, (uint32_t)N, 64));
            const double dr = fabs(Uf - U);
            if (dr > mdrift) mdrift = dr;
            U = Uf;
            ++n_rc;
        }
        double Ws = 0.0, Hs = 0.0;
        if (a.sW) {
            virial_part<KIND>(q, X, N, per, L, w, lane, (uint32_t)N, 64, Ws, Hs);
            Ws = wave_sum(Ws);
            Hs = wave_sum(Hs);
        }
        if (lane == 0) {
            const uint64_t row = (uint64_t)(soff + sw) * a.n_chains + ic;
            a.sU[row] = U;
            a.sL[row] = L;
            a.sA[row] = make_uint2(acc, accv);
            if (a.sW) {
                a.sW[row] = Ws;
                a.sH[row] = Hs;
            }
        }
    }
    if (lane == 0) {
        DevState &o = a.states[ic];
        o.L = L;
        o.U = U;
        o.sweep = sweep;
        o.cnt[0] += n_tr;
        o.cnt[1] += acc_total;
        o.cnt[2] += n_wall;
        o.cnt[3] += n_core;
        o.cnt[4] += n_terms;
        o.cnt[5] += n_vt;
        o.cnt[6] += n_av;
        o.cnt[7] += n_rc;
        o.max_drift = mdrift;
    }
    __syncthreads();
    if (!GX)
        for (uint32_t k = lane; k < (uint32_t)N; k += 64) a.x[ch.x_off + k] = smem[k];
}

template <bool GX>
__global__ void __launch_bounds__(64) k_sweep_plain(SweepArgs a)
{
    extern __shared__ double smem[];
    const uint32_t soff = a.ctrl->sample_off;
    const uint32_t ic = a.plain_list[blockIdx.x];
    switch (a.chains[ic].kind) {
    case POT_LJ: run_plain<POT_LJ, GX>(a, ic, smem, soff); break;
    case POT_HARMONIC: run_plain<POT_HARMONIC, GX>(a, ic, smem, soff); break;
    case POT_HRSW: run_plain<POT_HRSW, GX>(a, ic, smem, soff); break;
    default: run_plain<POT_YUKAWA, GX>(a, ic, smem, soff); break;
    }
}

__global__ void k_begin(Ctrl *c, uint32_t n)
{
    c->queue = 0;
    c->sample_off = c->next_off;
    c->next_off = c->next_off + n;
}

// ------------------------------------------------------- differentiation

// One workgroup per configuration.  Pair arithmetic in Real, sums in FP64.
// obs = (U, W, dU/dtheta0, dU/dtheta1, dW/dtheta0, dW/dtheta1); grad = dU/dx.
template <typename Real, int KIND>
__device__ void run_obs(const DevChain &ch, const double *X, double L, double sc, double *obs,
                        double *grad)
{
    __shared__ double red[4 * 6];
    const uint32_t tid = threadIdx.x;
    const Pot q = ch.pot;
    const int64_t N = ch.N;
    // a compressed configuration (sc < 1) may hold gaps below the hard core:
    // walk to the cutoff (or the NBN limit) without the window limit
    const uint32_t w = sc < 1.0 ? min(ch.nbn_lim, (uint32_t)(N > 0 ? N - 1 : 0)) : ch.w;
    const bool per = (ch.flags & F_PERIODIC) != 0;
    const Real sh[3] = {Real(ch.sh[0]), Real(ch.sh[1]), Real(ch.sh[2])};
    // position j (index possibly past either end on a periodic line) of the
    // configuration scaled by sc to length L: x_j sc, images x sc -+ L
    auto pos = [&](int64_t j, double &xj) -> bool {
        if (j >= 0 && j < N) {
            xj = X[j] * sc;
            return true;
        }
        if (!per) return false;
        xj = (j < 0) ? X[j + N] * sc - L : X[j - N] * sc + L;
        return true;
    };
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t i = tid; i < (uint32_t)N; i += 256) {
        const double xi = X[i] * sc;
        double g = 0.0;
        for (uint32_t k = 1; k <= w; ++k) {
            double xj;
            if (!pos((int64_t)i + k, xj)) break;
            const double d = xj - xi;
            if (d > q.rc) break;
            // a rescaled configuration can overlap hard cores: infinite energy
            if ((KIND == POT_HRSW && d < q.p0) || (KIND == POT_YUKAWA && d < q.p2)) {
                acc[0] += INFINITY;
                continue;
            }
            if (!inside64<KIND>(q, d)) continue;
            Real o[6];
            pair_obs<KIND, Real>(q, Real(d), sh, o);
#pragma unroll
            for (int m = 0; m < 6; ++m) acc[m] += (double)o[m];
            g += (double)o[1] / d;
        }
        if (grad) {
            for (uint32_t k = 1; k <= w; ++k) {
                double xj;
                if (!pos((int64_t)i - k, xj)) break;
                const double d = xi - xj;
                if (d > q.rc) break;
                if (!inside64<KIND>(q, d)) continue;
                Real o[6];
                pair_obs<KIND, Real>(q, Real(d), sh, o);
                g -= (double)o[1] / d;
            }
            grad[i] = g;
        }
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) acc[m] = wave_sum(acc[m]);
    if ((tid & 63) == 0)
        for (int m = 0; m < 6; ++m) red[(tid >> 6) * 6 + m] = acc[m];
    __syncthreads();
    if (tid == 0)
        for (int m = 0; m < 6; ++m) obs[m] = ((red[m] + red[6 + m]) + red[12 + m]) + red[18 + m];
}

template <typename Real>
__global__ void __launch_bounds__(256) k_observables(ObsArgs a)
{
    const uint32_t c = blockIdx.x;
    const DevChain &ch = a.standalone ? a.one : a.chains[a.list[c]];
    const double *X;
    double L, *grad;
    double *obs;
    double sc = 1.0;
    if (a.standalone) {
        X = a.x + (uint64_t)c * ch.N;
        L = a.L[c];
        grad = a.grad ? a.grad + (uint64_t)c * ch.N : nullptr;
        obs = a.obs + (uint64_t)c * 6;
    } else {
        const uint32_t ic = a.list[c];
        X = a.x + ch.x_off;
        L = a.L[ic];
        if (a.Lnew) {
            sc = a.Lnew[ic] / L;
            L = a.Lnew[ic];
        }
        grad = a.grad ? a.grad + (a.goff ? a.goff[ic] : ch.x_off) : nullptr;
        obs = a.obs + (uint64_t)(a.row ? a.row[ic] : ic) * 6;
    }
    switch (ch.kind) {
    case POT_LJ: run_obs<Real, POT_LJ>(ch, X, L, sc, obs, grad); break;
    case POT_HARMONIC: run_obs<Real, POT_HARMONIC>(ch, X, L, sc, obs, grad); break;
    case POT_HRSW: run_obs<Real, POT_HRSW>(ch, X, L, sc, obs, grad); break;
    default: run_obs<Real, POT_YUKAWA>(ch, X, L, sc, obs, grad); break;
    }
}

// Copies every chain's current length out of the state array (observables
// of engine chains).
__global__ void k_lengths(const DevState *st, double *L, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) L[i] = st[i].L;
}

// Sample rows [sweep][chain] (internal chain order) to the caller's layout
// [chain][sweep] (external order), through 32 x 32 tiles in LDS so that the
// reads (along chains) and the writes (along sweeps) are both contiguous.
// Grid-stride over tiles.
__global__ void __launch_bounds__(256) k_pack_samples(const double *sU, const double *sL, const uint2 *sA,
                                                      const uint32_t *ext, uint32_t nc, uint32_t ns,
                                                      SampleRec *out)
{
    __shared__ double tU[32][33], tL[32][33];
    __shared__ uint2 tA[32][33];
    const uint32_t tx = threadIdx.x & 31u, ty = threadIdx.x >> 5;
    const uint32_t tk = (nc + 31u) / 32u, ts = (ns + 31u) / 32u;
    const uint64_t tiles = (uint64_t)tk * ts;
    for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t k0 = (uint32_t)(tile % tk) * 32u, s0 = (uint32_t)(tile / tk) * 32u;
        for (uint32_t j = ty; j < 32u; j += 8u) {
            const uint32_t sw = s0 + j, k = k0 + tx;
            if (sw < ns && k < nc) {
                const uint64_t r = (uint64_t)sw * nc + k;
                tU[j][tx] = sU[r];
                tL[j][tx] = sL[r];
                tA[j][tx] = sA[r];
            }
        }
        __syncthreads();
        for (uint32_t j = ty; j < 32u; j += 8u) {
            const uint32_t k = k0 + j, sw = s0 + tx;
            if (k < nc && sw < ns) {
                SampleRec o;
                o.U = tU[tx][j];
                o.L = tL[tx][j];
                o.acc_d = tA[tx][j].x;
                o.acc_v = tA[tx][j].y;
                out[(uint64_t)ext[k] * ns + sw] = o;
            }
        }
        __syncthreads();
    }
}

// Virial rows [sweep][chain] to out[(ext * ns + sweep) * 2 + {0, 1}], through
// 32 x 32 LDS tiles as k_pack_samples.
__global__ void __launch_bounds__(256) k_pack_virials(const double *sW, const double *sH, const uint32_t *ext,
                                                      uint32_t nc, uint32_t ns, double *out)
{
    __shared__ double tW[32][33], tH[32][33];
    const uint32_t tx = threadIdx.x & 31u, ty = threadIdx.x >> 5;
    const uint32_t tk = (nc + 31u) / 32u, ts = (ns + 31u) / 32u;
    const uint64_t tiles = (uint64_t)tk * ts;
    for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t k0 = (uint32_t)(tile % tk) * 32u, s0 = (uint32_t)(tile / tk) * 32u;
        for (uint32_t j = ty; j < 32u; j += 8u) {
            const uint32_t sw = s0 + j, k = k0 + tx;
            if (sw < ns && k < nc) {
                const uint64_t r = (uint64_t)sw * nc + k;
                tW[j][tx] = sW[r];
                tH[j][tx] = sH[r];
            }
        }
        __syncthreads();
        for (uint32_t j = ty; j < 32u; j += 8u) {
            const uint32_t k = k0 + j, sw = s0 + tx;
            if (k < nc && sw < ns) {
                const uint64_t o = ((uint64_t)ext[k] * ns + sw) * 2;
                out[o] = tW[tx][j];
                out[o + 1] = tH[tx][j];
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ histograms

// Position folded into [-L/2, L/2) on a periodic line; unchanged with walls.
__device__ __forceinline__ double fold(double x, double L, bool per)
{
    return per ? x - L * floor((x + L / 2.0) / L) : x;
}

__device__ __forceinline__ int64_t bin_of(double x, double width, uint32_t nb)
{
    return (int64_t)floor(x / width + nb / 2.0);
}

// One sample of the reference's density and pair-separation histograms
// (fgrho, src/jmmMCState.cpp:1069-1128): rho bin floor(x/rbw + rhonb/2) of
// every particle; for every pair closer than gnb*gbw, separation bin
// floor(d/gbw) counted once in the segment floor(x/gsw + gns/2) of each
// member.  Periodic chains: positions folded into [-L/2, L/2), each pair
// once at its shorter separation (below L/2).  Grid: chains x parts, the
// parts of a chain stride over its particles.
__global__ void __launch_bounds__(256) k_hist(HistArgs a)
{
    const uint32_t c = blockIdx.x / a.parts, part = blockIdx.x % a.parts;
    const DevChain &ch = a.chains[c];
    const uint32_t N = ch.N;
    const bool per = (ch.flags & F_PERIODIC) != 0;
    const double L = a.states[c].L;
    const double *X = a.x + ch.x_off;
    const double gmax = a.gnb * a.gbw;
    unsigned long long *rho = a.rho + (uint64_t)c * a.rhonb;
    unsigned long long *g = a.g + (uint64_t)c * a.gns * a.gnb;
    for (uint32_t i = part * 256 + threadIdx.x; i < N; i += 256 * a.parts) {
        const double xi = X[i];
        const double fi = fold(xi, L, per);
        const int64_t rb = bin_of(fi, a.rbw, a.rhonb);
        if (rb >= 0 && rb < (int64_t)a.rhonb) atomicAdd(&rho[rb], 1ull);
        if (!a.gns || !a.gnb) continue;
        const int64_t si = bin_of(fi, a.gsw, a.gns);
        for (uint32_t k = 1; k < N; ++k) {
            const uint32_t j0 = i + k;
            double d, xj;
            if (j0 < N) {
                xj = X[j0];
                d = xj - xi;
            } else {
                if (!per) break;
                xj = X[j0 - N];
                d = (xj + L) - xi;
            }
            if (!(d < gmax) || (per && !(d < L / 2.0))) break;
            const int64_t gb = (int64_t)floor(d / a.gbw);
            if (gb < 0 || gb >= (int64_t)a.gnb) continue;
            const int64_t sj = bin_of(fold(xj, L, per), a.gsw, a.gns);
            if (si >= 0 && si < (int64_t)a.gns) atomicAdd(&g[(uint64_t)si * a.gnb + gb], 1ull);
            if (sj >= 0 && sj < (int64_t)a.gns) atomicAdd(&g[(uint64_t)sj * a.gnb + gb], 1ull);
        }
    }
}

// Positions of every chain scaled to a new length: x -> x Lnew/L (moveVolume,
// src/jmmMCState.cpp:2831-2916).  The length itself is set by the host.
__global__ void __launch_bounds__(256) k_rescale(const DevChain *chains, const DevState *st,
                                                 double *x, const double *Lnew, uint32_t parts)
{
    const uint32_t c = blockIdx.x / parts, part = blockIdx.x % parts;
    const DevChain &ch = chains[c];
    const double sc = Lnew[c] / st[c].L;
    double *X = x + ch.x_off;
    for (uint32_t i = part * 256 + threadIdx.x; i < ch.N; i += 256 * parts) X[i] = X[i] * sc;
}

// ------------------------------------------------------------ launchers

hipError_t launch_begin(Ctrl *c, uint32_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_begin, dim3(1), dim3(1), 0, s, c, n);
    return hipGetLastError();
}

template <typename Real>
static const void *pick_shape(int T, int gx)
{
    if (T == 1024) return gx ? (const void *)k_sweep_batched<Real, 1024, true>
                             : (const void *)k_sweep_batched<Real, 1024, false>;
    if (T == 512) return gx ? (const void *)k_sweep_batched<Real, 512, true>
                            : (const void *)k_sweep_batched<Real, 512, false>;
    return gx ? (const void *)k_sweep_batched<Real, 256, true>
              : (const void *)k_sweep_batched<Real, 256, false>;
}

const void *batched_kernel(int fp32, int T, int gx)
{
    return fp32 ? pick_shape<float>(T, gx) : pick_shape<double>(T, gx);
}

hipError_t launch_batched(const SweepArgs &a, int fp32, int T, int gx, uint32_t grid, size_t lds,
                          hipStream_t s)
{
    const void *fn = batched_kernel(fp32, T, gx);
    void *args[] = {(void *)&a};
    return hipLaunchKernel(fn, dim3(grid), dim3(T), args, lds, s);
}

const void *plain_kernel(int gx)
{
    return gx ? (const void *)k_sweep_plain<true> : (const void *)k_sweep_plain<false>;
}

hipError_t launch_plain(const SweepArgs &a, int gx, uint32_t grid, size_t lds, hipStream_t s)
{
    if (gx) hipLaunchKernelGGL(k_sweep_plain<true>, dim3(grid), dim3(64), lds, s, a);
    else hipLaunchKernelGGL(k_sweep_plain<false>, dim3(grid), dim3(64), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_observables(const ObsArgs &a, int fp32, uint32_t grid, hipStream_t s)
{
    if (fp32) hipLaunchKernelGGL(k_observables<float>, dim3(grid), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_observables<double>, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_rescale(const DevChain *chains, const DevState *st, double *x, const double *Lnew,
                          uint32_t n_chains, uint32_t parts, hipStream_t s)
{
    hipLaunchKernelGGL(k_rescale, dim3(n_chains * parts), dim3(256), 0, s, chains, st, x, Lnew, parts);
    return hipGetLastError();
}

const void *segmented_kernel(int fp32, int tw, int lpb)
{
    if (lpb == 2 && tw == 512)
        return fp32 ? (const void *)k_sweep_seg<float, 512, 2> : (const void *)k_sweep_seg<double, 512, 2>;
    if (lpb == 2)
        return fp32 ? (const void *)k_sweep_seg<float, 256, 2> : (const void *)k_sweep_seg<double, 256, 2>;
    return fp32 ? (const void *)k_sweep_seg<float, 256, 1> : (const void *)k_sweep_seg<double, 256, 1>;
}

// An ordinary launch: the segments' synchronisation needs every workgroup
// resident at once, which jmm_create checks against the occupancy of the
// whole device (the launch follows the stream's previous work, so the
// device is free when it starts), and a wait that never completes ends
// after its bounded polls with the error flag set.  A cooperative launch
// would add only a runtime check of the same residency, and a rocprofv3
// kernel trace of a program using one ended in a segmentation fault with
// no results written.
hipError_t launch_segmented(const SegArgs &a, int fp32, int tw, int lpb, size_t lds, hipStream_t s)
{
    void *args[] = {(void *)&a};
    return hipLaunchKernel(segmented_kernel(fp32, tw, lpb), dim3(a.n_segs), dim3(tw), args, lds, s);
}

hipError_t launch_hist(const HistArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_hist, dim3(a.n_chains * a.parts), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_lengths(const DevState *st, double *L, uint32_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_lengths, dim3((n + 255) / 256), dim3(256), 0, s, st, L, n);
    return hipGetLastError();
}

hipError_t launch_pack_virials(const double *sW, const double *sH, const uint32_t *ext, uint32_t n_chains,
                               uint32_t n_sweeps, double *out, hipStream_t s)
{
    const uint64_t tiles = (uint64_t)((n_chains + 31u) / 32u) * ((n_sweeps + 31u) / 32u);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(tiles, 1u << 16);
    if (!grid) return hipSuccess;
    hipLaunchKernelGGL(k_pack_virials, dim3(grid), dim3(256), 0, s, sW, sH, ext, n_chains, n_sweeps, out);
    return hipGetLastError();
}

hipError_t launch_pack_samples(const double *sU, const double *sL, const uint2 *sA, const uint32_t *ext,
                               uint32_t n_chains, uint32_t n_sweeps, SampleRec *out, hipStream_t s)
{
    const uint64_t tiles = (uint64_t)((n_chains + 31u) / 32u) * ((n_sweeps + 31u) / 32u);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(tiles, 1u << 16);
    if (!grid) return hipSuccess;
    hipLaunchKernelGGL(k_pack_samples, dim3(grid), dim3(256), 0, s, sU, sL, sA, ext, n_chains, n_sweeps, out);
    return hipGetLastError();
}

}  // namespace jmm
