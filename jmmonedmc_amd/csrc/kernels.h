// kernels.h -- device data layout and the gfx950 kernels of the engine.
//
// HBM layout (one engine):
//   x        double[sum N]     chain positions, sorted, chains in bundle order
//   chains   DevChain[n]       read-only per-chain constants
//   states   DevState[n]       L, U, sweep counter, cumulative counters
//   sU, sL   double[cap][n]    per-sweep samples, one row per sweep (coalesced)
//   sA       uint2[cap][n]     accepted displacements / box-length trials
//   sW, sH   double[cap][n]    virial and hypervirial after each sweep (optional)
//   bundles  DevBundle[m]      work items: chains packed into one workgroup
//   slots    DevSlot[...]      chain -> lane range and LDS offset in a bundle
//
// Sweep schedule (shared with the oracle): chain c's N particles form B_c
// blocks (B_c a power of two or whole waves, every block at least M_min
// particles); in phase t every block updates its t-th particle.  Trials in
// one phase are at least M_min apart in index order, and M_min is chosen so
// that no two of them can interact or read each other's particle: every
// neighbour gap stays at least the hard core g_min (moves that would break
// it are rejected), so a partner within the cutoff rc is at most
// w = floor(rc/g_min) places away, and M_min = w + 1 (or NBN + 1, and at
// least 2 because the hard-core test reads both nearest partners).  A phase
// is therefore equivalent to running its trials one after another in any
// order.
//
// LDS layout of a chain inside a bundle (batched forms): ghost-left (wg),
// B blocks at an odd stride Sp, ghost-right (wg); see LdsChain.
#pragma once

#include "jmm_device.h"

namespace jmm {

enum : uint32_t { F_PERIODIC = 1u, F_NPT = 2u, F_VOLLOG = 4u };

struct DevChain {
    uint64_t seed;
    uint64_t x_off;   // first position of this chain in x[]
    uint64_t log_off; // first byte of this chain's accept log
    uint32_t N, B, w, smax;
    uint32_t chain_id, eci, flags, kind;
    uint32_t wg, Sp;  // LDS layout: ghost width max(w, 1), odd block stride >= smax
    uint32_t lds_len; // LDS doubles of the chain's region: 2 wg + B Sp
    uint32_t nbn_lim;  // NBN partner limit (UINT32_MAX: none)
    double beta, P, max_step, max_dv, gmin, Lmin;
    double sh[3];  // shifted form at the cutoff: E, dE/dtheta0, dE/dtheta1
    Pot pot;
    float betaf, pad[3];
};

struct DevState {
    double L, U;
    uint64_t sweep;
    uint64_t cnt[8];  // trials, acc_d, wall, core, pair_terms, vol_trials, acc_v, recomputes
    double max_drift;
};

struct DevBundle {
    uint32_t slot0, nslots, lanes, smax;
    uint32_t lds_doubles, kind, global_x, pad;
    uint64_t x_off;
};

struct DevSlot {
    uint32_t chain, lane_off, B, lds_off;
};

struct Ctrl {
    uint32_t queue, sample_off, next_off, pad;
};

struct SweepArgs {
    const DevChain *chains;
    DevState *states;
    double *x;
    const DevBundle *bundles;
    const DevSlot *slots;
    const uint32_t *plain_list;  // plain form: chains of this launch
    Ctrl *ctrl;
    double *sU, *sL;
    uint2 *sA;
    double *sW, *sH;  // nullable: per-sweep virial and hypervirial rows
    uint8_t *acclog;
    uint32_t n_items, n_sweeps, n_chains, cap;
};

struct ObsArgs {
    const DevChain *chains;  // engine form (nullable)
    DevChain one;            // standalone form
    const double *x;
    const double *L;
    const uint32_t *list;    // engine form: chains of this launch
    const double *Lnew;      // engine form, nullable: evaluate at x * Lnew/L, length Lnew
    double *obs;
    double *grad;
    const uint32_t *row;     // engine form, nullable: output row of each chain (else its index)
    const uint64_t *goff;    // engine form, nullable: gradient offset of each chain (else x_off)
    uint32_t n;
    uint32_t standalone;
};

// A chain too large for one workgroup's LDS is split into S segments of Bs
// consecutive blocks, one workgroup each (one launch, all resident).
struct DevSeg {
    uint32_t chain;  // internal chain index
    uint32_t s, S;   // segment index, segments of the chain
    uint32_t Bs;     // blocks (lanes) in each segment
    uint32_t first;  // index of the chain's segment 0 in the per-segment arrays
    uint32_t pad[3];
};

// Per-segment partial sums: [segment][sweep parity][field][wave of the
// segment's workgroup], fields dU, box-length dU, accepts, overlaps,
// recomputed U, final counters, virial, hypervirial, the FP32 form's
// configuration energy from the box-length pass.
constexpr uint32_t SEG_FIELDS = 9, SEG_WAVES = 4;

struct SegArgs {
    SweepArgs a;
    const DevSeg *segs;
    double *pub_head, *pub_tail;  // [segment][wgmax]: first / last wg positions
    unsigned long long *ctr;      // [segment]: synchronisation points passed
    double *part;                 // [segment][2][SEG_FIELDS][SEG_WAVES]
    uint32_t *err;                // set when a wait times out
    uint32_t wgmax, n_segs;
};

// rho(x) and g(x) histograms of every chain's current configuration.
struct HistArgs {
    const DevChain *chains;
    const DevState *states;
    const double *x;
    unsigned long long *rho;  // [chain][rhonb]
    unsigned long long *g;    // [chain][gns][gnb]
    double rbw, gsw, gbw;
    uint32_t n_chains, rhonb, gns, gnb, parts;
};

}  // namespace jmm

namespace jmm {
hipError_t launch_begin(Ctrl *c, uint32_t n, hipStream_t s);
const void *batched_kernel(int fp32, int T, int gx);
hipError_t launch_batched(const SweepArgs &a, int fp32, int T, int gx, uint32_t grid, size_t lds,
                          hipStream_t s);
const void *plain_kernel(int gx);
hipError_t launch_plain(const SweepArgs &a, int gx, uint32_t grid, size_t lds, hipStream_t s);
hipError_t launch_observables(const ObsArgs &a, int fp32, uint32_t grid, hipStream_t s);
hipError_t launch_lengths(const DevState *st, double *L, uint32_t n, hipStream_t s);
// Samples in the caller's layout (jmm_sample): out[ext[k] * n_sweeps + s]
// from the engine's [sweep][chain] arrays.
struct SampleRec {
    double U, L;
    uint32_t acc_d, acc_v;
};
static_assert(sizeof(SampleRec) == 24, "SampleRec must match jmm_sample");
hipError_t launch_pack_samples(const double *sU, const double *sL, const uint2 *sA, const uint32_t *ext,
                               uint32_t n_chains, uint32_t n_sweeps, SampleRec *out, hipStream_t s);
// Virial rows to the caller's layout: out[(ext[k] * n_sweeps + s) * 2 + {0, 1}]
// = (W, HV) of sweep s.
hipError_t launch_pack_virials(const double *sW, const double *sH, const uint32_t *ext, uint32_t n_chains,
                               uint32_t n_sweeps, double *out, hipStream_t s);
hipError_t launch_hist(const HistArgs &a, hipStream_t s);
// tw: workgroup width (256 or 512); lpb: lanes per block (1 or 2; 512 only with 2)
const void *segmented_kernel(int fp32, int tw, int lpb);
hipError_t launch_segmented(const SegArgs &a, int fp32, int tw, int lpb, size_t lds, hipStream_t s);
hipError_t launch_rescale(const DevChain *chains, const DevState *st, double *x, const double *Lnew,
                          uint32_t n_chains, uint32_t parts, hipStream_t s);
}  // namespace jmm
