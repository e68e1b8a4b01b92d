// engine.hip -- host side of the engine: chain validation, the trial
// schedule, bundle packing, device buffers, sweep-block launches (optionally
// replayed from a captured hipGraph), sample read-back and the C ABI declared
// in include/jmm_engine.h.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <limits>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "jmm_engine.h"
#include "kernels.h"

using namespace jmm;

static int refresh_energies(jmm_engine *e, int only);

// Staging of jmm_read_samples_noalloc: the chain-order map and the
// transposed sample records, in one file-scope device array of
// SCRATCH_BYTES.  Every GPU the library runs on holds its own copy until
// the process ends; all calls on one GPU share it (calls that overlap on
// one GPU race).  In host code the array's name is only a stand-in for the
// symbol: its device address is taken per device with hipGetSymbolAddress
// at engine creation (that device current) and used from then on.
constexpr size_t SCRATCH_BYTES = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // device address of `scratch`, by device ordinal
}  // namespace

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail(JMM_EHIP, "%s: %s", #x, hipGetErrorString(e_));  \
    } while (0)

double gmin_of(const jmm_chain_spec &s)
{
    switch (s.potential) {
    case JMM_POT_LJ: return s.pot_p[2] > 0.0 ? s.pot_p[2] : 0.5 * s.pot_p[1];
    case JMM_POT_HRSW: return s.pot_p[0];
    case JMM_POT_YUKAWA: return s.pot_p[2];
    default: return std::numeric_limits<double>::denorm_min();  // harmonic: gap <= 0
    }
}

// The trial schedule (also restated by the oracle).  Every state keeps each
// gap between neighbours (and across the periodic wrap) at least g_min: the
// lattice and set_state are checked, a displacement that would bring its
// particle within g_min of either neighbour is rejected before any energy
// work, and a length move that would do so is rejected.  So partner k of a
// particle is at least k g_min away, from its old position and from any new
// position that passes the hard-core test (the neighbour's gap adds to the
// k - 1 gaps beyond it); the same holds for the scaled configuration of a
// length move that is not rejected for an overlap.  w = largest index
// distance of a partner that can lie within rc: floor(rc / g_min), taken
// with a relative guard of 1e-3 so that partner w + 1 is beyond rc by at
// least 1e-3 g_min, far more than the rounding of FP64 separations (or NBN,
// or N - 1).  M_min = w + 1 (at least 2, as the hard-core test reads both
// nearest partners): two trials at least M_min indices apart read none of
// each other's particles and, in every order, stay beyond rc of each other,
// so they commute.  B = power-of-two number of blocks with every block at
// least M_min long, capped by bcap; a bcap that is not a power of two (a
// multiple of 64) allows B to be any multiple of 64 up to it (whole waves
// per chain), or a power of two below 64.
void schedule_of(const jmm_chain_spec &s, uint64_t &B, uint64_t &M, uint64_t &w)
{
    const uint64_t N = s.N;
    const uint64_t big = N > 0 ? N : 1;
    uint64_t wc = big, mc = big + 1;
    const double g = gmin_of(s);
    if (std::isfinite(s.cutoff) && g > 1e-300) {
        const double a = std::floor(s.cutoff / g + 1e-3);
        if (a < (double)big) wc = (uint64_t)a;
        if (a + 1.0 < (double)(big + 1)) mc = (uint64_t)a + 1;
    }
    if (s.nbn >= 0) {
        if ((uint64_t)s.nbn < wc) wc = (uint64_t)s.nbn;
        if ((uint64_t)s.nbn + 1 < mc) mc = (uint64_t)s.nbn + 1;
    }
    if (mc < 2) mc = 2;  // the hard-core test reads both nearest partners
    if (N >= 1 && wc > N - 1) wc = N - 1;
    uint64_t nb = 1;
    if (mc >= 1 && mc <= N) {
        uint64_t q = N / mc;
        const uint64_t cap = s.bcap ? s.bcap : 256;
        if (q > cap) q = cap;
        if ((cap & (cap - 1)) != 0 && q >= 64) nb = q / 64 * 64;  // whole waves
        else while (nb * 2 <= q) nb *= 2;
    }
    B = nb;
    M = mc;
    w = wc;
}

int validate(const jmm_chain_spec &s, int idx)
{
    if (s.N < 1 || s.N > (1ull << 30)) return fail(JMM_EINVAL, "chain %d: N=%llu out of range", idx, (unsigned long long)s.N);
    if (!(s.T > 0.0)) return fail(JMM_EINVAL, "chain %d: T must be > 0", idx);
    if (!(s.max_step >= 0.0) || !std::isfinite(s.max_step)) return fail(JMM_EINVAL, "chain %d: bad max_step", idx);
    if (s.potential < 0 || s.potential > 3) return fail(JMM_EINVAL, "chain %d: unknown potential %d", idx, s.potential);
    if (s.boundary != JMM_WALLS && s.boundary != JMM_PERIODIC) return fail(JMM_EINVAL, "chain %d: bad boundary", idx);
    if (s.ensemble != JMM_NLT && s.ensemble != JMM_NPT) return fail(JMM_EINVAL, "chain %d: bad ensemble", idx);
    if (s.vol_move != JMM_VOL_LINEAR && s.vol_move != JMM_VOL_LOG) return fail(JMM_EINVAL, "chain %d: bad vol_move", idx);
    if (!(s.cutoff > 0.0)) return fail(JMM_EINVAL, "chain %d: cutoff must be > 0", idx);
    if (s.bcap > 1024 || ((s.bcap & (s.bcap - 1)) != 0 && s.bcap % 64 != 0))
        return fail(JMM_EINVAL, "chain %d: bcap must be a power of two or a multiple of 64, at most 1024", idx);
    if (!(s.L0 > 0.0) || !std::isfinite(s.L0)) return fail(JMM_EINVAL, "chain %d: L0 must be finite and > 0", idx);
    if (s.potential == JMM_POT_LJ && !(s.pot_p[1] > 0.0)) return fail(JMM_EINVAL, "chain %d: LJ sigma must be > 0", idx);
    if (s.potential == JMM_POT_HRSW) {
        if (!(s.pot_p[0] > 0.0) || !(s.pot_p[2] >= s.pot_p[0]))
            return fail(JMM_EINVAL, "chain %d: hard-rod/square-well needs 0 < rod <= well edge", idx);
        if (s.cutoff != s.pot_p[2]) return fail(JMM_EINVAL, "chain %d: hard-rod/square-well cutoff must equal the well edge", idx);
    }
    if (s.potential == JMM_POT_YUKAWA && !(s.pot_p[2] > 0.0)) return fail(JMM_EINVAL, "chain %d: Yukawa hard core must be > 0", idx);
    if (s.potential == JMM_POT_LJ && (s.pot_p[2] < 0.0 || s.pot_p[2] >= s.cutoff))
        return fail(JMM_EINVAL, "chain %d: LJ exclusion distance must be 0 (0.5 sigma) or inside the cutoff", idx);
    if (s.boundary == JMM_PERIODIC) {
        if (!std::isfinite(s.cutoff)) return fail(JMM_EINVAL, "chain %d: periodic chains need a finite cutoff", idx);
        if (!(s.L0 > 2.0 * (s.cutoff + s.max_step)))
            return fail(JMM_EINVAL, "chain %d: periodic length must exceed 2 (cutoff + max_step)", idx);
    }
    if (s.N >= 2 && (s.L0 / (double)s.N) < gmin_of(s))
        return fail(JMM_EINVAL, "chain %d: lattice spacing L0/N is inside the hard core", idx);
    if (s.ensemble == JMM_NPT && !(s.max_dv >= 0.0)) return fail(JMM_EINVAL, "chain %d: bad max_dv", idx);
    return 0;
}

Pot make_pot(const jmm_chain_spec &s)
{
    Pot q{};
    q.rc = s.cutoff;
    q.p0 = s.pot_p[0];
    q.p1 = s.pot_p[1];
    q.p2 = s.pot_p[2];
    const double sig3 = s.pot_p[1] * s.pot_p[1] * s.pot_p[1];
    q.sig6 = sig3 * sig3;
    q.e4 = 4.0 * s.pot_p[0];
    q.shift = 0.0;
    if (s.shifted && std::isfinite(s.cutoff)) {
        switch (s.potential) {
        case JMM_POT_LJ: q.shift = phi64<POT_LJ>(q, s.cutoff); break;
        case JMM_POT_YUKAWA: q.shift = phi64<POT_YUKAWA>(q, s.cutoff); break;
        default: break;  // harmonic / square well: no shifted form
        }
    }
    q.rcf = (float)s.cutoff;
    q.shiftf = (float)q.shift;
    q.e4f = (float)q.e4;
    q.sig6f = (float)q.sig6;
    q.a12f = (float)(q.e4 * (q.sig6 * q.sig6));
    q.a6f = (float)(q.e4 * q.sig6);
    q.p0f = (float)q.p0;
    q.p1f = (float)q.p1;
    q.p2f = (float)q.p2;
    return q;
}

// Shifted-form values at the cutoff for the observables (E, dE/dtheta0,
// dE/dtheta1), computed with the unshifted pair function.
void shift_obs(const jmm_chain_spec &s, const Pot &q, double sh[3])
{
    sh[0] = sh[1] = sh[2] = 0.0;
    if (!s.shifted || !std::isfinite(s.cutoff)) return;
    Pot q0 = q;
    q0.shift = 0.0;
    const double z[3] = {0, 0, 0};
    double o[6];
    if (s.potential == JMM_POT_LJ) pair_obs<POT_LJ, double>(q0, s.cutoff, z, o);
    else if (s.potential == JMM_POT_YUKAWA) pair_obs<POT_YUKAWA, double>(q0, s.cutoff, z, o);
    else return;
    sh[0] = o[0];
    sh[1] = o[2];
    sh[2] = o[3];
}

DevChain make_chain(const jmm_chain_spec &s)
{
    DevChain c{};
    uint64_t B, M, w;
    schedule_of(s, B, M, w);
    c.seed = s.seed;
    c.N = (uint32_t)s.N;
    c.B = (uint32_t)B;
    c.w = (uint32_t)w;
    c.nbn_lim = s.nbn >= 0 ? (uint32_t)s.nbn : UINT32_MAX;
    c.smax = (uint32_t)((s.N + B - 1) / B);
    c.wg = c.w > 0 ? c.w : 1;
    c.Sp = c.smax | 1u;
    c.lds_len = 2 * c.wg + c.B * c.Sp;
    c.chain_id = s.chain_id;
    c.eci = s.eci;
    c.flags = (s.boundary == JMM_PERIODIC ? F_PERIODIC : 0u) | (s.ensemble == JMM_NPT ? F_NPT : 0u) |
              (s.vol_move == JMM_VOL_LOG ? F_VOLLOG : 0u);
    c.kind = (uint32_t)s.potential;
    c.beta = 1.0 / s.T;
    c.betaf = (float)c.beta;
    c.P = s.P;
    c.max_step = s.max_step;
    c.max_dv = s.max_dv;
    c.gmin = gmin_of(s);
    c.Lmin = s.boundary == JMM_PERIODIC ? 2.0 * (s.cutoff + s.max_step) : 0.0;
    c.pot = make_pot(s);
    shift_obs(s, c.pot, c.sh);
    return c;
}

// The device address of `scratch` on the current device, taken once per
// device.
int scratch_address(int device, void **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p)
            return fail(JMM_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(err));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return 0;
}

template <typename T>
int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) return fail(JMM_ENOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
    return 0;
}

}  // namespace

struct Group {
    int gx = 0;
    int kind = 0;
    uint32_t n_items = 0;     // bundles (batched) or chains (plain)
    uint32_t grid = 0;
    size_t lds = 0;
    DevBundle *d_bundles = nullptr;
    uint32_t *d_list = nullptr;
};

struct jmm_engine {
    int n = 0, form = 0, device = 0, T = 256;
    hipStream_t stream = nullptr;
    std::vector<jmm_chain_spec> specs;
    std::vector<DevChain> hch;          // internal order
    std::vector<uint32_t> int_of_ext, ext_of_int;
    std::vector<uint64_t> ext_off;      // external chain -> offset in an ext-ordered flat array
    uint64_t total_N = 0;
    DevChain *d_chains = nullptr;
    DevState *d_states = nullptr;
    double *d_x = nullptr;
    DevSlot *d_slots = nullptr;
    Ctrl *d_ctrl = nullptr;
    double *d_sU = nullptr, *d_sL = nullptr;
    uint2 *d_sA = nullptr;
    double *d_sW = nullptr, *d_sH = nullptr;  // virial rows (opts.sample_virial)
    SampleRec *d_packed = nullptr;  // samples in the caller's layout (first read allocates)
    double *d_packed_v = nullptr;   // virials in the caller's layout (first read allocates)
    unsigned char *scratch = nullptr;  // device address of this device's `scratch` array
    uint32_t *d_ext = nullptr;      // external index of each internal chain
    uint64_t *d_extoff = nullptr;   // external position offset of each internal chain
    uint8_t *d_log = nullptr;
    double *d_obs = nullptr, *d_grad = nullptr, *d_Ltmp = nullptr, *d_Lnew = nullptr;
    uint32_t *d_all = nullptr;
    uint32_t cap = 1024, write_pos = 0, block_sweeps = 100;
    std::vector<Group> grp;
    int use_graph = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    // HIP events around the sweep-block launches not yet folded into the
    // totals below (kept at most EV_KEEP pairs: older ones are read and
    // released as the queue advances)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    uint64_t ev_sweeps = 0, ev_done = 0;
    double ev_ms_done = 0.0;
    uint64_t sweeps_run = 0;            // sweeps launched since creation
    // chains split over several workgroups (segments)
    std::vector<jmm_layout> layout;  // external index
    uint32_t n_segs = 0, seg_wgmax = 1, seg_ctrl = 0, n_ctrl = 0;
    size_t seg_lds_bytes = 0;
    int seg_tw = 256, seg_lpb = 1;  // split chains: workgroup width, lanes per block
    DevSeg *d_segs = nullptr;
    double *d_pub_head = nullptr, *d_pub_tail = nullptr, *d_part = nullptr;
    unsigned long long *d_ctr = nullptr;
    uint32_t *d_seg_err = nullptr;
    std::vector<double> step_cap;       // max_step at creation, per internal chain
    // rho(x)/g(x) sampling (jmm_hist_configure)
    uint32_t hist_every = 0;
    jmm_hist_spec hs{};
    unsigned long long *d_rho = nullptr, *d_g = nullptr;
    uint64_t hist_samples = 0;
};

namespace {

constexpr size_t EV_KEEP = 64;

// Folds the elapsed time of the oldest event pairs into the engine's totals
// and releases them: every pair whose end event has completed, and, while
// more than EV_KEEP remain, the oldest after waiting for it.
int fold_events(jmm_engine *e, bool all)
{
    size_t k = 0;
    for (; k < e->ev.size(); ++k) {
        auto &p = e->ev[k];
        hipError_t q = hipEventQuery(p.second);
        if (q == hipErrorNotReady) {
            if (!all && e->ev.size() - k <= EV_KEEP) break;
            q = hipEventSynchronize(p.second);
        }
        if (q != hipSuccess) return fail(JMM_EHIP, "kernel timer event: %s", hipGetErrorString(q));
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, p.first, p.second));
        e->ev_ms_done += t;
        e->ev_done++;
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    e->ev.erase(e->ev.begin(), e->ev.begin() + (std::ptrdiff_t)k);
    return 0;
}

int launch_block(jmm_engine *e, uint32_t m)
{
    for (size_t g = 0; g < e->grp.size(); ++g) {
        Group &G = e->grp[g];
        if (!G.n_items) continue;
        HIPCHK(launch_begin(&e->d_ctrl[g], m, e->stream));
        SweepArgs a{};
        a.chains = e->d_chains;
        a.states = e->d_states;
        a.x = e->d_x;
        a.bundles = G.d_bundles;
        a.slots = e->d_slots;
        a.plain_list = G.d_list;
        a.ctrl = &e->d_ctrl[g];
        a.sU = e->d_sU;
        a.sL = e->d_sL;
        a.sA = e->d_sA;
        a.sW = e->d_sW;
        a.sH = e->d_sH;
        a.acclog = e->d_log;
        a.n_items = G.n_items;
        a.n_sweeps = m;
        a.n_chains = (uint32_t)e->n;
        a.cap = e->cap;
        if (e->form == JMM_FORM_PLAIN_FP64) {
            HIPCHK(launch_plain(a, G.gx, G.grid, G.lds, e->stream));
        } else {
            HIPCHK(launch_batched(a, e->form == JMM_FORM_BATCHED_FP32, e->T, G.gx, G.grid, G.lds,
                                  e->stream));
        }
    }
    if (e->n_segs) {
        HIPCHK(launch_begin(&e->d_ctrl[e->seg_ctrl], m, e->stream));
        SegArgs g{};
        g.a.chains = e->d_chains;
        g.a.states = e->d_states;
        g.a.x = e->d_x;
        g.a.ctrl = &e->d_ctrl[e->seg_ctrl];
        g.a.sU = e->d_sU;
        g.a.sL = e->d_sL;
        g.a.sA = e->d_sA;
        g.a.sW = e->d_sW;
        g.a.sH = e->d_sH;
        g.a.acclog = e->d_log;
        g.a.n_sweeps = m;
        g.a.n_chains = (uint32_t)e->n;
        g.a.cap = e->cap;
        g.segs = e->d_segs;
        g.pub_head = e->d_pub_head;
        g.pub_tail = e->d_pub_tail;
        g.ctr = e->d_ctr;
        g.part = e->d_part;
        g.err = e->d_seg_err;
        g.wgmax = e->seg_wgmax;
        g.n_segs = e->n_segs;
        HIPCHK(launch_segmented(g, e->form == JMM_FORM_BATCHED_FP32, e->seg_tw, e->seg_lpb,
                                e->seg_lds_bytes, e->stream));
    }
    return 0;
}

// Split chains of the segmented path wait for each other with bounded
// polls; a wait that ran out sets this flag.
int check_segments(jmm_engine *e)
{
    if (!e->n_segs) return 0;
    uint32_t err = 0;
    HIPCHK(hipMemcpy(&err, e->d_seg_err, sizeof err, hipMemcpyDeviceToHost));
    if (err) return fail(JMM_EHIP, "segmented chains: a synchronisation wait timed out; results are invalid");
    return 0;
}

int set_lds_attr(const void *fn, size_t lds)
{
    if (lds > 65536) {
        hipError_t err = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (err != hipSuccess) return fail(JMM_EHIP, "hipFuncSetAttribute(%zu bytes LDS): %s", lds, hipGetErrorString(err));
    }
    return 0;
}

// Pack chains into bundles: chains of equal (potential, ensemble, boundary,
// recompute interval) share a workgroup; block counts are powers of two and
// are placed largest first, so every chain's lane range is aligned to its own
// size (the segmented reductions rely on this).
struct PackOut {
    std::vector<DevBundle> b[2];
    std::vector<DevSlot> slots;
    std::vector<uint32_t> order;  // internal -> external
};

PackOut pack(const std::vector<jmm_chain_spec> &specs, const std::vector<DevChain> &ch, int T,
             size_t lds_doubles, size_t lds_max_doubles, const std::vector<uint32_t> &segS)
{
    const int n = (int)specs.size();
    std::vector<uint32_t> idx;
    for (int c = 0; c < n; ++c)
        if (!segS[c]) idx.push_back((uint32_t)c);
    // chains of one bundle share everything but N, T, P, L0, seed and stream
    // (the kernel reads the shared part once per bundle)
    auto key = [&](uint32_t c) {
        const jmm_chain_spec &s = specs[c];
        return std::make_tuple(s.potential, s.ensemble, s.boundary, s.eci, s.vol_move, s.nbn,
                               s.shifted, s.max_step, s.max_dv, s.cutoff, s.pot_p[0], s.pot_p[1],
                               s.pot_p[2], s.pot_p[3]);
    };
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
        if (key(a) != key(b)) return key(a) < key(b);
        if (ch[a].B != ch[b].B) return ch[a].B > ch[b].B;
        if (ch[a].smax != ch[b].smax) return ch[a].smax > ch[b].smax;
        return ch[a].N > ch[b].N;
    });
    PackOut o;
    // global-x chains first (they run in their own group), then LDS bundles
    std::vector<uint32_t> gx, lds;
    for (uint32_t c : idx) ((size_t)ch[c].lds_len > lds_max_doubles ? gx : lds).push_back(c);
    uint64_t xoff = 0;
    auto open_bundle = [&](int g, uint32_t c) {
        DevBundle bd{};
        bd.slot0 = (uint32_t)o.slots.size();
        bd.kind = ch[c].kind;
        bd.global_x = g;
        bd.x_off = xoff;
        o.b[g].push_back(bd);
    };
    for (uint32_t c : gx) {
        open_bundle(1, c);
        DevBundle &bd = o.b[1].back();
        o.slots.push_back({(uint32_t)o.order.size(), 0u, ch[c].B, 0u});
        bd.nslots = 1;
        bd.lanes = ch[c].B;
        bd.smax = ch[c].smax;
        bd.lds_doubles = 0;
        o.order.push_back(c);
        xoff += ch[c].N;
    }
    for (size_t k = 0; k < lds.size(); ++k) {
        const uint32_t c = lds[k];
        bool fresh = o.b[0].empty();
        if (!fresh) {
            const DevBundle &bd = o.b[0].back();
            const uint32_t c0 = o.order[o.slots[bd.slot0].chain];
            fresh = key(c0) != key(c) || bd.lanes + ch[c].B > (uint32_t)T ||
                    (size_t)bd.lds_doubles + ch[c].lds_len > lds_doubles;
        }
        if (fresh) open_bundle(0, c);
        DevBundle &bd = o.b[0].back();
        o.slots.push_back({(uint32_t)o.order.size(), bd.lanes, ch[c].B, bd.lds_doubles});
        bd.nslots++;
        bd.lanes += ch[c].B;
        bd.smax = std::max(bd.smax, ch[c].smax);
        bd.lds_doubles += ch[c].lds_len;
        o.order.push_back(c);
        xoff += ch[c].N;
    }
    return o;
}

double bundle_cost(const DevBundle &b, const std::vector<DevChain> &chi, const std::vector<DevSlot> &slots)
{
    const DevChain &c = chi[slots[b.slot0].chain];
    return (double)b.smax * (double)(c.w + 2) + ((c.flags & F_NPT) ? (double)c.smax * c.w : 0.0);
}

}  // namespace

extern "C" {

int jmm_abi_version(void) { return JMM_ABI_VERSION; }

const char *jmm_last_error(void) { return g_err.c_str(); }

int jmm_schedule(const jmm_chain_spec *spec, uint64_t *B, uint64_t *M_min, uint64_t *w)
{
    if (!spec || !B || !M_min || !w) return fail(JMM_EINVAL, "jmm_schedule: null argument");
    schedule_of(*spec, *B, *M_min, *w);
    return 0;
}

int jmm_destroy(jmm_engine *e)
{
    if (!e) return 0;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    for (auto &p : e->ev) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    (void)hipFree(e->d_chains);
    (void)hipFree(e->d_states);
    (void)hipFree(e->d_x);
    (void)hipFree(e->d_slots);
    (void)hipFree(e->d_ctrl);
    (void)hipFree(e->d_sU);
    (void)hipFree(e->d_sL);
    (void)hipFree(e->d_sA);
    (void)hipFree(e->d_sW);
    (void)hipFree(e->d_sH);
    (void)hipFree(e->d_packed_v);
    (void)hipFree(e->d_log);
    (void)hipFree(e->d_obs);
    (void)hipFree(e->d_grad);
    (void)hipFree(e->d_Ltmp);
    (void)hipFree(e->d_Lnew);
    (void)hipFree(e->d_all);
    (void)hipFree(e->d_rho);
    (void)hipFree(e->d_g);
    (void)hipFree(e->d_packed);
    (void)hipFree(e->d_ext);
    (void)hipFree(e->d_extoff);
    (void)hipFree(e->d_segs);
    (void)hipFree(e->d_pub_head);
    (void)hipFree(e->d_pub_tail);
    (void)hipFree(e->d_part);
    (void)hipFree(e->d_ctr);
    (void)hipFree(e->d_seg_err);
    for (auto &G : e->grp) {
        (void)hipFree(G.d_bundles);
        (void)hipFree(G.d_list);
    }
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return 0;
}

int jmm_create(const jmm_chain_spec *specs, int n_chains, const jmm_engine_opts *opts,
               jmm_engine **out)
{
    if (!out) return fail(JMM_EINVAL, "jmm_create: out is NULL");
    *out = nullptr;
    if (!specs || n_chains <= 0) return fail(JMM_EINVAL, "jmm_create: need at least one chain");
    jmm_engine_opts o{};
    if (opts) o = *opts;
    if (o.form < 0 || o.form > 2) return fail(JMM_EINVAL, "jmm_create: unknown form %d", o.form);
    for (int c = 0; c < n_chains; ++c)
        if (int rc = validate(specs[c], c)) return rc;

    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (o.device < 0 || o.device >= ndev) return fail(JMM_EINVAL, "jmm_create: device %d of %d", o.device, ndev);
    HIPCHK(hipSetDevice(o.device));
    void *scr = nullptr;
    if (int rc = scratch_address(o.device, &scr)) return rc;

    jmm_engine *e = new jmm_engine();
    e->scratch = (unsigned char *)scr;
    e->n = n_chains;
    e->form = o.form;
    e->device = o.device;
    e->cap = o.sample_capacity ? o.sample_capacity : 1024;
    e->block_sweeps = o.block_sweeps ? o.block_sweeps : 100;
    e->use_graph = o.use_graph ? 1 : 0;
    e->specs.assign(specs, specs + n_chains);
    auto bail = [&](int rc) {
        jmm_destroy(e);
        return rc;
    };
    // from here on a failing HIP call releases the partly built engine
#define CREATE_CHK(x)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return bail(fail(JMM_EHIP, "%s: %s", #x, hipGetErrorString(e_))); \
    } while (0)
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(JMM_EHIP, "hipStreamCreate failed"));

    std::vector<DevChain> ch(n_chains);
    uint32_t maxB = 1;
    for (int c = 0; c < n_chains; ++c) {
        ch[c] = make_chain(specs[c]);
        maxB = std::max(maxB, ch[c].B);
    }
    int lds_max = 0;
    CREATE_CHK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, o.device));
    int n_cu = 0;
    CREATE_CHK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, o.device));

    // chains whose LDS image does not fit one workgroup: split into S
    // segments of B/S blocks (S the smallest power of two whose segment image
    // fits and has at most 256 lanes); a chain that cannot be split stays
    // whole and runs from HBM
    std::vector<uint32_t> segS(n_chains, 0);
    if (e->form != JMM_FORM_PLAIN_FP64) {
        const size_t dev_max = (size_t)lds_max / 8 - 64;
        const size_t thr = o.seg_lds ? std::min<size_t>(o.seg_lds / 8, dev_max) : dev_max;
        for (int c = 0; c < n_chains; ++c) {
            if ((size_t)ch[c].lds_len <= thr) continue;
            // segments are power-of-two shares of the blocks: a chain whose
            // block count is not a power of two stays whole (LDS or HBM)
            if ((ch[c].B & (ch[c].B - 1)) != 0) continue;
            for (uint32_t S = 2; S <= ch[c].B; S *= 2) {
                const uint32_t Bs = ch[c].B / S;
                const size_t img = 2 * (size_t)ch[c].wg + (size_t)Bs * ch[c].Sp;
                if (Bs <= 256 && img <= thr) {
                    segS[c] = S;
                    break;
                }
            }
        }
    }
    maxB = 1;
    for (int c = 0; c < n_chains; ++c)
        if (!segS[c]) maxB = std::max(maxB, ch[c].B);

    // internal chain order and bundles (segmented chains last)
    std::vector<DevBundle> bundles[2];
    std::vector<DevSlot> slots;
    std::vector<uint32_t> order(n_chains);
    if (e->form == JMM_FORM_PLAIN_FP64) {
        std::iota(order.begin(), order.end(), 0u);
    } else {
        // workgroup width: 256, 512 or 1024 lanes, at least the largest
        // chain's block count (1024 lanes cap the kernel at 128 VGPRs)
        const uint32_t want = std::max<uint32_t>(o.wg_threads ? o.wg_threads : 256u, maxB);
        e->T = want <= 256 ? 256 : want <= 512 ? 512 : 1024;
        if (maxB > 1024) return bail(fail(JMM_EINVAL, "jmm_create: more than 1024 blocks in a chain"));
        size_t budget = o.lds_budget ? o.lds_budget : 65536;
        budget = std::min<size_t>(budget, (size_t)lds_max);
        PackOut p = pack(e->specs, ch, e->T, budget / 8, (size_t)lds_max / 8 - 64, segS);
        order = p.order;
        for (int c = 0; c < n_chains; ++c)
            if (segS[c]) order.push_back((uint32_t)c);
        slots = p.slots;
        for (int g = 0; g < 2; ++g) bundles[g] = p.b[g];
    }
    e->layout.assign(n_chains, jmm_layout{});
    for (int c = 0; c < n_chains; ++c) {
        jmm_layout &l = e->layout[c];
        l.B = ch[c].B;
        l.w = ch[c].w;
        l.segments = 1;
        l.lanes = 1;
        if (e->form == JMM_FORM_PLAIN_FP64) {
            l.B = 1;
            l.lds_bytes = sizeof(double) * specs[c].N;
            l.mode = l.lds_bytes > (uint64_t)lds_max ? JMM_MODE_HBM : JMM_MODE_LDS;
            if (l.mode == JMM_MODE_HBM) l.lds_bytes = 0;
        } else if (segS[c]) {
            l.mode = JMM_MODE_SPLIT;
            l.segments = segS[c];
            l.lds_bytes = sizeof(double) * (2 * (uint64_t)ch[c].wg + (uint64_t)(ch[c].B / segS[c]) * ch[c].Sp);
        } else if ((size_t)ch[c].lds_len > (size_t)lds_max / 8 - 64) {
            l.mode = JMM_MODE_HBM;
        } else {
            l.mode = JMM_MODE_LDS;
            l.lds_bytes = sizeof(double) * (uint64_t)ch[c].lds_len;
        }
    }
    e->ext_of_int = order;
    e->int_of_ext.assign(n_chains, 0);
    for (int k = 0; k < n_chains; ++k) e->int_of_ext[order[k]] = (uint32_t)k;
    e->ext_off.assign(n_chains, 0);
    uint64_t acc = 0;
    for (int c = 0; c < n_chains; ++c) {
        e->ext_off[c] = acc;
        acc += specs[c].N;
    }
    e->total_N = acc;

    // internal-order chain constants, positions, states
    e->hch.resize(n_chains);
    std::vector<double> hx(e->total_N);
    std::vector<DevState> hst(n_chains);
    uint64_t xoff = 0;
    for (int k = 0; k < n_chains; ++k) {
        const uint32_t c = order[k];
        DevChain d = ch[c];
        d.x_off = xoff;
        d.log_off = xoff * e->cap;
        const jmm_chain_spec &s = specs[c];
        double *x = hx.data() + xoff;
        for (uint64_t i = 0; i < s.N; ++i) x[i] = (((double)i + 0.5) / (double)s.N - 0.5) * s.L0;
        DevState st{};
        st.L = s.L0;
        st.U = 0.0;  // set from the GPU energy kernel below
        hst[k] = st;
        e->hch[k] = d;
        xoff += s.N;
    }
    if (int rc = dalloc(&e->d_chains, n_chains)) return bail(rc);
    if (int rc = dalloc(&e->d_states, n_chains)) return bail(rc);
    if (int rc = dalloc(&e->d_x, e->total_N)) return bail(rc);
    if (int rc = dalloc(&e->d_sU, (size_t)e->cap * n_chains)) return bail(rc);
    if (int rc = dalloc(&e->d_sL, (size_t)e->cap * n_chains)) return bail(rc);
    if (int rc = dalloc(&e->d_sA, (size_t)e->cap * n_chains)) return bail(rc);
    if (o.log_accepts)
        if (int rc = dalloc(&e->d_log, (size_t)e->cap * e->total_N)) return bail(rc);
    if (o.sample_virial) {
        if (int rc = dalloc(&e->d_sW, (size_t)e->cap * n_chains)) return bail(rc);
        if (int rc = dalloc(&e->d_sH, (size_t)e->cap * n_chains)) return bail(rc);
    }
    e->step_cap.resize(n_chains);
    for (int k = 0; k < n_chains; ++k) e->step_cap[k] = e->hch[k].max_step;
    CREATE_CHK(hipMemcpy(e->d_chains, e->hch.data(), sizeof(DevChain) * n_chains, hipMemcpyHostToDevice));
    CREATE_CHK(hipMemcpy(e->d_states, hst.data(), sizeof(DevState) * n_chains, hipMemcpyHostToDevice));
    CREATE_CHK(hipMemcpy(e->d_x, hx.data(), sizeof(double) * e->total_N, hipMemcpyHostToDevice));
    if (int rc = refresh_energies(e, -1)) return bail(rc);
    // launch groups: plain form by LDS fit; batched forms by (memory mode,
    // potential), one persistent launch each
    if (e->form == JMM_FORM_PLAIN_FP64) {
        std::vector<uint32_t> lists[2];
        size_t maxlds[2] = {0, 0};
        for (int k = 0; k < n_chains; ++k) {
            const size_t need = sizeof(double) * e->hch[k].N;
            const int g = need > (size_t)lds_max ? 1 : 0;
            lists[g].push_back((uint32_t)k);
            if (!g) maxlds[0] = std::max(maxlds[0], need);
        }
        for (int g = 0; g < 2; ++g) {
            if (lists[g].empty()) continue;
            Group G;
            G.gx = g;
            G.n_items = (uint32_t)lists[g].size();
            G.grid = G.n_items;
            G.lds = maxlds[g];
            e->grp.push_back(G);
            Group &Gr = e->grp.back();
            if (int rc = set_lds_attr(plain_kernel(g), Gr.lds)) return bail(rc);
            if (int rc = dalloc(&Gr.d_list, Gr.n_items)) return bail(rc);
            CREATE_CHK(hipMemcpy(Gr.d_list, lists[g].data(), sizeof(uint32_t) * Gr.n_items, hipMemcpyHostToDevice));
        }
    } else {
        if (int rc = dalloc(&e->d_slots, slots.size())) return bail(rc);
        CREATE_CHK(hipMemcpy(e->d_slots, slots.data(), sizeof(DevSlot) * slots.size(), hipMemcpyHostToDevice));
        for (int g = 0; g < 2; ++g) {
            std::vector<DevBundle> bv = bundles[g];
            if (bv.empty()) continue;
            // longest-processing-time first, all potentials in one queue
            std::stable_sort(bv.begin(), bv.end(), [&](const DevBundle &a, const DevBundle &b) {
                return bundle_cost(a, e->hch, slots) > bundle_cost(b, e->hch, slots);
            });
            Group G;
            G.gx = g;
            G.n_items = (uint32_t)bv.size();
            size_t maxd = 0;
            for (auto &b : bv) maxd = std::max<size_t>(maxd, b.lds_doubles);
            G.lds = g ? 0 : maxd * sizeof(double);
            const void *fn = batched_kernel(e->form == JMM_FORM_BATCHED_FP32, e->T, g);
            if (int rc = set_lds_attr(fn, G.lds)) return bail(rc);
            int occ = 0;
            CREATE_CHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, e->T, G.lds));
            if (occ < 1) occ = 1;
            G.grid = std::min<uint32_t>(G.n_items, (uint32_t)(occ * n_cu));
            e->grp.push_back(G);  // owned by the engine from here: destroy releases it
            Group &Gb = e->grp.back();
            if (int rc = dalloc(&Gb.d_bundles, Gb.n_items)) return bail(rc);
            CREATE_CHK(hipMemcpy(Gb.d_bundles, bv.data(), sizeof(DevBundle) * Gb.n_items, hipMemcpyHostToDevice));
        }
    }
    // segmented chains: one launch, one workgroup per segment, all resident
    {
        std::vector<DevSeg> segv;
        uint32_t wgmax = 1, bsmax = 1;
        size_t img = 0;
        for (int k = 0; k < n_chains; ++k) {
            const uint32_t S = segS[order[k]];
            if (!S) continue;
            const DevChain &d = e->hch[k];
            const uint32_t first = (uint32_t)segv.size();
            for (uint32_t si = 0; si < S; ++si) {
                DevSeg g{};
                g.chain = (uint32_t)k;
                g.s = si;
                g.S = S;
                g.Bs = d.B / S;
                g.first = first;
                segv.push_back(g);
            }
            wgmax = std::max(wgmax, d.wg);
            bsmax = std::max(bsmax, d.B / S);
            img = std::max(img, 2 * (size_t)d.wg + (size_t)(d.B / S) * d.Sp);
        }
        if (!segv.empty()) {
            const int fp32 = e->form == JMM_FORM_BATCHED_FP32;
            if (o.seg_lanes > 2) return bail(fail(JMM_EINVAL, "jmm_create: seg_lanes must be 0, 1 or 2"));
            e->seg_lpb = o.seg_lanes ? (int)o.seg_lanes : 1;
            e->seg_tw = e->seg_lpb * bsmax > 256 ? 512 : 256;
            for (int c = 0; c < n_chains; ++c)
                if (segS[c]) e->layout[c].lanes = (uint32_t)e->seg_lpb;
            e->n_segs = (uint32_t)segv.size();
            e->seg_wgmax = wgmax;
            e->seg_lds_bytes = img * sizeof(double);
            const void *fn = segmented_kernel(fp32, e->seg_tw, e->seg_lpb);
            if (int rc = set_lds_attr(fn, e->seg_lds_bytes)) return bail(rc);
            int occ = 0;
            CREATE_CHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, e->seg_tw, e->seg_lds_bytes));
            if ((uint64_t)occ * (uint64_t)n_cu < e->n_segs)
                return bail(fail(JMM_EINVAL, "jmm_create: split chains need %u workgroups resident at once, "
                                             "the device holds %d", e->n_segs, occ * n_cu));
            if (int rc = dalloc(&e->d_segs, segv.size())) return bail(rc);
            CREATE_CHK(hipMemcpy(e->d_segs, segv.data(), sizeof(DevSeg) * segv.size(), hipMemcpyHostToDevice));
            if (int rc = dalloc(&e->d_pub_head, (size_t)e->n_segs * wgmax)) return bail(rc);
            if (int rc = dalloc(&e->d_pub_tail, (size_t)e->n_segs * wgmax)) return bail(rc);
            if (int rc = dalloc(&e->d_part, (size_t)e->n_segs * 2 * SEG_FIELDS * SEG_WAVES)) return bail(rc);
            if (int rc = dalloc(&e->d_ctr, e->n_segs)) return bail(rc);
            if (int rc = dalloc(&e->d_seg_err, 1)) return bail(rc);
            std::vector<unsigned long long> z(e->n_segs, 0ull);
            CREATE_CHK(hipMemcpy(e->d_ctr, z.data(), sizeof(unsigned long long) * z.size(), hipMemcpyHostToDevice));
            const uint32_t zero = 0;
            CREATE_CHK(hipMemcpy(e->d_seg_err, &zero, sizeof zero, hipMemcpyHostToDevice));
            e->seg_ctrl = (uint32_t)e->grp.size();
            e->use_graph = 0;  // split chains are launched directly, not from a graph
        }
    }
    e->n_ctrl = (uint32_t)e->grp.size() + (e->n_segs ? 1u : 0u);
    if (int rc = dalloc(&e->d_ctrl, std::max<size_t>(1, e->n_ctrl))) return bail(rc);
    {
        std::vector<Ctrl> zc(std::max<size_t>(1, e->n_ctrl));
        CREATE_CHK(hipMemcpy(e->d_ctrl, zc.data(), sizeof(Ctrl) * zc.size(), hipMemcpyHostToDevice));
    }

    if (e->use_graph) {
        CREATE_CHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = launch_block(e, e->block_sweeps);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(e->stream, &g);
        e->graph = g;  // released by destroy, also when the capture failed
        if (rc) return bail(rc);
        if (ce != hipSuccess) return bail(fail(JMM_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ce)));
        CREATE_CHK(hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0));
    }
    CREATE_CHK(hipStreamSynchronize(e->stream));
    *out = e;
    return 0;
}
#undef CREATE_CHK

static int launch_hist_sample(jmm_engine *e)
{
    HistArgs h{};
    h.chains = e->d_chains;
    h.states = e->d_states;
    h.x = e->d_x;
    h.rho = e->d_rho;
    h.g = e->d_g;
    h.rbw = e->hs.rbw;
    h.gsw = e->hs.gsw;
    h.gbw = e->hs.gbw;
    h.n_chains = (uint32_t)e->n;
    h.rhonb = e->hs.rhonb;
    h.gns = e->hs.gns;
    h.gnb = e->hs.gnb;
    uint64_t nmax = 1;
    for (const auto &c : e->hch) nmax = std::max<uint64_t>(nmax, c.N);
    h.parts = (uint32_t)std::min<uint64_t>(64, (nmax + 255) / 256);
    HIPCHK(launch_hist(h, e->stream));
    e->hist_samples++;
    return 0;
}

int jmm_hist_configure(jmm_engine *e, const jmm_hist_spec *spec, uint32_t every)
{
    if (!e || (every && !spec)) return fail(JMM_EINVAL, "jmm_hist_configure: null argument");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    (void)hipFree(e->d_rho);
    (void)hipFree(e->d_g);
    e->d_rho = nullptr;
    e->d_g = nullptr;
    e->hist_every = 0;
    e->hist_samples = 0;
    if (!every) return 0;
    const jmm_hist_spec &h = *spec;
    if (h.rhonb == 0 || !(h.rbw > 0.0) || (h.gns && (!(h.gsw > 0.0) || h.gnb == 0 || !(h.gbw > 0.0))))
        return fail(JMM_EINVAL, "jmm_hist_configure: bin counts and widths must be positive");
    const uint64_t nr = (uint64_t)e->n * h.rhonb, ng = (uint64_t)e->n * h.gns * h.gnb;
    if (int rc = dalloc(&e->d_rho, nr)) return rc;
    if (ng)
        if (int rc = dalloc(&e->d_g, ng)) return rc;
    HIPCHK(hipMemsetAsync(e->d_rho, 0, sizeof(unsigned long long) * nr, e->stream));
    if (ng) HIPCHK(hipMemsetAsync(e->d_g, 0, sizeof(unsigned long long) * ng, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->hs = h;
    e->hist_every = every;
    return 0;
}

int jmm_hist_sample(jmm_engine *e)
{
    if (!e) return fail(JMM_EINVAL, "jmm_hist_sample: null engine");
    if (!e->hist_every) return fail(JMM_ESTATE, "jmm_hist_sample: histograms are not configured");
    HIPCHK(hipSetDevice(e->device));
    return launch_hist_sample(e);
}

int jmm_hist_read(jmm_engine *e, uint64_t *rho, uint64_t *g, uint64_t *n_samples, int reset)
{
    if (!e || !n_samples) return fail(JMM_EINVAL, "jmm_hist_read: null argument");
    if (!e->hist_every) return fail(JMM_ESTATE, "jmm_hist_read: histograms are not configured");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    const uint64_t rn = e->hs.rhonb, gn = (uint64_t)e->hs.gns * e->hs.gnb;
    if (rho) {
        std::vector<unsigned long long> h((size_t)e->n * rn);
        HIPCHK(hipMemcpy(h.data(), e->d_rho, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
        for (int k = 0; k < e->n; ++k)
            for (uint64_t b = 0; b < rn; ++b) rho[(uint64_t)e->ext_of_int[k] * rn + b] = h[(uint64_t)k * rn + b];
    }
    if (g && gn) {
        std::vector<unsigned long long> h((size_t)e->n * gn);
        HIPCHK(hipMemcpy(h.data(), e->d_g, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
        for (int k = 0; k < e->n; ++k)
            for (uint64_t b = 0; b < gn; ++b) g[(uint64_t)e->ext_of_int[k] * gn + b] = h[(uint64_t)k * gn + b];
    }
    *n_samples = e->hist_samples;
    if (reset) {
        HIPCHK(hipMemsetAsync(e->d_rho, 0, sizeof(unsigned long long) * e->n * rn, e->stream));
        if (gn) HIPCHK(hipMemsetAsync(e->d_g, 0, sizeof(unsigned long long) * e->n * gn, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        e->hist_samples = 0;
    }
    return 0;
}

int jmm_run_sweeps(jmm_engine *e, uint32_t n_sweeps)
{
    if (!e) return fail(JMM_EINVAL, "jmm_run_sweeps: null engine");
    if ((uint64_t)e->write_pos + n_sweeps > e->cap)
        return fail(JMM_EFULL, "sample buffer holds %u sweeps, %u pending: read samples first", e->cap, e->write_pos);
    HIPCHK(hipSetDevice(e->device));
    while (n_sweeps > 0) {
        uint32_t m = std::min(n_sweeps, e->block_sweeps);
        // with histograms on, blocks end on every sampling sweep
        if (e->hist_every)
            m = std::min<uint32_t>(m, e->hist_every - (uint32_t)(e->sweeps_run % e->hist_every));
        if (e->ev.size() >= EV_KEEP)
            if (int rc = fold_events(e, false)) return rc;
        hipEvent_t a, b;
        HIPCHK(hipEventCreate(&a));
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            return fail(JMM_EHIP, "hipEventCreate failed");
        }
        e->ev.push_back({a, b});
        HIPCHK(hipEventRecord(a, e->stream));
        if (e->gexec && m == e->block_sweeps) {
            HIPCHK(hipGraphLaunch(e->gexec, e->stream));
        } else {
            if (int rc = launch_block(e, m)) return rc;
        }
        HIPCHK(hipEventRecord(b, e->stream));
        e->ev_sweeps += m;
        e->write_pos += m;
        e->sweeps_run += m;
        n_sweeps -= m;
        if (e->hist_every && e->sweeps_run % e->hist_every == 0) {
            if (int rc = launch_hist_sample(e)) return rc;
        }
    }
    return 0;
}

int jmm_synchronize(jmm_engine *e)
{
    if (!e) return fail(JMM_EINVAL, "jmm_synchronize: null engine");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    return check_segments(e);
}

int jmm_kernel_time(jmm_engine *e, uint64_t *launches, double *ms, uint64_t *sweeps, int reset)
{
    if (!e) return fail(JMM_EINVAL, "jmm_kernel_time: null engine");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (int rc = fold_events(e, true)) return rc;
    if (launches) *launches = e->ev_done;
    if (ms) *ms = e->ev_ms_done;
    if (sweeps) *sweeps = e->ev_sweeps;
    if (reset) {
        e->ev_done = 0;
        e->ev_ms_done = 0.0;
        e->ev_sweeps = 0;
    }
    return 0;
}

// Device copies of the internal -> external chain maps (index and position
// offset), for kernels that write results in the caller's order.
static int ensure_ext(jmm_engine *e)
{
    if (e->d_ext) return 0;
    std::vector<uint64_t> off(e->n);
    for (int k = 0; k < e->n; ++k) off[k] = e->ext_off[e->ext_of_int[k]];
    if (int rc = dalloc(&e->d_ext, (size_t)e->n)) return rc;
    if (int rc = dalloc(&e->d_extoff, (size_t)e->n)) return rc;
    HIPCHK(hipMemcpy(e->d_ext, e->ext_of_int.data(), sizeof(uint32_t) * e->n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_extoff, off.data(), sizeof(uint64_t) * e->n, hipMemcpyHostToDevice));
    return 0;
}

static_assert(sizeof(jmm_sample) == sizeof(SampleRec) && offsetof(jmm_sample, L) == offsetof(SampleRec, L) &&
                  offsetof(jmm_sample, acc_d) == offsetof(SampleRec, acc_d) &&
                  offsetof(jmm_sample, acc_v) == offsetof(SampleRec, acc_v),
              "the device sample record must have the ABI's layout");

int jmm_read_samples(jmm_engine *e, jmm_sample *out, uint32_t max_sweeps, uint32_t *n_out)
{
    return jmm_read_samples_virial(e, out, nullptr, max_sweeps, n_out);
}

int jmm_read_samples_virial(jmm_engine *e, jmm_sample *out, double *virial, uint32_t max_sweeps, uint32_t *n_out)
{
    if (!e || !n_out) return fail(JMM_EINVAL, "jmm_read_samples: null argument");
    if (virial && !e->d_sW) return fail(JMM_ESTATE, "jmm_read_samples_virial: engine created without sample_virial");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (int rc = check_segments(e)) return rc;
    const uint32_t n = e->write_pos;
    if (n > max_sweeps) return fail(JMM_EINVAL, "jmm_read_samples: %u sweeps pending, room for %u", n, max_sweeps);
    if (n && !out) return fail(JMM_EINVAL, "jmm_read_samples: out is NULL");
    if (n) {
        // transposed to [chain][sweep] on the device, then one pitched copy
        // into the caller's rows of max_sweeps samples
        if (int rc = ensure_ext(e)) return rc;
        if (!e->d_packed)
            if (int rc = dalloc(&e->d_packed, (size_t)e->cap * e->n)) return rc;
        HIPCHK(launch_pack_samples(e->d_sU, e->d_sL, e->d_sA, e->d_ext, (uint32_t)e->n, n, e->d_packed,
                                   e->stream));
        if (virial) {
            if (!e->d_packed_v)
                if (int rc = dalloc(&e->d_packed_v, (size_t)e->cap * e->n * 2)) return rc;
            HIPCHK(launch_pack_virials(e->d_sW, e->d_sH, e->d_ext, (uint32_t)e->n, n, e->d_packed_v, e->stream));
        }
        HIPCHK(hipStreamSynchronize(e->stream));
        HIPCHK(hipMemcpy2D(out, (size_t)max_sweeps * sizeof(jmm_sample), e->d_packed, (size_t)n * sizeof(SampleRec),
                           (size_t)n * sizeof(SampleRec), (size_t)e->n, hipMemcpyDeviceToHost));
        if (virial)
            HIPCHK(hipMemcpy2D(virial, (size_t)max_sweeps * 2 * sizeof(double), e->d_packed_v,
                               (size_t)n * 2 * sizeof(double), (size_t)n * 2 * sizeof(double), (size_t)e->n,
                               hipMemcpyDeviceToHost));
    }
    e->write_pos = 0;
    std::vector<Ctrl> zc(std::max<size_t>(1, e->n_ctrl));
    HIPCHK(hipMemcpy(e->d_ctrl, zc.data(), sizeof(Ctrl) * zc.size(), hipMemcpyHostToDevice));
    *n_out = n;
    return 0;
}

int jmm_read_samples_noalloc(jmm_engine *e, jmm_sample *out, uint32_t max_sweeps, uint32_t *n_out)
{
    if (!e || !n_out) return fail(JMM_EINVAL, "jmm_read_samples_noalloc: null argument");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (int rc = check_segments(e)) return rc;
    const uint32_t n = e->write_pos;
    if (n > max_sweeps)
        return fail(JMM_EINVAL, "jmm_read_samples_noalloc: %u sweeps pending, room for %u", n, max_sweeps);
    if (n && !out) return fail(JMM_EINVAL, "jmm_read_samples_noalloc: out is NULL");
    // staging in `scratch`: the chain-order map, then the records
    const size_t map_bytes = ((sizeof(uint32_t) * (size_t)e->n + 255) / 256) * 256;
    const size_t rec_bytes = sizeof(SampleRec) * (size_t)n * (size_t)e->n;
    if (map_bytes + rec_bytes > SCRATCH_BYTES)
        return fail(JMM_EINVAL,
                    "jmm_read_samples_noalloc: %u sweeps of %d chains need %zu bytes of staging, more than "
                    "the %zu it holds; read after fewer sweeps or use jmm_read_samples",
                    n, e->n, map_bytes + rec_bytes, SCRATCH_BYTES);
    if (n) {
        uint32_t *d_map = (uint32_t *)e->scratch;
        SampleRec *d_rec = (SampleRec *)(e->scratch + map_bytes);
        HIPCHK(hipMemcpyAsync(d_map, e->ext_of_int.data(), sizeof(uint32_t) * e->n, hipMemcpyHostToDevice,
                              e->stream));
        HIPCHK(launch_pack_samples(e->d_sU, e->d_sL, e->d_sA, d_map, (uint32_t)e->n, n, d_rec, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        HIPCHK(hipMemcpy2D(out, (size_t)max_sweeps * sizeof(jmm_sample), d_rec, (size_t)n * sizeof(SampleRec),
                           (size_t)n * sizeof(SampleRec), (size_t)e->n, hipMemcpyDeviceToHost));
    }
    e->write_pos = 0;
    HIPCHK(hipMemsetAsync(e->d_ctrl, 0, sizeof(Ctrl) * std::max<size_t>(1, e->n_ctrl), e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    *n_out = n;
    return 0;
}

int jmm_read_accept_log(jmm_engine *e, uint8_t *out, uint32_t max_sweeps, uint32_t *n_out)
{
    if (!e || !n_out) return fail(JMM_EINVAL, "jmm_read_accept_log: null argument");
    if (!e->d_log) return fail(JMM_ESTATE, "engine was created without log_accepts");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (int rc = check_segments(e)) return rc;
    const uint32_t n = e->write_pos;
    if (n > max_sweeps) return fail(JMM_EINVAL, "jmm_read_accept_log: %u sweeps pending, room for %u", n, max_sweeps);
    for (int k = 0; k < e->n && n; ++k) {
        const uint32_t c = e->ext_of_int[k];
        const uint64_t N = e->specs[c].N;
        HIPCHK(hipMemcpy(out + e->ext_off[c] * max_sweeps, e->d_log + e->hch[k].log_off,
                         (size_t)n * N, hipMemcpyDeviceToHost));
    }
    *n_out = n;
    return 0;
}

int jmm_get_state(jmm_engine *e, int chain, double *x, double *L, double *U, uint64_t *sweep)
{
    if (!e || chain < 0 || chain >= e->n) return fail(JMM_EINVAL, "jmm_get_state: bad chain");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (int rc = check_segments(e)) return rc;
    const uint32_t k = e->int_of_ext[chain];
    if (x) HIPCHK(hipMemcpy(x, e->d_x + e->hch[k].x_off, sizeof(double) * e->hch[k].N, hipMemcpyDeviceToHost));
    DevState st;
    HIPCHK(hipMemcpy(&st, e->d_states + k, sizeof st, hipMemcpyDeviceToHost));
    if (L) *L = st.L;
    if (U) *U = st.U;
    if (sweep) *sweep = st.sweep;
    return 0;
}

int jmm_chain_layout(jmm_engine *e, int chain, jmm_layout *out)
{
    if (!e || !out || chain < 0 || chain >= e->n) return fail(JMM_EINVAL, "jmm_chain_layout: bad argument");
    *out = e->layout[chain];
    return 0;
}

int jmm_get_states(jmm_engine *e, double *x, double *L, double *U)
{
    if (!e) return fail(JMM_EINVAL, "jmm_get_states: null engine");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (int rc = check_segments(e)) return rc;
    if (x) {
        std::vector<double> hx(e->total_N);
        HIPCHK(hipMemcpy(hx.data(), e->d_x, sizeof(double) * e->total_N, hipMemcpyDeviceToHost));
        for (int k = 0; k < e->n; ++k) {
            const uint32_t c = e->ext_of_int[k];
            std::memcpy(x + e->ext_off[c], hx.data() + e->hch[k].x_off, sizeof(double) * e->hch[k].N);
        }
    }
    if (L || U) {
        std::vector<DevState> st(e->n);
        HIPCHK(hipMemcpy(st.data(), e->d_states, sizeof(DevState) * e->n, hipMemcpyDeviceToHost));
        for (int k = 0; k < e->n; ++k) {
            const uint32_t c = e->ext_of_int[k];
            if (L) L[c] = st[k].L;
            if (U) U[c] = st[k].U;
        }
    }
    return 0;
}

int jmm_set_state(jmm_engine *e, int chain, const double *x, double L, uint64_t sweep)
{
    if (!e || chain < 0 || chain >= e->n || !x) return fail(JMM_EINVAL, "jmm_set_state: bad argument");
    HIPCHK(hipSetDevice(e->device));
    const jmm_chain_spec &s = e->specs[chain];
    const uint32_t k = e->int_of_ext[chain];
    const DevChain &d = e->hch[k];
    const bool per = s.boundary == JMM_PERIODIC;
    if (!(L > 0.0) || (per && !(L > d.Lmin))) return fail(JMM_EINVAL, "jmm_set_state: bad length %g", L);
    for (uint64_t i = 0; i < s.N; ++i) {
        if (!std::isfinite(x[i])) return fail(JMM_EINVAL, "jmm_set_state: position %llu not finite", (unsigned long long)i);
        if (!per && std::fabs(x[i]) > L / 2.0) return fail(JMM_EINVAL, "jmm_set_state: position %llu outside the walls", (unsigned long long)i);
        if (i + 1 < s.N && !(x[i + 1] - x[i] >= d.gmin))
            return fail(JMM_EINVAL, "jmm_set_state: positions %llu,%llu unsorted or inside the hard core", (unsigned long long)i, (unsigned long long)i + 1);
    }
    if (per && s.N >= 2 && !(x[0] + L - x[s.N - 1] >= d.gmin))
        return fail(JMM_EINVAL, "jmm_set_state: wrap-around pair inside the hard core");
    HIPCHK(hipStreamSynchronize(e->stream));
    DevState st;
    HIPCHK(hipMemcpy(&st, e->d_states + k, sizeof st, hipMemcpyDeviceToHost));
    st.L = L;
    st.U = 0.0;
    st.sweep = sweep;
    HIPCHK(hipMemcpy(e->d_x + d.x_off, x, sizeof(double) * s.N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_states + k, &st, sizeof st, hipMemcpyHostToDevice));
    return refresh_energies(e, (int)k);
}

int jmm_set_states(jmm_engine *e, const double *x, const double *L, const double *U, const uint64_t *sweep)
{
    if (!e || !x || !L) return fail(JMM_EINVAL, "jmm_set_states: null argument");
    HIPCHK(hipSetDevice(e->device));
    for (int c = 0; c < e->n; ++c) {
        const jmm_chain_spec &s = e->specs[c];
        const DevChain &d = e->hch[e->int_of_ext[c]];
        const bool per = s.boundary == JMM_PERIODIC;
        const double *xc = x + e->ext_off[c];
        if (!(L[c] > 0.0) || !std::isfinite(L[c]) || (per && !(L[c] > d.Lmin)))
            return fail(JMM_EINVAL, "jmm_set_states: chain %d: bad length %g", c, L[c]);
        if (U && !std::isfinite(U[c])) return fail(JMM_EINVAL, "jmm_set_states: chain %d: energy not finite", c);
        for (uint64_t i = 0; i < s.N; ++i) {
            if (!std::isfinite(xc[i])) return fail(JMM_EINVAL, "jmm_set_states: chain %d: position %llu not finite", c, (unsigned long long)i);
            if (!per && std::fabs(xc[i]) > L[c] / 2.0)
                return fail(JMM_EINVAL, "jmm_set_states: chain %d: position %llu outside the walls", c, (unsigned long long)i);
            if (i + 1 < s.N && !(xc[i + 1] - xc[i] >= d.gmin))
                return fail(JMM_EINVAL, "jmm_set_states: chain %d: positions %llu,%llu unsorted or inside the hard core", c,
                            (unsigned long long)i, (unsigned long long)i + 1);
        }
        if (per && s.N >= 2 && !(xc[0] + L[c] - xc[s.N - 1] >= d.gmin))
            return fail(JMM_EINVAL, "jmm_set_states: chain %d: wrap-around pair inside the hard core", c);
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    std::vector<double> hx(e->total_N);
    std::vector<DevState> st(e->n);
    HIPCHK(hipMemcpy(st.data(), e->d_states, sizeof(DevState) * e->n, hipMemcpyDeviceToHost));
    for (int k = 0; k < e->n; ++k) {
        const uint32_t c = e->ext_of_int[k];
        std::memcpy(hx.data() + e->hch[k].x_off, x + e->ext_off[c], sizeof(double) * e->hch[k].N);
        st[k].L = L[c];
        st[k].U = U ? U[c] : 0.0;
        if (sweep) st[k].sweep = sweep[c];
    }
    HIPCHK(hipMemcpy(e->d_x, hx.data(), sizeof(double) * e->total_N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_states, st.data(), sizeof(DevState) * e->n, hipMemcpyHostToDevice));
    return U ? 0 : refresh_energies(e, -1);
}

int jmm_set_steps(jmm_engine *e, const double *max_step, const double *max_dv)
{
    if (!e) return fail(JMM_EINVAL, "jmm_set_steps: null engine");
    HIPCHK(hipSetDevice(e->device));
    for (int c = 0; c < e->n; ++c) {
        const uint32_t k = e->int_of_ext[c];
        if (max_step && (!(max_step[c] >= 0.0) || max_step[c] > e->step_cap[k]))
            return fail(JMM_EINVAL, "jmm_set_steps: chain %d: max_step %g outside [0, %g] (the creation value)",
                        c, max_step ? max_step[c] : 0.0, e->step_cap[k]);
        if (max_dv && !(max_dv[c] >= 0.0 && std::isfinite(max_dv[c])))
            return fail(JMM_EINVAL, "jmm_set_steps: chain %d: bad max_dv", c);
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int c = 0; c < e->n; ++c) {
        const uint32_t k = e->int_of_ext[c];
        if (max_step) {
            e->hch[k].max_step = max_step[c];
            e->specs[c].max_step = max_step[c];
        }
        if (max_dv) {
            e->hch[k].max_dv = max_dv[c];
            e->specs[c].max_dv = max_dv[c];
        }
    }
    HIPCHK(hipMemcpy(e->d_chains, e->hch.data(), sizeof(DevChain) * e->n, hipMemcpyHostToDevice));
    return 0;
}

int jmm_perturb_energy(jmm_engine *e, int chain, double delta)
{
    if (!e || chain < 0 || chain >= e->n) return fail(JMM_EINVAL, "jmm_perturb_energy: bad chain");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    const uint32_t k = e->int_of_ext[chain];
    DevState st;
    HIPCHK(hipMemcpy(&st, e->d_states + k, sizeof st, hipMemcpyDeviceToHost));
    st.U += delta;
    HIPCHK(hipMemcpy(e->d_states + k, &st, sizeof st, hipMemcpyHostToDevice));
    return 0;
}

static void fill_counters(const DevState &st, jmm_counters *out)
{
    out->sweeps = st.sweep;
    out->trials = st.cnt[0];
    out->acc_d = st.cnt[1];
    out->wall_rejects = st.cnt[2];
    out->core_rejects = st.cnt[3];
    out->pair_terms = st.cnt[4];
    out->vol_trials = st.cnt[5];
    out->acc_v = st.cnt[6];
    out->recomputes = st.cnt[7];
    out->max_drift = st.max_drift;
    out->U = st.U;
    out->L = st.L;
}

int jmm_get_counters(jmm_engine *e, int chain, jmm_counters *out)
{
    if (!e || chain < -1 || chain >= e->n || !out) return fail(JMM_EINVAL, "jmm_get_counters: bad argument");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (chain >= 0) {
        DevState st;
        HIPCHK(hipMemcpy(&st, e->d_states + e->int_of_ext[chain], sizeof st, hipMemcpyDeviceToHost));
        fill_counters(st, out);
        return 0;
    }
    std::vector<DevState> st(e->n);
    HIPCHK(hipMemcpy(st.data(), e->d_states, sizeof(DevState) * e->n, hipMemcpyDeviceToHost));
    for (int k = 0; k < e->n; ++k) fill_counters(st[k], out + e->ext_of_int[k]);
    return 0;
}

// obs/grad: host arrays (copied out after the kernel, synchronous), or with
// on_device the caller's device arrays, written by the kernel on the
// engine's stream and not waited for.
static int observables_impl(jmm_engine *e, double *obs, double *grad, int fp32,
                            const double *Lnew_dev = nullptr, bool on_device = false)
{
    HIPCHK(hipSetDevice(e->device));
    if (int rc = ensure_ext(e)) return rc;
    if (!e->d_obs)
        if (int rc = dalloc(&e->d_obs, (size_t)e->n * 6)) return rc;
    if (!e->d_Ltmp)
        if (int rc = dalloc(&e->d_Ltmp, (size_t)e->n)) return rc;
    if (!e->d_all) {
        if (int rc = dalloc(&e->d_all, (size_t)e->n)) return rc;
        std::vector<uint32_t> all(e->n);
        std::iota(all.begin(), all.end(), 0u);
        HIPCHK(hipMemcpy(e->d_all, all.data(), sizeof(uint32_t) * e->n, hipMemcpyHostToDevice));
    }
    if (grad && !on_device && !e->d_grad)
        if (int rc = dalloc(&e->d_grad, e->total_N)) return rc;
    HIPCHK(launch_lengths(e->d_states, e->d_Ltmp, (uint32_t)e->n, e->stream));
    ObsArgs a{};
    a.chains = e->d_chains;
    a.x = e->d_x;
    a.L = e->d_Ltmp;
    a.list = e->d_all;
    a.Lnew = Lnew_dev;
    a.obs = on_device ? obs : e->d_obs;
    a.grad = on_device ? grad : (grad ? e->d_grad : nullptr);
    // results in the caller's chain order, copied out as they are
    a.row = e->d_ext;
    a.goff = e->d_extoff;
    a.n = (uint32_t)e->n;
    a.standalone = 0;
    HIPCHK(launch_observables(a, fp32, (uint32_t)e->n, e->stream));
    if (on_device) return 0;
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(obs, e->d_obs, sizeof(double) * (size_t)e->n * 6, hipMemcpyDeviceToHost));
    if (grad) HIPCHK(hipMemcpy(grad, e->d_grad, sizeof(double) * e->total_N, hipMemcpyDeviceToHost));
    return 0;
}

// FP64 energies of the current configurations (the GPU observables kernel)
// written into the chain states: every chain (only = -1) or one internal
// chain.  Plays the role of fad's step-0 energy (src/Main.cpp:66-68).
static int refresh_energies(jmm_engine *e, int only)
{
    std::vector<double> obs((size_t)e->n * 6);
    if (int rc = observables_impl(e, obs.data(), nullptr, 0)) return rc;
    std::vector<DevState> st(e->n);
    HIPCHK(hipMemcpy(st.data(), e->d_states, sizeof(DevState) * e->n, hipMemcpyDeviceToHost));
    for (int k = 0; k < e->n; ++k)
        if (only < 0 || k == only) st[k].U = obs[(size_t)e->ext_of_int[k] * 6];
    HIPCHK(hipMemcpy(e->d_states, st.data(), sizeof(DevState) * e->n, hipMemcpyHostToDevice));
    return 0;
}

// Per-chain lengths (external order) copied to the device in internal order.
static int upload_lengths(jmm_engine *e, const double *L_ext)
{
    if (!e->d_Lnew)
        if (int rc = dalloc(&e->d_Lnew, (size_t)e->n)) return rc;
    std::vector<double> h(e->n);
    for (int k = 0; k < e->n; ++k) h[k] = L_ext[e->ext_of_int[k]];
    HIPCHK(hipMemcpy(e->d_Lnew, h.data(), sizeof(double) * e->n, hipMemcpyHostToDevice));
    return 0;
}

int jmm_scaled_energy(jmm_engine *e, const double *L_try, double *U_out)
{
    if (!e || !L_try || !U_out) return fail(JMM_EINVAL, "jmm_scaled_energy: null argument");
    for (int c = 0; c < e->n; ++c)
        if (!(L_try[c] > 0.0) || !std::isfinite(L_try[c]))
            return fail(JMM_EINVAL, "jmm_scaled_energy: chain %d: bad length %g", c, L_try[c]);
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (int rc = upload_lengths(e, L_try)) return rc;
    std::vector<double> obs((size_t)e->n * 6);
    if (int rc = observables_impl(e, obs.data(), nullptr, 0, e->d_Lnew)) return rc;
    for (int c = 0; c < e->n; ++c) U_out[c] = obs[(size_t)c * 6];
    return 0;
}

int jmm_set_lengths(jmm_engine *e, const double *L_new)
{
    if (!e || !L_new) return fail(JMM_EINVAL, "jmm_set_lengths: null argument");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    std::vector<DevState> st(e->n);
    HIPCHK(hipMemcpy(st.data(), e->d_states, sizeof(DevState) * e->n, hipMemcpyDeviceToHost));
    // the scaled configuration must keep the hard core and (periodic) the
    // length bound the window is built for
    std::vector<double> x(e->total_N);
    HIPCHK(hipMemcpy(x.data(), e->d_x, sizeof(double) * e->total_N, hipMemcpyDeviceToHost));
    for (int c = 0; c < e->n; ++c) {
        const uint32_t k = e->int_of_ext[c];
        const DevChain &d = e->hch[k];
        const bool per = (d.flags & F_PERIODIC) != 0;
        const double Ln = L_new[c];
        if (!(Ln > 0.0) || !std::isfinite(Ln) || (per && !(Ln > d.Lmin)))
            return fail(JMM_EINVAL, "jmm_set_lengths: chain %d: bad length %g", c, Ln);
        const double sc = Ln / st[k].L;
        const double *xc = x.data() + d.x_off;
        double gmin_seen = INFINITY;
        for (uint64_t i = 0; i + 1 < d.N; ++i) gmin_seen = std::min(gmin_seen, (xc[i + 1] - xc[i]) * sc);
        if (per && d.N >= 1) gmin_seen = std::min(gmin_seen, (xc[0] + st[k].L - xc[d.N - 1]) * sc);
        if (!(gmin_seen >= d.gmin) && d.N >= 2)
            return fail(JMM_EINVAL, "jmm_set_lengths: chain %d: length %g puts a pair inside the hard core", c, Ln);
        if (!per && d.N >= 1 && (std::fabs(xc[0] * sc) > Ln / 2.0 || std::fabs(xc[d.N - 1] * sc) > Ln / 2.0))
            return fail(JMM_EINVAL, "jmm_set_lengths: chain %d: positions outside the walls", c);
    }
    if (int rc = upload_lengths(e, L_new)) return rc;
    uint64_t nmax = 1;
    for (const auto &d : e->hch) nmax = std::max<uint64_t>(nmax, d.N);
    HIPCHK(launch_rescale(e->d_chains, e->d_states, e->d_x, e->d_Lnew, (uint32_t)e->n,
                          (uint32_t)std::min<uint64_t>(64, (nmax + 255) / 256), e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int c = 0; c < e->n; ++c) st[e->int_of_ext[c]].L = L_new[c];
    HIPCHK(hipMemcpy(e->d_states, st.data(), sizeof(DevState) * e->n, hipMemcpyHostToDevice));
    return refresh_energies(e, -1);
}

int jmm_observables(jmm_engine *e, double *obs, double *grad)
{
    if (!e || !obs) return fail(JMM_EINVAL, "jmm_observables: null argument");
    HIPCHK(hipSetDevice(e->device));
    return observables_impl(e, obs, grad, e->form == JMM_FORM_BATCHED_FP32);
}

int jmm_observables_device(jmm_engine *e, double *obs_dev, double *grad_dev)
{
    if (!e || !obs_dev) return fail(JMM_EINVAL, "jmm_observables_device: null argument");
    HIPCHK(hipSetDevice(e->device));
    return observables_impl(e, obs_dev, grad_dev, e->form == JMM_FORM_BATCHED_FP32, nullptr, true);
}

int jmm_full_energy(jmm_engine *e, double *U_out)
{
    if (!e || !U_out) return fail(JMM_EINVAL, "jmm_full_energy: null argument");
    HIPCHK(hipSetDevice(e->device));
    std::vector<double> obs((size_t)e->n * 6);
    if (int rc = observables_impl(e, obs.data(), nullptr, 0)) return rc;
    for (int c = 0; c < e->n; ++c) U_out[c] = obs[(size_t)c * 6];
    return 0;
}

int jmm_pair_observables(const jmm_chain_spec *spec, int n_configs, const double *x_dev,
                         const double *L_dev, double *obs_dev, double *grad_dev, int precision,
                         void *stream)
{
    if (!spec || n_configs <= 0 || !x_dev || !L_dev || !obs_dev)
        return fail(JMM_EINVAL, "jmm_pair_observables: null or empty argument");
    if (precision != 32 && precision != 64) return fail(JMM_EINVAL, "jmm_pair_observables: precision must be 32 or 64");
    jmm_chain_spec s = *spec;
    s.L0 = std::isfinite(s.L0) && s.L0 > 0 ? s.L0 : 1.0;
    if (s.N < 1 || s.potential < 0 || s.potential > 3 || !(s.cutoff > 0.0))
        return fail(JMM_EINVAL, "jmm_pair_observables: bad potential description");
    ObsArgs a{};
    a.one = make_chain(s);
    // configurations from the caller may hold gaps below the hard core: walk
    // every partner to the cutoff (or the NBN limit), not the chain's window
    a.one.w = std::min<uint64_t>(a.one.nbn_lim, s.N - 1);
    a.x = x_dev;
    a.L = L_dev;
    a.obs = obs_dev;
    a.grad = grad_dev;
    a.n = (uint32_t)n_configs;
    a.standalone = 1;
    HIPCHK(launch_observables(a, precision == 32, (uint32_t)n_configs, (hipStream_t)stream));
    return 0;
}

}  // extern "C"
