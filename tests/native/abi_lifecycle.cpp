// The C ABI's lifecycle under AddressSanitizer and LeakSanitizer (host code
// of the library instrumented): every entry point, the error paths, engines
// created and destroyed repeatedly, a creation that fails after it has
// allocated, and device memory returned after destroy.
//
//   abi_lifecycle errors   argument and specification errors (no GPU needed)
//   abi_lifecycle full     everything, on device 0
//
// Exit status 0 when every check passes.  The program runs LeakSanitizer's
// check itself once every engine is destroyed, then leaves with _exit: the
// ROCm runtime's own teardown at exit frees host memory after the
// sanitizer's device allocator has been told that the runtime is unloaded,
// and when such a free pushes a quarantined device allocation out, the
// sanitizer's internal check aborts the process (after main has returned,
// with the program's buffered output lost).  Leaks are found the same way
// before that teardown, since nothing of the engine remains by then.
#include "jmm_engine.h"

#include <hip/hip_runtime_api.h>
#include <sanitizer/lsan_interface.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

static int g_fails = 0;

#define EXPECT(cond, ...)                                                  \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);      \
            std::fprintf(stderr, __VA_ARGS__);                             \
            std::fprintf(stderr, " [last error: %s]\n", jmm_last_error()); \
            ++g_fails;                                                     \
        }                                                                  \
    } while (0)

static jmm_chain_spec lj(uint64_t N, double L0, int npt, int periodic, uint32_t id)
{
    jmm_chain_spec s;
    std::memset(&s, 0, sizeof s);
    s.potential = JMM_POT_LJ;
    s.boundary = periodic ? JMM_PERIODIC : JMM_WALLS;
    s.ensemble = npt ? JMM_NPT : JMM_NLT;
    s.vol_move = JMM_VOL_LOG;
    s.nbn = -1;
    s.shifted = 1;
    s.N = N;
    s.T = 1.0;
    s.P = 0.5;
    s.L0 = L0;
    s.max_step = 0.5;
    s.max_dv = 0.01;
    s.cutoff = 3.0;
    s.pot_p[0] = 1.0;
    s.pot_p[1] = 1.0;
    s.seed = 1000 + id;
    s.chain_id = id;
    s.eci = 7;
    return s;
}

static jmm_chain_spec hrsw(uint64_t N, double L0, uint32_t id)
{
    jmm_chain_spec s = lj(N, L0, 1, 1, id);
    s.potential = JMM_POT_HRSW;
    s.shifted = 0;
    s.cutoff = 1.5;
    s.pot_p[0] = 1.0;
    s.pot_p[1] = 1.0;
    s.pot_p[2] = 1.5;
    return s;
}

static void errors_mode()
{
    EXPECT(jmm_abi_version() == JMM_ABI_VERSION, "abi version");
    jmm_engine *e = nullptr;
    jmm_chain_spec good = lj(64, 128.0, 0, 1, 0);
    EXPECT(jmm_create(nullptr, 1, nullptr, &e) == JMM_EINVAL && !e, "null specs");
    EXPECT(jmm_create(&good, 0, nullptr, &e) == JMM_EINVAL && !e, "no chains");
    EXPECT(jmm_create(&good, 1, nullptr, nullptr) == JMM_EINVAL, "null out");
    EXPECT(std::strlen(jmm_last_error()) > 0, "message set");
    jmm_chain_spec bad = good;
    bad.N = 0;
    EXPECT(jmm_create(&bad, 1, nullptr, &e) == JMM_EINVAL && !e, "N = 0");
    bad = good;
    bad.potential = 7;
    EXPECT(jmm_create(&bad, 1, nullptr, &e) == JMM_EINVAL && !e, "unknown potential");
    bad = good;
    bad.T = -1.0;
    EXPECT(jmm_create(&bad, 1, nullptr, &e) == JMM_EINVAL && !e, "negative T");
    jmm_chain_spec two[2] = {good, good};
    two[1].L0 = 1.0;  // 64 particles cannot fit a length of 1
    EXPECT(jmm_create(two, 2, nullptr, &e) == JMM_EINVAL && !e, "overfull chain");
    jmm_engine_opts o;
    std::memset(&o, 0, sizeof o);
    o.form = 9;
    EXPECT(jmm_create(&good, 1, &o, &e) == JMM_EINVAL && !e, "unknown form");
    uint64_t B = 0, M = 0, w = 0;
    EXPECT(jmm_schedule(&good, &B, &M, &w) == 0 && B >= 1 && M >= 2, "schedule");
    EXPECT(jmm_schedule(nullptr, &B, &M, &w) == JMM_EINVAL, "schedule null");
    EXPECT(jmm_destroy(nullptr) == 0, "destroy null");
    // calls on a null engine report errors, never crash
    uint32_t n = 0;
    double d = 0.0;
    jmm_sample smp;
    jmm_counters cnt;
    jmm_layout lay;
    EXPECT(jmm_run_sweeps(nullptr, 1) != 0, "run null");
    EXPECT(jmm_synchronize(nullptr) != 0, "sync null");
    EXPECT(jmm_full_energy(nullptr, &d) != 0, "energy null");
    EXPECT(jmm_observables_device(nullptr, nullptr, nullptr) != 0, "device observables null");
    EXPECT(jmm_read_samples(nullptr, &smp, 1, &n) != 0, "samples null");
    EXPECT(jmm_get_state(nullptr, 0, nullptr, &d, &d, nullptr) != 0, "state null");
    EXPECT(jmm_get_states(nullptr, nullptr, &d, &d) != 0, "states null");
    EXPECT(jmm_set_state(nullptr, 0, &d, 1.0, 0) != 0, "set state null");
    EXPECT(jmm_set_states(nullptr, &d, &d, nullptr, nullptr) != 0, "set states null");
    EXPECT(jmm_read_samples_virial(nullptr, &smp, &d, 1, &n) != 0, "virial samples null");
    EXPECT(jmm_get_counters(nullptr, 0, &cnt) != 0, "counters null");
    EXPECT(jmm_chain_layout(nullptr, 0, &lay) != 0, "layout null");
    EXPECT(jmm_pair_observables(nullptr, 1, nullptr, nullptr, nullptr, nullptr, 64, nullptr) != 0,
           "pair observables null");
}

static size_t device_free()
{
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return 0;
    return f;
}

// One engine through every call.
static void exercise(int form, uint32_t seg_lds)
{
    std::vector<jmm_chain_spec> specs = {lj(3000, 6000.0, 1, 1, 1), hrsw(300, 600.0, 2),
                                         lj(200, 400.0, 1, 0, 3), lj(64, 128.0, 0, 1, 4)};
    specs[2].vol_move = JMM_VOL_LINEAR;
    specs[2].max_dv = 1.0;
    specs[2].shifted = 0;
    const int n = (int)specs.size();
    jmm_engine_opts o;
    std::memset(&o, 0, sizeof o);
    o.form = form;
    o.sample_capacity = 64;
    o.block_sweeps = 8;
    o.log_accepts = 1;
    o.seg_lds = seg_lds;
    o.sample_virial = 1;
    jmm_engine *e = nullptr;
    EXPECT(jmm_create(specs.data(), n, &o, &e) == 0 && e, "create form %d", form);
    if (!e) return;
    uint64_t total = 0;
    for (auto &s : specs) total += s.N;
    jmm_layout lay;
    EXPECT(jmm_chain_layout(e, 0, &lay) == 0, "layout");
    if (seg_lds && form != JMM_FORM_PLAIN_FP64) EXPECT(lay.mode == JMM_MODE_SPLIT, "chain 0 split");
    EXPECT(jmm_chain_layout(e, n, &lay) == JMM_EINVAL, "layout range");
    EXPECT(jmm_run_sweeps(e, 20) == 0, "run");
    EXPECT(jmm_synchronize(e) == 0, "sync");
    uint64_t launches = 0, sweeps = 0;
    double ms = 0.0;
    EXPECT(jmm_kernel_time(e, &launches, &ms, &sweeps, 1) == 0 && sweeps == 20, "kernel time");
    std::vector<uint8_t> log(total * 64);
    uint32_t nlog = 0;
    EXPECT(jmm_read_accept_log(e, log.data(), 64, &nlog) == 0 && nlog == 20, "accept log");
    std::vector<jmm_sample> smp(n * 64);
    std::vector<double> vir((size_t)n * 64 * 2);
    uint32_t ns = 0;
    EXPECT(jmm_read_samples_virial(e, smp.data(), vir.data(), 64, &ns) == 0 && ns == 20, "samples");
    std::vector<double> x(total), Ls(n), Us(n);
    EXPECT(jmm_get_states(e, x.data(), Ls.data(), Us.data()) == 0, "states");
    std::vector<double> x0(specs[0].N);
    double L = 0.0, U = 0.0;
    uint64_t sw = 0;
    EXPECT(jmm_get_state(e, 0, x0.data(), &L, &U, &sw) == 0 && sw == 20, "state");
    EXPECT(jmm_set_state(e, 0, x0.data(), L, sw) == 0, "set state");
    EXPECT(jmm_set_state(e, n, x0.data(), L, sw) == JMM_EINVAL, "set state range");
    EXPECT(jmm_get_states(e, x.data(), Ls.data(), Us.data()) == 0, "states again");
    std::vector<uint64_t> sws(n, 20);
    EXPECT(jmm_set_states(e, x.data(), Ls.data(), Us.data(), sws.data()) == 0, "set states");
    EXPECT(jmm_set_states(e, x.data(), Ls.data(), nullptr, nullptr) == 0, "set states, energies recomputed");
    std::vector<double> xbad(x);
    std::swap(xbad[0], xbad[1]);
    EXPECT(jmm_set_states(e, xbad.data(), Ls.data(), nullptr, nullptr) == JMM_EINVAL, "set states unsorted");
    std::vector<jmm_counters> cnt(n);
    EXPECT(jmm_get_counters(e, -1, cnt.data()) == 0 && cnt[0].sweeps == 20, "counters");
    std::vector<double> Uf(n);
    EXPECT(jmm_full_energy(e, Uf.data()) == 0 && std::fabs(Uf[0] - cnt[0].U) <= 1e-6 * (1.0 + std::fabs(Uf[0])),
           "full energy %g vs %g", Uf[0], cnt[0].U);
    std::vector<double> obs(n * 6), grad(total);
    EXPECT(jmm_observables(e, obs.data(), grad.data()) == 0, "observables");
    {
        // the same into device arrays: equal bits
        double *od = nullptr, *gd = nullptr;
        if (hipMalloc((void **)&od, sizeof(double) * n * 6) == hipSuccess &&
            hipMalloc((void **)&gd, sizeof(double) * total) == hipSuccess) {
            EXPECT(jmm_observables_device(e, od, gd) == 0 && jmm_synchronize(e) == 0, "observables on device");
            std::vector<double> obs2(n * 6), grad2(total);
            EXPECT(hipMemcpy(obs2.data(), od, sizeof(double) * n * 6, hipMemcpyDeviceToHost) == hipSuccess &&
                       hipMemcpy(grad2.data(), gd, sizeof(double) * total, hipMemcpyDeviceToHost) == hipSuccess,
                   "observables on device: copy");
            EXPECT(std::memcmp(obs2.data(), obs.data(), sizeof(double) * n * 6) == 0 &&
                       std::memcmp(grad2.data(), grad.data(), sizeof(double) * total) == 0,
                   "observables on device equal the host copy");
        } else {
            EXPECT(false, "hipMalloc for the device observables");
        }
        if (od) (void)hipFree(od);
        if (gd) (void)hipFree(gd);
    }
    jmm_hist_spec hs;
    std::memset(&hs, 0, sizeof hs);
    hs.rhonb = 16;
    hs.rbw = 50.0;
    hs.gns = 2;
    hs.gnb = 8;
    hs.gsw = 100.0;
    hs.gbw = 0.5;
    EXPECT(jmm_hist_configure(e, &hs, 5) == 0, "hist configure");
    EXPECT(jmm_hist_sample(e) == 0, "hist sample");
    std::vector<uint64_t> rho((size_t)n * hs.rhonb), g((size_t)n * hs.gns * hs.gnb);
    uint64_t nh = 0;
    EXPECT(jmm_hist_read(e, rho.data(), g.data(), &nh, 1) == 0 && nh >= 1, "hist read");
    std::vector<double> steps(n, 0.3), dvs(n, 0.005);
    EXPECT(jmm_set_steps(e, steps.data(), dvs.data()) == 0, "set steps");
    std::vector<double> Lt(Ls), Usc(n);
    for (auto &v : Lt) v *= 1.001;
    EXPECT(jmm_scaled_energy(e, Lt.data(), Usc.data()) == 0, "scaled energy");
    EXPECT(jmm_set_lengths(e, Lt.data()) == 0, "set lengths");
    EXPECT(jmm_perturb_energy(e, 1, 0.25) == 0, "perturb");
    EXPECT(jmm_run_sweeps(e, 40) == 0, "run more");
    EXPECT(jmm_run_sweeps(e, 30) == JMM_EFULL, "sample buffer full");
    EXPECT(jmm_read_samples(e, smp.data(), 64, &ns) == 0 && ns == 40, "drain");
    EXPECT(jmm_run_sweeps(e, 8) == 0, "run after drain");
    EXPECT(jmm_destroy(e) == 0, "destroy");
}

static void pair_observables()
{
    const int nc = 3;
    jmm_chain_spec s = lj(128, 256.0, 0, 1, 9);
    std::vector<double> x(nc * s.N), L(nc, s.L0);
    for (int c = 0; c < nc; ++c)
        for (uint64_t i = 0; i < s.N; ++i) x[c * s.N + i] = (((double)i + 0.5) / (double)s.N - 0.5) * s.L0;
    double *dx = nullptr, *dL = nullptr, *dobs = nullptr, *dgrad = nullptr;
    EXPECT(hipMalloc((void **)&dx, sizeof(double) * x.size()) == hipSuccess, "alloc x");
    EXPECT(hipMalloc((void **)&dL, sizeof(double) * nc) == hipSuccess, "alloc L");
    EXPECT(hipMalloc((void **)&dobs, sizeof(double) * nc * 6) == hipSuccess, "alloc obs");
    EXPECT(hipMalloc((void **)&dgrad, sizeof(double) * x.size()) == hipSuccess, "alloc grad");
    if (dx && dL && dobs && dgrad) {
        (void)hipMemcpy(dx, x.data(), sizeof(double) * x.size(), hipMemcpyHostToDevice);
        (void)hipMemcpy(dL, L.data(), sizeof(double) * nc, hipMemcpyHostToDevice);
        for (int prec : {32, 64})
            EXPECT(jmm_pair_observables(&s, nc, dx, dL, dobs, dgrad, prec, nullptr) == 0, "pair obs %d", prec);
        EXPECT(jmm_pair_observables(&s, nc, dx, dL, dobs, dgrad, 16, nullptr) == JMM_EINVAL, "precision 16");
        EXPECT(hipDeviceSynchronize() == hipSuccess, "sync");
    }
    (void)hipFree(dx);
    (void)hipFree(dL);
    (void)hipFree(dobs);
    (void)hipFree(dgrad);
}

// Creation that fails after allocating: more split workgroups than the
// device holds at once.
static void failing_create()
{
    std::vector<jmm_chain_spec> many(2000, lj(4096, 8192.0, 1, 1, 0));
    for (size_t c = 0; c < many.size(); ++c) many[c].chain_id = (uint32_t)c;
    jmm_engine_opts o;
    std::memset(&o, 0, sizeof o);
    o.form = JMM_FORM_BATCHED_FP32;
    o.seg_lds = 8 * (2 * 7 + 16 * 17);
    jmm_engine *e = nullptr;
    EXPECT(jmm_create(many.data(), (int)many.size(), &o, &e) == JMM_EINVAL && !e,
           "too many split workgroups");
    jmm_chain_spec s = lj(64, 128.0, 0, 1, 0);
    std::memset(&o, 0, sizeof o);
    o.device = 4096;
    EXPECT(jmm_create(&s, 1, &o, &e) == JMM_EINVAL && !e, "device out of range");
}

static void cycle()
{
    exercise(JMM_FORM_BATCHED_FP32, 0);
    exercise(JMM_FORM_BATCHED_FP32, 8 * (2 * 7 + 16 * 13));  // chain 0 split into segments
    exercise(JMM_FORM_BATCHED_FP64, 0);
    exercise(JMM_FORM_PLAIN_FP64, 0);
    pair_observables();
    failing_create();
}

static void full_mode()
{
    errors_mode();
    // two cycles first: the runtime's own device allocations (code objects,
    // scratch backing, pools) happen here and are kept by the runtime.
    // Memory the engine failed to return would grow with every later cycle:
    // six more cycles may cost at most one 2 MiB allocation granule.
    cycle();
    cycle();
    const size_t m0 = device_free();
    for (int rep = 0; rep < 3; ++rep) cycle();
    const size_t m1 = device_free();
    for (int rep = 0; rep < 3; ++rep) cycle();
    const size_t m2 = device_free();
    std::printf("device free memory after 2, 5 and 8 cycles: %zu %zu %zu\n", m0, m1, m2);
    EXPECT(m2 + (2u << 20) >= m0, "device memory not returned: %zu bytes over six cycles", m0 - m2);
}

int main(int argc, char **argv)
{
    const char *mode = argc > 1 ? argv[1] : "errors";
    if (!std::strcmp(mode, "full")) full_mode();
    else errors_mode();
    std::printf("%s: %d failed checks\n", mode, g_fails);
    std::fflush(stdout);
    __lsan_do_leak_check();     // reports and exits non-zero on unsuppressed leaks
    std::fflush(stderr);
    _exit(g_fails ? 1 : 0);
}
