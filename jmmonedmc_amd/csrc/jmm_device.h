// jmm_device.h -- arithmetic shared by the gfx950 kernels and the engine's
// host set-up code: Philox4x32-10, 53-bit uniforms, the deterministic exp of
// the log-length move, and the pair potentials in two forms:
//   phi64  FP64, one IEEE rounding per operation in a fixed order (the file
//          is compiled with -ffp-contract=off), so positions, per-pair energies
//          and per-trial energy changes are reproducible bit for bit;
//   phi32  FP32 pair arithmetic on an FP64-formed separation (the caller
//          subtracts positions in FP64 and decides the cutoff in FP64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define JMM_HD __host__ __device__ __forceinline__

namespace jmm {

enum Kind : int { POT_LJ = 0, POT_HARMONIC = 1, POT_HRSW = 2, POT_YUKAWA = 3 };

// Philox4x32-10 (Salmon et al., SC'11).  ctr = (slot, sweep, chain_id, 0),
// key = seed.  Replaces the reference's GSL taus2 stream
// (src/jmmMCState.cpp:779-781, draws at :1762-1763, :1370, :2252).
JMM_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                          uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// 53-bit uniform in [0, 1) from the high 27 bits of a and 26 bits of b.
JMM_HD double u53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Uniform in (0, 1) from one word: (w + 1/2) 2^-32, exact in FP64.
JMM_HD double u32o(uint32_t w)
{
    return ((double)w + 0.5) * (1.0 / 4294967296.0);
}

// exp for the log-length move's scale factor, built only from IEEE
// add/mul/floor so host and device agree bit for bit: Cody-Waite reduction
// x = k ln2 + r with a two-constant ln2, degree-13 Taylor polynomial in
// Horner form, exact 2^k scaling.
JMM_HD double det_exp(double x)
{
    if (x > 709.0) return INFINITY;
    if (x < -708.0) return 0.0;
    const double inv_ln2 = 1.4426950408889634;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double kd = floor(x * inv_ln2 + 0.5);
    const double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
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
    const int64_t k = (int64_t)kd;
    union {
        uint64_t u;
        double d;
    } sc;
    sc.u = (uint64_t)(k + 1023) << 52;
    return p * sc.d;
}

// Per-chain potential constants, prepared once on the host.
struct Pot {
    double rc;     // cutoff (inclusive for LJ/Yukawa, strict for harmonic/HRSW)
    double shift;  // phi(rc) for the shifted form, else 0
    double e4;     // LJ: 4 eps
    double sig6;   // LJ: sigma^6
    double p0, p1, p2;
    float rcf, shiftf, e4f, sig6f, p0f, p1f, p2f;
    float a12f, a6f;  // LJ: 4 eps sigma^12 and 4 eps sigma^6 (FP32 forms)
    float pad;
};

// FP64 pair energy, operation for operation the oracle's sequence
// (LJ: src/pot.cpp:47-56 generalised to eps, sigma; harmonic:
// src/pot.cpp:116-131).
template <int KIND>
JMM_HD double phi64(const Pot &q, double d)
{
    if (KIND == POT_LJ) {
        if (!(d <= q.rc)) return 0.0;
        const double d3 = d * d * d;
        const double i6 = (1.0 / (d3 * d3)) * q.sig6;
        const double i12 = i6 * i6;
        return (q.e4 * i12 - q.e4 * i6) - q.shift;
    } else if (KIND == POT_HARMONIC) {
        if (d <= 0.0) return 10e10;
        if (d < q.rc) {
            const double m = d - 1.0;
            return m * m;
        }
        return 0.0;
    } else if (KIND == POT_HRSW) {
        if (d < q.p0) return INFINITY;
        return d < q.p2 ? -q.p1 : 0.0;
    } else {
        if (d < q.p2) return INFINITY;
        if (!(d <= q.rc)) return 0.0;
        return (q.p0 * (exp(-q.p1 * d) / d)) - q.shift;
    }
}

// Virial W = -d phi'(d) and hypervirial H = -d W'(d) of one pair inside the
// cutoff, FP64 (the reference's phi[1] and phi[6], src/pot.cpp:61-66, for
// LJ: 48/d^12 - 24/d^6 and 576/d^12 - 144/d^6 at eps = sigma = 1).  The hard
// rod + square well has no finite virial away from contact: zero.
template <int KIND>
JMM_HD void virial64(const Pot &q, double d, double &W, double &H)
{
    if (KIND == POT_LJ) {
        const double d3 = d * d * d;
        const double i6 = (1.0 / (d3 * d3)) * q.sig6;
        const double a12 = q.e4 * (i6 * i6), a6 = q.e4 * i6;
        W = 12.0 * a12 - 6.0 * a6;
        H = 144.0 * a12 - 36.0 * a6;
    } else if (KIND == POT_YUKAWA) {
        const double E = q.p0 * (exp(-q.p1 * d) / d);
        const double kd = q.p1 * d;
        W = E * (kd + 1.0);
        H = E * ((kd + 1.0) * (kd + 1.0) - kd);
    } else if (KIND == POT_HARMONIC) {
        W = -2.0 * d * (d - 1.0);
        H = 4.0 * d * d - 2.0 * d;
    } else {
        W = 0.0;
        H = 0.0;
    }
}

// Inside-the-cutoff test in FP64 (the FP32 form uses it so that both forms
// sum the same pairs).
template <int KIND>
JMM_HD bool inside64(const Pot &q, double d)
{
    if (KIND == POT_HARMONIC || KIND == POT_HRSW) return d < (KIND == POT_HRSW ? q.p2 : q.rc);
    return d <= q.rc;
}

// Hardware reciprocal and exponential on the device (v_rcp_f32, v_exp_f32);
// plain library forms on the host.
JMM_HD float frcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

JMM_HD float fexp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __expf(x);
#else
    return expf(x);
#endif
}

// FP32 pair energy of a separation d whose cutoff test `in` the caller has
// made; evaluated without branches and selected, so that a wavefront keeps
// one instruction stream.  Hard cores are handled before any pair is
// evaluated.
template <int KIND>
JMM_HD float phi32(const Pot &q, float d, bool in)
{
    float v;
    if (KIND == POT_LJ) {
        // 4 eps ((sigma/d)^12 - (sigma/d)^6) - shift as
        // s6 (4 eps sigma^12 s6 - 4 eps sigma^6) - shift with s6 = d^-6, two FMAs
        const float ir2 = frcp(d * d);
        const float s6 = ir2 * ir2 * ir2;
        v = fmaf(s6, fmaf(q.a12f, s6, -q.a6f), -q.shiftf);
    } else if (KIND == POT_HARMONIC) {
        const float m = d - 1.0f;
        v = d <= 0.0f ? 10e10f : m * m;
    } else if (KIND == POT_HRSW) {
        v = -q.p1f;
    } else {
        v = q.p0f * fexp(-q.p1f * d) * frcp(d) - q.shiftf;
    }
    return in ? v : 0.0f;
}

// Per-trial constants of the FP32 box-length pass (a scaling of every
// separation by s), rounded from FP64: for LJ c12 = 4 eps sigma^12
// (s^-12 - 1), c6 = 4 eps sigma^6 (s^-6 - 1), so that phi(s d) - phi(d) =
// s6 (c12 s6 - c6) with s6 = d^-6 (one reciprocal per pair, no
// cancellation).
struct BoxK {
    float c12, c6;
};

// Change of one pair's FP32 energy when its separation goes from d0 to d1
// (d1 = s d0 up to the FP64 roundings of the scaled positions), given the
// FP64 inside-the-cutoff flags of both: a pair entering the cutoff adds
// phi(d1) - shift, a pair leaving it removes phi(d0) - shift, a pair inside
// at both changes by phi(d1) - phi(d0): for LJ the scaling identity's
// difference, for Yukawa and the harmonic well the difference of the two
// FP32 energies (two exponentials per Yukawa pair cost less than a series
// for the bracket, measured).  e0 receives the pair's energy at d0
// (phi(d0) - shift, zero outside the cutoff), so the same pass sums the
// configuration's energy.  The hard rod + square well is handled by the
// callers (exact counts).
template <int KIND>
JMM_HD float box_term32(const Pot &q, const BoxK &bk, float f0, float f1, bool in0, bool in1, float &e0)
{
    float dlt, r0;
    if (KIND == POT_LJ) {
        const float ir2 = frcp(f0 * f0);
        const float s6 = ir2 * ir2 * ir2;
        dlt = s6 * fmaf(bk.c12, s6, -bk.c6);
        r0 = fmaf(s6, fmaf(q.a12f, s6, -q.a6f), -q.shiftf);
    } else if (KIND == POT_YUKAWA) {
        const float v0 = q.p0f * fexp(-q.p1f * f0) * frcp(f0) - q.shiftf;
        const float v1 = q.p0f * fexp(-q.p1f * f1) * frcp(f1) - q.shiftf;
        e0 = in0 ? v0 : 0.0f;
        return (in1 ? v1 : 0.0f) - e0;
    } else {
        const float m0 = f0 - 1.0f, m1 = f1 - 1.0f;
        dlt = m1 * m1 - m0 * m0;
        r0 = m0 * m0;
    }
    e0 = in0 ? r0 : 0.0f;
    const float enter = in1 ? r0 : -r0;
    return (in1 ? dlt : 0.0f) + (in0 != in1 ? enter : 0.0f);
}

// FP32 inside-the-cutoff test (strict for harmonic and the square well,
// inclusive otherwise).
template <int KIND>
JMM_HD bool inside32(const Pot &q, float d)
{
    if (KIND == POT_HRSW) return d < q.p2f;
    if (KIND == POT_HARMONIC) return d < q.rcf;
    return d <= q.rcf;
}

// Pair observables for the backward kernel: E, W = -d phi', dE/dtheta0,
// dE/dtheta1, dW/dtheta0, dW/dtheta1 (LJ theta = (eps, sigma); Yukawa
// theta = (A, kappa); hard-core deltas excluded).  `sh` holds the shifted
// form's values at the cutoff (E, dE/dtheta0, dE/dtheta1), zero otherwise.
template <int KIND, typename R>
JMM_HD void pair_obs(const Pot &q, R d, const R sh[3], R o[6])
{
    for (int k = 0; k < 6; ++k) o[k] = R(0);
    if (KIND == POT_LJ) {
        const R d3 = d * d * d;
        const R i6 = (R(1) / (d3 * d3)) * R(q.sig6);
        const R i12 = i6 * i6;
        const R e4 = R(q.e4);
        const R a12 = e4 * i12, a6 = e4 * i6;
        const R sig = R(q.p1);
        o[0] = (a12 - a6) - sh[0];
        o[1] = R(12) * a12 - R(6) * a6;
        o[2] = (R(4) * i12 - R(4) * i6) - sh[1];
        o[3] = o[1] / sig - sh[2];
        o[4] = R(12) * (R(4) * i12) - R(6) * (R(4) * i6);
        o[5] = (R(144) * a12 - R(36) * a6) / sig;
    } else if (KIND == POT_YUKAWA) {
        const R A = R(q.p0), k = R(q.p1);
        const R e = exp(-k * d);
        const R ed = e / d;
        const R E = A * ed;
        o[0] = E - sh[0];
        o[1] = E * (k * d + R(1));
        o[2] = ed - sh[1];
        o[3] = -A * e - sh[2];
        o[4] = ed * (k * d + R(1));
        o[5] = -k * d * d * E;
    } else if (KIND == POT_HARMONIC) {
        const R m = d - R(1);
        o[0] = m * m;
        o[1] = R(-2) * d * m;
    } else {
        o[0] = R(-q.p1);
        o[2] = R(-1);
    }
}

}  // namespace jmm
