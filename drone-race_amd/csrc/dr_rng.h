// Counter-based noise for the MPPI rollouts: Philox4x32-10 and a Box-Muller
// transform built only from IEEE-754 basic operations (add, mul, fma,
// correctly rounded div and sqrt) so that host code compiled with
// -ffp-contract=off produces the same bits as the gfx950 device code.
//
// The reference draws its sampling noise with cuRAND XORWOW
// (src/kernels.cu:10-31, include/state.h:11-19).  That stream lives in a
// closed library and is not reproduced; this engine uses a stateless
// counter-based generator instead, so no per-sample RNG state is stored in
// HBM and two rollouts of the same (key, counter) see the same noise.
//
//   Philox4x32-10: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
//   "Parallel Random Numbers: As Easy as 1, 2, 3", SC'11.
//   Box-Muller: G. E. P. Box, M. E. Muller, Ann. Math. Stat. 29 (1958).
#pragma once

#include <stdint.h>

#ifndef DR_HD
#if defined(__HIPCC__)
#define DR_HD __host__ __device__ __forceinline__
#else
#define DR_HD inline
#endif
#endif

namespace dr {

// Purposes of the Philox streams (fourth counter word).
enum : uint32_t
{
    STREAM_CONTROL = 0u, // control sampling noise, counter (k, t, step, 0)
    STREAM_TRACK = 1u,   // scenario generation (gate jitter, radii, sizes)
    STREAM_BRANCH = 2u,  // PR-MPPI branch beliefs and opponent velocity shifts
    STREAM_MODEL = 3u,   // PR-MPPI opponent-model draws, counter (m, 0, step, 3)
    STREAM_ENV = 4u,     // environment noise of a 2-D race rollout, counter (k, 4 t + block, step, 4)
    STREAM_VERIFY = 5u,  // plan verification: model draw (s | set << 30, 0xFFFFFFFF, step, 5),
                         // environment noise (s | set << 30, 4 t + block, step, 5)
};

struct U4
{
    uint32_t x, y, z, w;
};

DR_HD uint32_t f2u(float f)
{
    union { float f; uint32_t u; } c;
    c.f = f;
    return c.u;
}

DR_HD float u2f(uint32_t u)
{
    union { float f; uint32_t u; } c;
    c.u = u;
    return c.f;
}

// a ^ b ^ c.  On the device one three-input bitwise instruction
// (v_bitop3_b32 with the truth table of a three-way XOR, 0x96) instead of two
// XORs; on the host two XORs.  The same bits either way.
DR_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
#else
    return a ^ b ^ c;
#endif
}

DR_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r)
    {
        if (r > 0)
        {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        U4 n;
        // the first round's words may be uniform across a wave (the counter
        // words of a step): plain XORs there keep them in scalar registers
        n.x = r == 0 ? hi1 ^ c.y ^ k0 : xor3(hi1, c.y, k0);
        n.y = lo1;
        n.z = r == 0 ? hi0 ^ c.w ^ k1 : xor3(hi0, c.w, k1);
        n.w = lo0;
        c = n;
    }
    return c;
}

// (0, 1]: (top 24 bits + 1) * 2^-24, exact in binary32.
DR_HD float uniform_open0(uint32_t x)
{
    return (float)((x >> 8) + 1u) * 5.9604644775390625e-08f;
}

// [0, 1): top 24 bits * 2^-24, exact in binary32.
DR_HD float uniform01(uint32_t x)
{
    return (float)(x >> 8) * 5.9604644775390625e-08f;
}

// Natural log for normal positive x.  x = 2^e * m with m in [sqrt(1/2), sqrt(2));
// ln m = 2 atanh(s), s = (m - 1) / (m + 1), series to s^11 (|s| < 0.1716,
// truncation < 2e-9 relative).  ln 2 is split into a short high part so that
// e * ln2_hi is exact.
DR_HD float logf_ieee(float x)
{
    uint32_t b = f2u(x);
    int e = (int)(b >> 23) - 127;
    float m = u2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f)
    {
        m = m * 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float q = 0.0909090936f;         // 1/11
    q = __builtin_fmaf(q, z, 0.111111112f); // 1/9
    q = __builtin_fmaf(q, z, 0.142857149f); // 1/7
    q = __builtin_fmaf(q, z, 0.200000003f); // 1/5
    q = __builtin_fmaf(q, z, 0.333333343f); // 1/3
    const float r = (s * z) * q;
    const float lnm = 2.0f * (s + r);
    const float fe = (float)e;
    return __builtin_fmaf(fe, 0.693145752f, __builtin_fmaf(fe, 1.42860677e-06f, lnm));
}

// sin and cos of 2*pi*k*2^-24 for a 24-bit integer k.  The nearest quarter
// turn is split off exactly in integer arithmetic; the remainder angle
// a in [-pi/4, pi/4) goes through Taylor polynomials (sin to a^9, cos to
// a^10; truncation below 3e-9).
DR_HD void sincos_turns24(uint32_t k, float* s_out, float* c_out)
{
    const uint32_t qu = (k + (1u << 21)) >> 22;                 // round(4 k / 2^24)
    const int32_t rem = (int32_t)k - (int32_t)(qu << 22);        // [-2^21, 2^21)
    const float a = (float)rem * 3.74507028e-07f;                // pi/2 * 2^-22
    const float a2 = a * a;
    float ps = 2.75573188e-06f;                                  // 1/9!
    ps = __builtin_fmaf(ps, a2, -1.98412701e-04f);                // -1/7!
    ps = __builtin_fmaf(ps, a2, 8.33333377e-03f);                 // 1/5!
    ps = __builtin_fmaf(ps, a2, -0.166666672f);                   // -1/3!
    const float sn = __builtin_fmaf(a * a2, ps, a);
    float pc = -2.75573188e-07f;                                 // -1/10!
    pc = __builtin_fmaf(pc, a2, 2.48015876e-05f);                 // 1/8!
    pc = __builtin_fmaf(pc, a2, -1.38888892e-03f);                // -1/6!
    pc = __builtin_fmaf(pc, a2, 4.16666679e-02f);                 // 1/4!
    pc = __builtin_fmaf(pc, a2, -0.5f);
    const float cs = __builtin_fmaf(a2, pc, 1.0f);
    // rotate by the quarter turns: odd quarters swap (sin, cos) -> (cos, -sin),
    // quarters 2 and 3 negate both (selects, no branches)
    const bool odd = (qu & 1u) != 0u, half = (qu & 2u) != 0u;
    const float s1 = odd ? cs : sn;
    const float c1 = odd ? -sn : cs;
    *s_out = half ? -s1 : s1;
    *c_out = half ? -c1 : c1;
}

// Two standard normals from two 32-bit words (Box-Muller).
DR_HD void box_muller(uint32_t a, uint32_t b, float* z0, float* z1)
{
    const float u1 = uniform_open0(a);
    const float r = __builtin_sqrtf(-2.0f * logf_ieee(u1));
    float s, c;
    sincos_turns24(b >> 8, &s, &c);
    *z0 = r * c;
    *z1 = r * s;
}

// Four standard normals for (key, counter).
DR_HD void normal4(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, float z[4])
{
    U4 c;
    c.x = c0;
    c.y = c1;
    c.z = c2;
    c.w = c3;
    const U4 o = philox4x32_10(c, k0, k1);
    box_muller(o.x, o.y, &z[0], &z[1]);
    box_muller(o.z, o.w, &z[2], &z[3]);
}

// exp(x) for the MPPI weights, x <= 0 in practice.  Cody-Waite reduction
// x = n ln2 + r, |r| <= ln2/2, degree-7 Taylor polynomial for e^r
// (truncation < 3e-9), scale by 2^n through the exponent bits.  Results below
// e^-86 (about 4e-38, just above FLT_MIN) are flushed to zero so that no
// subnormal is ever produced; x > 88 gives +inf.
DR_HD float expf_ieee(float x)
{
    if (x < -86.0f)
        return 0.0f;
    if (x > 88.0f)
        return __builtin_inff();
    const float nf = __builtin_rintf(x * 1.44269502f);
    const float r = __builtin_fmaf(nf, -1.42860677e-06f, __builtin_fmaf(nf, -0.693145752f, x));
    float p = 1.98412701e-04f;               // 1/7!
    p = __builtin_fmaf(p, r, 1.38888892e-03f); // 1/6!
    p = __builtin_fmaf(p, r, 8.33333377e-03f); // 1/5!
    p = __builtin_fmaf(p, r, 4.16666679e-02f); // 1/4!
    p = __builtin_fmaf(p, r, 0.166666672f);    // 1/3!
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    const int n = (int)nf;
    return p * u2f((uint32_t)(n + 127) << 23);
}

// Inverse CDF of the belief at u in (0, 1] (prmppi_kernels.cu:34-47,
// sampleModelFromBelief state.h:118-131): running FP32 sum, first i with
// u <= sum, else the last model.
DR_HD int model_from_belief(float u, const float* belief, int n_branches)
{
    float acc = 0.0f;
    for (int i = 0; i < n_branches; ++i)
    {
        acc = acc + belief[i];
        if (u <= acc)
            return i;
    }
    return n_branches - 1;
}

// Opponent model m of a sampled PR-MPPI step.
DR_HD int sample_branch(uint32_t k0, uint32_t k1, uint32_t m, uint32_t step, const float* belief, int n_branches)
{
    const U4 r = philox4x32_10(U4{m, 0u, step, STREAM_MODEL}, k0, k1);
    return model_from_belief(uniform_open0(r.x), belief, n_branches);
}

} // namespace dr
