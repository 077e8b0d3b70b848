// ORACLE — test infrastructure only.
//
// CPU restatement of the MPPI control step, used by tests/, by the smoke
// check and by bench.py's CPU-baseline leg, and never by the product path.
// It is compiled with g++ (-ffp-contract=off) into oracle/liboracle.so.
//
// Templated on the scalar type: S = float reproduces the GPU engine's FP32
// arithmetic operation by operation (same noise bits, same dynamics bits);
// S = double is the FP64 oracle whose distance to the FP32 results is
// reported.  The noise itself (Philox4x32-10 + Box-Muller) is always
// generated in FP32, so both oracles consume the same samples.
//
// Reference logic restated (pablo-gr-rey/drone-race):
//   rollout loop, discount, terminal cost    src/kernels.cu:100-155
//   running / terminal cost structure        include/costs.cuh:7-39
//   weights, masked min (plain case)         src/kernels.cu:221-301, 307-408
//   weighted update nom += sum(w eps)/sum(w) src/kernels.cu:376-403
//   L2/box clamp of the nominal              src/kernels.cu:434-456
//   action pick and shift with zero fill     src/mppi_controller.cu:392-397, 430-436
//   temperature adaptation                   src/mppi_controller.cu:413-424
//   gate crossing and race progress          include/environments/env_dronerace.h:374-394, 412-470
//   PR-MPPI step (oracle_pr.cpp)             src/prmppi_kernels.cu:23-393, src/prmppi_controller.cu:142-349,
//                                            include/costs.cuh:44-86
// The quadrotor dynamics have no reference counterpart; they are pinned by
// the physical invariants tested in tests/test_oracle_quad.py.
//
// Parity pins for each part (tests/):
//   Philox4x32-10        Random123 known-answer vectors (tests/golden/philox_kat.json)
//   log/exp/sincos       libm in FP64 over dense sweeps, <= 2 ulp
//   Box-Muller normals   closed form from libm; moments and KS test
//   Quad13 step          hover equilibrium, free fall, quaternion norm, zero-rate attitude
//   gate crossing        hand-derived crossing / miss / margin cases
//   running cost         closed-form values at chosen distances
//   MPPI update          brute-force FP64 softmax on tiny K; sum of weights; min sample w = 1
#pragma once

#include <math.h>
#include <stdint.h>

#include "../include/dr_api.h"

namespace orc {

// ---------------------------------------------------------------- noise (FP32)

struct W4
{
    uint32_t v[4];
};

inline W4 philox(W4 c, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    for (int round = 0; round < 10; ++round)
    {
        if (round)
        {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t a = (uint64_t)M0 * c.v[0];
        const uint64_t b = (uint64_t)M1 * c.v[2];
        W4 n;
        n.v[0] = (uint32_t)(b >> 32) ^ c.v[1] ^ k0;
        n.v[1] = (uint32_t)b;
        n.v[2] = (uint32_t)(a >> 32) ^ c.v[3] ^ k1;
        n.v[3] = (uint32_t)a;
        c = n;
    }
    return c;
}

inline uint32_t bits(float f)
{
    union { float f; uint32_t u; } x;
    x.f = f;
    return x.u;
}

inline float from_bits(uint32_t u)
{
    union { float f; uint32_t u; } x;
    x.u = u;
    return x.f;
}

// ln x for normal x > 0 via ln m = 2 atanh((m-1)/(m+1)), m in [sqrt(1/2), sqrt(2)).
inline float log32(float x)
{
    const uint32_t b = bits(x);
    int ex = (int)(b >> 23) - 127;
    float m = from_bits((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f)
    {
        m = m * 0.5f;
        ex = ex + 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float q = fmaf(0.0909090936f, z, 0.111111112f);
    q = fmaf(q, z, 0.142857149f);
    q = fmaf(q, z, 0.200000003f);
    q = fmaf(q, z, 0.333333343f);
    const float r = (s * z) * q;
    const float lnm = 2.0f * (s + r);
    const float fe = (float)ex;
    return fmaf(fe, 0.693145752f, fmaf(fe, 1.42860677e-06f, lnm));
}

// sin, cos of 2 pi k / 2^24.
inline void sincos_turns(uint32_t k, float* sn, float* cs)
{
    const uint32_t quarter = (k + 0x200000u) >> 22;
    const int32_t rem = (int32_t)k - (int32_t)(quarter << 22);
    const float a = (float)rem * 3.74507028e-07f;
    const float a2 = a * a;
    float ps = fmaf(2.75573188e-06f, a2, -1.98412701e-04f);
    ps = fmaf(ps, a2, 8.33333377e-03f);
    ps = fmaf(ps, a2, -0.166666672f);
    const float s = fmaf(a * a2, ps, a);
    float pc = fmaf(-2.75573188e-07f, a2, 2.48015876e-05f);
    pc = fmaf(pc, a2, -1.38888892e-03f);
    pc = fmaf(pc, a2, 4.16666679e-02f);
    pc = fmaf(pc, a2, -0.5f);
    const float c = fmaf(a2, pc, 1.0f);
    const uint32_t qd = quarter & 3u;
    *sn = qd == 0u ? s : (qd == 1u ? c : (qd == 2u ? -s : -c));
    *cs = qd == 0u ? c : (qd == 1u ? -s : (qd == 2u ? -c : s));
}

inline void normal_pair(uint32_t a, uint32_t b, float* z0, float* z1)
{
    const float u1 = (float)((a >> 8) + 1u) * 5.9604644775390625e-08f;
    const float r = sqrtf(-2.0f * log32(u1));
    float s, c;
    sincos_turns(b >> 8, &s, &c);
    *z0 = r * c;
    *z1 = r * s;
}

inline void normals4(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, float z[4])
{
    W4 c = {{c0, c1, c2, c3}};
    const W4 o = philox(c, k0, k1);
    normal_pair(o.v[0], o.v[1], &z[0], &z[1]);
    normal_pair(o.v[2], o.v[3], &z[2], &z[3]);
}

// exp for x <= 88: Cody-Waite reduction, degree-7 Taylor, flush below e^-86.
inline float exp32(float x)
{
    if (x < -86.0f)
        return 0.0f;
    if (x > 88.0f)
        return INFINITY;
    const float n = rintf(x * 1.44269502f);
    const float r = fmaf(n, -1.42860677e-06f, fmaf(n, -0.693145752f, x));
    float p = fmaf(1.98412701e-04f, r, 1.38888892e-03f);
    p = fmaf(p, r, 8.33333377e-03f);
    p = fmaf(p, r, 4.16666679e-02f);
    p = fmaf(p, r, 0.166666672f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return p * from_bits((uint32_t)((int)n + 127) << 23);
}

// ------------------------------------------------------------- Quad13 model

template <class S>
inline S sq_root(S x)
{
    return sqrt(x);
}
template <>
inline float sq_root<float>(float x)
{
    return sqrtf(x);
}

// fused multiply-add, one rounding (the engine's fmaf)
inline float fm(float a, float b, float c) { return fmaf(a, b, c); }
inline double fm(double a, double b, double c) { return fma(a, b, c); }

// 1/sqrt(x) of the attitude renormalisation.  FP32: the engine's estimate
// from the bit pattern (Lomont's constant 0x5f375a86, Lomont 2003) and three
// Newton-Raphson steps in correction form y <- y + y (1/2 - (x/2) y^2), in
// the same operations; FP64: the exact value.
template <class S>
inline S inv_sqrt(S x)
{
    return (S)1 / sq_root(x);
}
template <>
inline float inv_sqrt<float>(float x)
{
    const float hx = 0.5f * x;
    float y = from_bits(0x5f375a86u - (bits(x) >> 1));
    y = fmaf(y, fmaf(-hx, y * y, 0.5f), y);
    y = fmaf(y, fmaf(-hx, y * y, 0.5f), y);
    y = fmaf(y, fmaf(-hx, y * y, 0.5f), y);
    return y;
}

template <class S>
inline S mn(S a, S b)
{
    return b < a ? b : a;
}
template <class S>
inline S mx(S a, S b)
{
    return a < b ? b : a;
}

template <class S>
struct Model
{
    S dt, half_dt, rate_k, inv_mass, gravity, drag, tmin, tmax, rmax, radius;

    // Constants are formed in FP32 first (as the engine does on the host),
    // then widened.
    explicit Model(const dr_quad_params& P)
    {
        const float h = 0.5f * P.dt, rk = P.dt * P.rate_gain, im = 1.0f / P.mass;
        dt = (S)P.dt;
        half_dt = (S)h;
        rate_k = (S)rk;
        inv_mass = (S)im;
        gravity = (S)P.gravity;
        drag = (S)P.drag;
        tmin = (S)P.thrust_min;
        tmax = (S)P.thrust_max;
        rmax = (S)P.rate_max;
        radius = (S)P.drone_radius;
    }
};

template <class S>
struct St
{
    S p[3], q[4], v[3], w[3];
    int gate, laps;
};

template <class S>
inline St<S> load_state(const dr_drone_state& d)
{
    St<S> s;
    for (int i = 0; i < 3; ++i)
    {
        s.p[i] = (S)d.p[i];
        s.v[i] = (S)d.v[i];
        s.w[i] = (S)d.w[i];
    }
    for (int i = 0; i < 4; ++i)
        s.q[i] = (S)d.q[i];
    s.gate = d.gate;
    s.laps = d.laps;
    return s;
}

template <class S>
inline S clamp(S x, S lo, S hi)
{
    return mn(mx(x, lo), hi);
}

template <class S>
inline void clamp_u(const Model<S>& M, const S in[4], S out[4])
{
    out[0] = clamp(in[0], M.tmin, M.tmax);
    for (int d = 1; d < 4; ++d)
        out[d] = clamp(in[d], -M.rmax, M.rmax);
}

template <class S>
inline void euler(const Model<S>& M, St<S>& s, const S u[4])
{
    const S qw = s.q[0], qx = s.q[1], qy = s.q[2], qz = s.q[3];
    const S wx = s.w[0], wy = s.w[1], wz = s.w[2];
    const S two = (S)2, one = (S)1;
    // R(q) e3
    const S bx = two * fm(qx, qz, qw * qy);
    const S by = two * fm(qy, qz, -(qw * qx));
    const S bz = fm(-two, fm(qx, qx, qy * qy), one);
    const S at = u[0] * M.inv_mass;
    const S acc[3] = {fm(at, bx, -(M.drag * s.v[0])), fm(at, by, -(M.drag * s.v[1])), fm(at, bz, -(M.drag * s.v[2])) - M.gravity};
    for (int i = 0; i < 3; ++i)
        s.p[i] = fm(M.dt, s.v[i], s.p[i]);
    for (int i = 0; i < 3; ++i)
        s.v[i] = fm(M.dt, acc[i], s.v[i]);
    // q <- q (x) (1, h w) / |(1, h w)| * (3 - |q|^2)/2, h = dt/2 (the Euler step q + h q (x) (0, w) renormalised)
    const S hx = M.half_dt * wx, hy = M.half_dt * wy, hz = M.half_dt * wz;
    const S rs = inv_sqrt(fm(hx, hx, fm(hy, hy, fm(hz, hz, one))));
    const S n[4] = {fm(-qx, hx, fm(-qy, hy, fm(-qz, hz, qw))), fm(qw, hx, fm(qy, hz, fm(-qz, hy, qx))),
                    fm(qw, hy, fm(qz, hx, fm(-qx, hz, qy))), fm(qw, hz, fm(qx, hy, fm(-qy, hx, qz)))};
    const S q2 = fm(qw, qw, fm(qx, qx, fm(qy, qy, qz * qz)));
    const S sc = rs * fm((S)-0.5f, q2, (S)1.5f);
    for (int i = 0; i < 4; ++i)
        s.q[i] = n[i] * sc;
    s.w[0] = fm(M.rate_k, u[1] - wx, wx);
    s.w[1] = fm(M.rate_k, u[2] - wy, wy);
    s.w[2] = fm(M.rate_k, u[3] - wz, wz);
}

// env_dronerace.h:412-470 in 3-D, division-free: lambda = num/den in [0, 1]
// iff num, den share a sign and |num| <= |den|; the crossing point x obeys
// den (x - c) = num (p1 - p0) - den (c - p0), compared against
// (radius*margin)^2 den^2.  |den| < 1e-10 never counts.
inline bool tiny(float x) { return !(fabsf(x) >= 1e-10f); }
inline bool tiny(double x) { return !(fabs(x) >= 1e-10); }

template <class S>
inline void gate_update(const dr_track& tr, S margin, const S p0[3], St<S>& s)
{
    const int nxt = (s.gate + 1) % tr.n_gates;
    const dr_gate& g = tr.gates[nxt];
    S a[3], c[3], n[3];
    for (int i = 0; i < 3; ++i)
    {
        a[i] = s.p[i] - p0[i];
        c[i] = (S)g.center[i] - p0[i];
        n[i] = (S)g.normal[i];
    }
    const S num = fm(n[0], c[0], fm(n[1], c[1], n[2] * c[2]));
    const S den = fm(n[0], a[0], fm(n[1], a[1], n[2] * a[2]));
    const S zero = (S)0;
    const bool cand = (den > zero) ? (!(num < zero) && num <= den) : (num <= zero && !(num < den));
    if (!cand || tiny(den))
        return;
    S w[3];
    for (int i = 0; i < 3; ++i)
        w[i] = fm(num, a[i], -(den * c[i]));
    const S rm = (S)g.radius * margin;
    if (fm(w[0], w[0], fm(w[1], w[1], w[2] * w[2])) <= (rm * rm) * (den * den))
    {
        s.gate = nxt;
        if (nxt == 0)
            s.laps = s.laps + 1;
    }
}

// env_dronerace.h:374-394
template <class S>
inline S progress(const dr_track& tr, const St<S>& s)
{
    const int nxt = (s.gate + 1) % tr.n_gates;
    const dr_gate& a = tr.gates[s.gate];
    const dr_gate& b = tr.gates[nxt];
    S g2 = 0, c2 = 0;
    S gd[3], cd[3];
    for (int i = 0; i < 3; ++i)
    {
        gd[i] = (S)b.center[i] - s.p[i];
        cd[i] = (S)b.center[i] - (S)a.center[i];
    }
    g2 = gd[0] * gd[0] + gd[1] * gd[1] + gd[2] * gd[2];
    c2 = cd[0] * cd[0] + cd[1] * cd[1] + cd[2] * cd[2];
    return (S)s.gate + (S)(tr.n_gates * s.laps) + (S)0.8f * ((S)1 - sq_root(g2 / c2));
}

// Cost constants formed in FP32 exactly as the engine does on the host.
template <class S>
struct CostK
{
    S r2_prox, inv_prox, danger, inv_danger, coll, far2;
    CostK(const dr_quad_params& P, const dr_cost_cfg& K)
    {
        const float r2 = K.opp_prox_radius * K.opp_prox_radius;
        const float ip = 1.0f / K.opp_prox_radius;
        const float dg = K.boundary_threshold_factor * P.drone_radius;
        const float idg = 1.0f / dg;
        const float cl = P.drone_radius * K.coll_dist_factor;
        const float reach = (dg > cl ? dg : cl) + P.drone_radius;
        r2_prox = (S)r2;
        inv_prox = (S)ip;
        danger = (S)dg;
        inv_danger = (S)idg;
        coll = (S)cl;
        far2 = (S)(reach * reach);
    }
};

// costs.cuh:7-28 structure with the quadrotor's boundary distance (see the
// engine's stage_cost for the terms).  Opponents beyond sqrt(far2) cannot
// reach the penalty band; their distance is +inf.
template <class S>
inline S running_cost(const Model<S>& M, const CostK<S>& Q, const dr_cost_cfg& K, const dr_track& tr, const St<S>& s,
                      const dr_opponent* opp, int n_opp, S tau, int* term)
{
    S cost = 0;
    S min_sq = (S)INFINITY;
    for (int j = 0; j < n_opp; ++j)
    {
        S d[3];
        for (int i = 0; i < 3; ++i)
            d[i] = s.p[i] - fm((S)opp[j].v[i], tau, (S)opp[j].p[i]);
        const S d2 = fm(d[0], d[0], fm(d[1], d[1], d[2] * d[2]));
        min_sq = mn(min_sq, d2);
        if (d2 < Q.r2_prox)
        {
            const S x = fm(-sq_root(d2), Q.inv_prox, (S)1);
            cost = fm((S)K.opp_prox_cost, x * x, cost);
        }
    }
    S d_opp = (S)INFINITY;
    if (min_sq < Q.far2)
        d_opp = sq_root(min_sq) - M.radius;
    const S d_ground = s.p[2] - (S)tr.ground_z;
    const S bd = mn(d_opp, d_ground);
    if (bd < Q.danger)
    {
        const S x = fm(-bd, Q.inv_danger, (S)1);
        cost = fm((S)K.boundary_cost, x * x, cost);
    }
    *term = 0;
    if (bd < Q.coll)
    {
        cost = cost + (S)K.outside_cost;
        *term = 1;
    }
    else if (s.laps >= tr.n_win_laps)
    {
        cost = cost - (S)K.win_cost;
        *term = 2;
    }
    return cost;
}

// PR-MPPI running cost and safety (costs.cuh:44-63, :78-86 structure; the
// engine's pr_stage_cost): the running cost without the collision term and
// safety = r*coll - bd with the exact boundary distance.  opp[j].v already
// holds the model's (shifted) velocity.
template <class S>
inline S pr_running_cost(const Model<S>& M, const CostK<S>& Q, const dr_cost_cfg& K, const dr_track& tr, const St<S>& s,
                         const dr_opponent* opp, int n_opp, S tau, S* safety, int* term)
{
    S cost = 0;
    S min_sq = (S)INFINITY;
    for (int j = 0; j < n_opp; ++j)
    {
        S d[3];
        for (int i = 0; i < 3; ++i)
            d[i] = s.p[i] - fm((S)opp[j].v[i], tau, (S)opp[j].p[i]);
        const S d2 = fm(d[0], d[0], fm(d[1], d[1], d[2] * d[2]));
        min_sq = mn(min_sq, d2);
        if (d2 < Q.r2_prox)
        {
            const S x = fm(-sq_root(d2), Q.inv_prox, (S)1);
            cost = fm((S)K.opp_prox_cost, x * x, cost);
        }
    }
    const S d_opp = sq_root(min_sq) - M.radius;
    const S d_ground = s.p[2] - (S)tr.ground_z;
    const S bd = mn(d_opp, d_ground);
    if (bd < Q.danger)
    {
        const S x = fm(-bd, Q.inv_danger, (S)1);
        cost = fm((S)K.boundary_cost, x * x, cost);
    }
    *safety = Q.coll - bd;
    *term = 0;
    if (bd < Q.coll)
        *term = 1;
    else if (s.laps >= tr.n_win_laps)
    {
        cost = cost - (S)K.win_cost;
        *term = 2;
    }
    return cost;
}

// Opponent model m of a sampled PR-MPPI step (prmppi_kernels.cu:23-51):
// u in (0, 1] from Philox counter (m, 0, step, 3); first i whose running
// FP32 belief sum reaches u, else the last model.
inline int sample_branch(uint32_t k0, uint32_t k1, uint32_t m, uint32_t step, const float* belief, int n_branches)
{
    const W4 r = philox(W4{{m, 0u, step, 3u}}, k0, k1);
    const float u = (float)((r.v[0] >> 8) + 1u) * 5.9604644775390625e-08f;
    float acc = 0.0f;
    for (int i = 0; i < n_branches; ++i)
    {
        acc = acc + belief[i];
        if (u <= acc)
            return i;
    }
    return n_branches - 1;
}

} // namespace orc

// ------------------------------------------------------------ op counting
// Scalar that counts FP operations (add, sub, mul, div, sqrt, min/max and
// comparisons each count 1) for the algorithmic work figure of the rollout.
namespace orc {

struct OpTally
{
    long long flops = 0;
};
extern thread_local OpTally g_tally;

struct Cn
{
    double v;
    Cn() : v(0) {}
    Cn(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    explicit operator int() const { return (int)v; }
};
inline Cn operator+(Cn a, Cn b) { ++g_tally.flops; return Cn(a.v + b.v); }
inline Cn operator-(Cn a, Cn b) { ++g_tally.flops; return Cn(a.v - b.v); }
inline Cn operator*(Cn a, Cn b) { ++g_tally.flops; return Cn(a.v * b.v); }
inline Cn operator/(Cn a, Cn b) { ++g_tally.flops; return Cn(a.v / b.v); }
inline Cn operator-(Cn a) { return Cn(-a.v); }
inline bool operator<(Cn a, Cn b) { ++g_tally.flops; return a.v < b.v; }
inline bool operator>(Cn a, Cn b) { ++g_tally.flops; return a.v > b.v; }
inline bool operator<=(Cn a, Cn b) { ++g_tally.flops; return a.v <= b.v; }
inline bool operator>=(Cn a, Cn b) { ++g_tally.flops; return a.v >= b.v; }
template <>
inline Cn sq_root<Cn>(Cn x) { ++g_tally.flops; return Cn(sqrt(x.v)); }
// the FP32 inverse square root's work: x/2, then per Newton step y^2 and two fmas (2 each)
template <>
inline Cn inv_sqrt<Cn>(Cn x) { g_tally.flops += 1 + 3 * 5; return Cn(1.0 / sqrt(x.v)); }
inline bool tiny(Cn x) { ++g_tally.flops; return !(fabs(x.v) >= 1e-10); }
inline Cn fm(Cn a, Cn b, Cn c) { g_tally.flops += 2; return Cn(fma(a.v, b.v, c.v)); }

} // namespace orc
