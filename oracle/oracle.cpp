// ORACLE — test infrastructure only (see oracle.h).  C entry points loaded
// with ctypes by tests/, the smoke check and bench.py's CPU-baseline leg.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "oracle.h"

using namespace orc;

thread_local OpTally orc::g_tally;

namespace {

template <class S>
void state_to_arr(const St<S>& s, double* a)
{
    for (int i = 0; i < 3; ++i)
        a[i] = (double)s.p[i];
    for (int i = 0; i < 4; ++i)
        a[3 + i] = (double)s.q[i];
    for (int i = 0; i < 3; ++i)
        a[7 + i] = (double)s.v[i];
    for (int i = 0; i < 3; ++i)
        a[10 + i] = (double)s.w[i];
}

template <class S>
St<S> arr_to_state(const double* a, int gate, int laps)
{
    St<S> s;
    for (int i = 0; i < 3; ++i)
        s.p[i] = (S)a[i];
    for (int i = 0; i < 4; ++i)
        s.q[i] = (S)a[3 + i];
    for (int i = 0; i < 3; ++i)
        s.v[i] = (S)a[7 + i];
    for (int i = 0; i < 3; ++i)
        s.w[i] = (S)a[10 + i];
    s.gate = gate;
    s.laps = laps;
    return s;
}

// One sample's rollout (src/kernels.cu:100-155): for t < T with early stop
// on a terminal state, control = clamp(nominal[t] + eps[t]), Euler step, gate
// update, running cost * decay (decay *= 0.97), then the terminal cost.  eps
// is NULL for the noise-free rollout of the nominal.
template <class S>
S rollout_one(const dr_quad_params& P, const Model<S>& M, const dr_cost_cfg& K, const dr_track& tr, const dr_drone_state& ego, const dr_opponent* opp, int n_opp,
              const float* nominal, int T, const float* eps, long long eps_stride)
{
    St<S> s = load_state<S>(ego);
    S acc = 0, decay = 1;
    const S margin = (S)K.gate_traversal_margin;
    const CostK<S> Q(P, K);
    for (int t = 0; t < T; ++t)
    {
        S uin[4], u[4];
        for (int d = 0; d < 4; ++d)
        {
            const float e = eps ? eps[(long long)t * eps_stride + d] : 0.0f;
            if (sizeof(S) == 4)
                uin[d] = (S)(nominal[t * 4 + d] + e);
            else
                uin[d] = (S)nominal[t * 4 + d] + (S)e;
        }
        clamp_u(M, uin, u);
        const S p0[3] = {s.p[0], s.p[1], s.p[2]};
        euler(M, s, u);
        gate_update(tr, margin, p0, s);
        int term = 0;
        const S tau = (S)(M.dt * (S)(t + 1));
        const S c = running_cost(M, Q, K, tr, s, opp, n_opp, tau, &term);
        acc = fm(c, decay, acc);
        decay = decay * (S)0.97f;
        if (term)
            break;
    }
    return acc + (-(S)K.final_adv_weight * progress(tr, s));
}

template <class S>
void rollout_all(const dr_quad_params& P, const dr_cost_cfg& K, const dr_track& tr, const dr_drone_state& ego, const dr_opponent* opp,
                 int n_opp, const float* nominal, int n_samples, int T, uint32_t key0, uint32_t key1, uint32_t step, const float sigma[4],
                 double* costs, float* noise_out, double* traj_cost, int n_threads)
{
    const Model<S> M(P);
    if (traj_cost)
        *traj_cost = (double)rollout_one<S>(P, M, K, tr, ego, opp, n_opp, nominal, T, nullptr, 0);
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : 1)
    {
        std::vector<float> eps((size_t)T * 4);
#pragma omp for schedule(static)
        for (int k = 0; k < n_samples; ++k)
        {
            for (int t = 0; t < T; ++t)
            {
                float z[4];
                normals4(key0, key1, (uint32_t)k, (uint32_t)t, step, 0u, z);
                for (int d = 0; d < 4; ++d)
                    eps[(size_t)t * 4 + d] = sigma[d] * z[d];
                if (noise_out)
                    for (int d = 0; d < 4; ++d)
                        noise_out[((size_t)k * T + t) * 4 + d] = eps[(size_t)t * 4 + d];
            }
            costs[k] = (double)rollout_one<S>(P, M, K, tr, ego, opp, n_opp, nominal, T, eps.data(), 4);
        }
    }
}

// Plain MPPI update (kernels.cu:307-408 with one branch and coefficient 1,
// clamp kernels.cu:434-456, action mppi_controller.cu:392-397, shift
// mppi_controller.cu:430-436).  FP32 mode: weights in FP32 with exp32, sums
// in FP64, increment rounded to FP32 and added in FP32.  FP64 mode: all FP64.
template <class S>
void update(const dr_quad_params& P, const double* costs, const float* noise, const float* nominal, int n_samples, int T, float lambda,
            const float fill[4], double* updated, double* action, double* shifted, double* min_cost, double* nu, double* weights)
{
    const Model<S> M(P);
    S mnc = (S)INFINITY;
    for (int k = 0; k < n_samples; ++k)
        mnc = mn(mnc, (S)costs[k]);
    std::vector<double> w(n_samples);
    double den = 0.0;
    for (int k = 0; k < n_samples; ++k)
    {
        if (sizeof(S) == 4)
            w[k] = (double)exp32(-((float)costs[k] - (float)mnc) / lambda);
        else
            w[k] = exp(-(costs[k] - (double)mnc) / (double)lambda);
        den += w[k];
        if (weights)
            weights[k] = w[k];
    }
    std::vector<S> u((size_t)T * 4);
    for (int t = 0; t < T; ++t)
        for (int d = 0; d < 4; ++d)
        {
            double num = 0.0;
            for (int k = 0; k < n_samples; ++k)
                num += w[k] * (double)noise[((size_t)k * T + t) * 4 + d];
            S v = (S)nominal[t * 4 + d];
            if (den > 1e-30)
            {
                if (sizeof(S) == 4)
                    v = (S)((float)v + (float)(num / den));
                else
                    v = v + (S)(num / den);
            }
            u[(size_t)t * 4 + d] = v;
        }
    for (int t = 0; t < T; ++t)
    {
        S o[4];
        clamp_u(M, &u[(size_t)t * 4], o);
        for (int d = 0; d < 4; ++d)
            u[(size_t)t * 4 + d] = o[d];
    }
    for (int i = 0; i < T * 4; ++i)
        updated[i] = (double)u[i];
    for (int d = 0; d < 4; ++d)
        action[d] = (double)u[d];
    for (int t = 0; t < T; ++t)
        for (int d = 0; d < 4; ++d)
            shifted[t * 4 + d] = (t + 1 < T) ? (double)u[(size_t)(t + 1) * 4 + d] : (double)fill[d];
    *min_cost = (double)mnc;
    *nu = sizeof(S) == 4 ? (double)(float)den : den;
}

// ---------------------------------------------------------------- adjoint
// Reverse-mode restatement of the rollout cost J(U) (noise-free) with
// checkpoints every `ck` steps and recomputation in between.  Same
// differentiation rules as the engine: clamp' = 1 on [lo, hi]; gate
// crossings and the terminal stop piecewise constant; min() takes the
// gradient of its active side.

template <class S>
struct G13
{
    S p[3], q[4], v[3], w[3];
};

template <class S>
void vjp_step(const Model<S>& M, const St<S>& x, const S u[4], G13<S>& gx, S gu[4])
{
    const S qw = x.q[0], qx = x.q[1], qy = x.q[2], qz = x.q[3];
    const S wx = x.w[0], wy = x.w[1], wz = x.w[2];
    const S two = (S)2, one = (S)1;
    const S b3[3] = {two * fm(qx, qz, qw * qy), two * fm(qy, qz, -(qw * qx)), fm(-two, fm(qx, qx, qy * qy), one)};
    const S at = u[0] * M.inv_mass;
    const S h = M.half_dt;
    const S hx = h * wx, hy = h * wy, hz = h * wz;
    const S rs = inv_sqrt(fm(hx, hx, fm(hy, hy, fm(hz, hz, one))));
    const S n0 = fm(-qx, hx, fm(-qy, hy, fm(-qz, hz, qw)));
    const S n1 = fm(qw, hx, fm(qy, hz, fm(-qz, hy, qx)));
    const S n2 = fm(qw, hy, fm(qz, hx, fm(-qx, hz, qy)));
    const S n3 = fm(qw, hz, fm(qx, hy, fm(-qy, hx, qz)));
    const S q2 = fm(qw, qw, fm(qx, qx, fm(qy, qy, qz * qz)));
    const S c = fm((S)-0.5f, q2, (S)1.5f);
    const S sc = rs * c;
    G13<S> g;
    const S kk = M.rate_k;
    for (int i = 0; i < 3; ++i)
    {
        gu[1 + i] = kk * gx.w[i];
        g.w[i] = (one - kk) * gx.w[i];
    }
    const S gn0 = sc * gx.q[0], gn1 = sc * gx.q[1], gn2 = sc * gx.q[2], gn3 = sc * gx.q[3];
    const S gsc = fm(n0, gx.q[0], fm(n1, gx.q[1], fm(n2, gx.q[2], n3 * gx.q[3])));
    const S gc = rs * gsc;
    const S gr2 = ((S)-0.5f * (rs * rs * rs)) * (c * gsc);
    g.q[0] = (gn0 + hx * gn1 + hy * gn2 + hz * gn3) - qw * gc;
    g.q[1] = (-hx * gn0 + gn1 - hz * gn2 + hy * gn3) - qx * gc;
    g.q[2] = (-hy * gn0 + hz * gn1 + gn2 - hx * gn3) - qy * gc;
    g.q[3] = (-hz * gn0 - hy * gn1 + hx * gn2 + gn3) - qz * gc;
    const S ghx = (-qx * gn0 + qw * gn1 + qz * gn2 - qy * gn3) + two * hx * gr2;
    const S ghy = (-qy * gn0 - qz * gn1 + qw * gn2 + qx * gn3) + two * hy * gr2;
    const S ghz = (-qz * gn0 + qy * gn1 - qx * gn2 + qw * gn3) + two * hz * gr2;
    g.w[0] = g.w[0] + h * ghx;
    g.w[1] = g.w[1] + h * ghy;
    g.w[2] = g.w[2] + h * ghz;
    S ga[3];
    for (int i = 0; i < 3; ++i)
    {
        ga[i] = M.dt * gx.v[i];
        g.p[i] = gx.p[i];
        g.v[i] = gx.v[i] + M.dt * gx.p[i] - M.drag * ga[i];
    }
    const S gat = ga[0] * b3[0] + ga[1] * b3[1] + ga[2] * b3[2];
    gu[0] = gat * M.inv_mass;
    const S gb0 = at * ga[0], gb1 = at * ga[1], gb2 = at * ga[2];
    g.q[0] = g.q[0] + two * (qy * gb0 - qx * gb1);
    g.q[1] = g.q[1] + two * (qz * gb0 - qw * gb1) - (S)4 * qx * gb2;
    g.q[2] = g.q[2] + two * (qw * gb0 + qz * gb1) - (S)4 * qy * gb2;
    g.q[3] = g.q[3] + two * (qx * gb0 + qy * gb1);
    gx = g;
}

template <class S>
void stage_grad(const Model<S>& M, const CostK<S>& Q, const dr_cost_cfg& K, const dr_track& tr, const St<S>& s, const dr_opponent* opp,
                int n_opp, S tau, S weight, S g[3])
{
    S min_sq = (S)INFINITY;
    int jmin = -1;
    S om[3] = {0, 0, 0};
    for (int j = 0; j < n_opp; ++j)
    {
        S o[3], d[3];
        for (int i = 0; i < 3; ++i)
        {
            o[i] = fm((S)opp[j].v[i], tau, (S)opp[j].p[i]);
            d[i] = s.p[i] - o[i];
        }
        const S d2 = fm(d[0], d[0], fm(d[1], d[1], d[2] * d[2]));
        if (d2 < min_sq)
        {
            min_sq = d2;
            jmin = j;
            for (int i = 0; i < 3; ++i)
                om[i] = o[i];
        }
        if (d2 < Q.r2_prox && d2 > (S)0)
        {
            const S dd = sq_root(d2);
            const S x = fm(-dd, Q.inv_prox, (S)1);
            const S f = weight * ((S)2 * (S)K.opp_prox_cost * x) * (-Q.inv_prox) / dd;
            for (int i = 0; i < 3; ++i)
                g[i] = fm(f, d[i], g[i]);
        }
    }
    S d_opp = (S)INFINITY;
    if (min_sq < Q.far2)
        d_opp = sq_root(min_sq) - M.radius;
    const S d_ground = s.p[2] - (S)tr.ground_z;
    const bool opp_side = d_opp < d_ground;
    const S bd = opp_side ? d_opp : d_ground;
    if (bd < Q.danger)
    {
        const S x = fm(-bd, Q.inv_danger, (S)1);
        const S dbd = weight * ((S)2 * (S)K.boundary_cost * x) * (-Q.inv_danger);
        if (opp_side && jmin >= 0)
        {
            const S dd = sq_root(min_sq);
            if (dd > (S)0)
            {
                const S f = dbd / dd;
                for (int i = 0; i < 3; ++i)
                    g[i] = fm(f, s.p[i] - om[i], g[i]);
            }
        }
        else
            g[2] = g[2] + dbd;
    }
}

template <class S>
void final_grad(const dr_cost_cfg& K, const dr_track& tr, const St<S>& s, S g[3])
{
    const int nxt = (s.gate + 1) % tr.n_gates;
    const dr_gate& a = tr.gates[s.gate];
    const dr_gate& b = tr.gates[nxt];
    S gd[3], cd[3];
    for (int i = 0; i < 3; ++i)
    {
        gd[i] = (S)b.center[i] - s.p[i];
        cd[i] = (S)b.center[i] - (S)a.center[i];
    }
    const S g2 = gd[0] * gd[0] + gd[1] * gd[1] + gd[2] * gd[2];
    const S c2 = cd[0] * cd[0] + cd[1] * cd[1] + cd[2] * cd[2];
    const S sr = sq_root(g2 / c2);
    if (sr > (S)0)
    {
        const S f = -(S)K.final_adv_weight * (S)0.8f / (sr * c2);
        for (int i = 0; i < 3; ++i)
            g[i] = fm(f, gd[i], g[i]);
    }
}

template <class S>
St<S> step_state(const Model<S>& M, const dr_track& tr, S margin, const St<S>& x, const float* U, S u_out[4])
{
    St<S> s = x;
    S uin[4];
    for (int d = 0; d < 4; ++d)
        uin[d] = (S)U[d];
    clamp_u(M, uin, u_out);
    const S p0[3] = {s.p[0], s.p[1], s.p[2]};
    euler(M, s, u_out);
    gate_update(tr, margin, p0, s);
    return s;
}

// J(U) and dJ/dU for one control sequence (T x 4); sig (may be NULL)
// receives a signature of the discrete decisions (clamp activity, gate
// index sequence, stop step, active cost branches) for kink detection.
template <class S>
S cost_grad(const dr_quad_params& P, const dr_cost_cfg& K, const dr_track& tr, const dr_drone_state& ego, const dr_opponent* opp, int n_opp,
            const float* U, int T, int ck, S* dJdU, uint64_t* sig)
{
    const Model<S> M(P);
    const CostK<S> Q(P, K);
    const S margin = (S)K.gate_traversal_margin;
    const int n_ck = (T + ck - 1) / ck;
    std::vector<St<S>> cks(n_ck);
    std::vector<S> dec(T + 1);
    St<S> s = load_state<S>(ego);
    S acc = 0, decay = 1;
    int t_end = T - 1;
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
    for (int t = 0; t < T; ++t)
    {
        if (t % ck == 0)
            cks[t / ck] = s;
        dec[t] = decay;
        S u[4];
        for (int d = 0; d < 4; ++d)
        {
            const S x = (S)U[t * 4 + d];
            const S lo = d == 0 ? M.tmin : -M.rmax, hi = d == 0 ? M.tmax : M.rmax;
            mix((uint64_t)((x < lo) ? 1 : (x > hi ? 2 : 0)));
        }
        s = step_state(M, tr, margin, s, U + t * 4, u);
        mix((uint64_t)(s.gate * 131 + s.laps));
        int term = 0;
        const S tau = (S)(M.dt * (S)(t + 1));
        const S c = running_cost(M, Q, K, tr, s, opp, n_opp, tau, &term);
        // active-branch signature of the running cost
        {
            S min_sq = (S)INFINITY;
            uint64_t bits = 0;
            for (int j = 0; j < n_opp; ++j)
            {
                S d[3];
                for (int i = 0; i < 3; ++i)
                    d[i] = s.p[i] - fm((S)opp[j].v[i], tau, (S)opp[j].p[i]);
                const S d2 = fm(d[0], d[0], fm(d[1], d[1], d[2] * d[2]));
                if (d2 < min_sq)
                {
                    min_sq = d2;
                    bits = (bits & ~0xffull) | (uint64_t)j;
                }
                if (d2 < Q.r2_prox)
                    bits |= 1ull << (8 + j);
            }
            S d_opp = (S)INFINITY;
            if (min_sq < Q.far2)
                d_opp = sq_root(min_sq) - M.radius;
            const S d_ground = s.p[2] - (S)tr.ground_z;
            const S bd = d_opp < d_ground ? d_opp : d_ground;
            bits |= (uint64_t)(min_sq < Q.far2) << 20;
            bits |= (uint64_t)(d_opp < d_ground) << 21;
            bits |= (uint64_t)(bd < Q.danger) << 22;
            mix(bits);
        }
        acc = fm(c, decay, acc);
        decay = decay * (S)0.97f;
        if (term)
        {
            t_end = t;
            mix(1000 + (uint64_t)t);
            break;
        }
    }
    const S J = acc + (-(S)K.final_adv_weight * progress(tr, s));
    if (sig)
        *sig = h;
    if (!dJdU)
        return J;
    for (int i = 0; i < T * 4; ++i)
        dJdU[i] = 0;
    G13<S> g;
    for (int i = 0; i < 3; ++i)
        g.p[i] = g.v[i] = g.w[i] = 0;
    for (int i = 0; i < 4; ++i)
        g.q[i] = 0;
    final_grad(K, tr, s, g.p);
    stage_grad(M, Q, K, tr, s, opp, n_opp, (S)(M.dt * (S)(t_end + 1)), dec[t_end], g.p);
    std::vector<St<S>> seg(ck + 1);
    for (int sg = t_end / ck; sg >= 0; --sg)
    {
        const int t0 = sg * ck;
        const int t1 = std::min(t0 + ck, t_end + 1);
        seg[0] = cks[sg];
        for (int t = t0; t < t1; ++t)
        {
            S u[4];
            seg[t - t0 + 1] = step_state(M, tr, margin, seg[t - t0], U + t * 4, u);
        }
        for (int t = t1 - 1; t >= t0; --t)
        {
            const St<S>& x = seg[t - t0];
            S uin[4], u[4], gu[4];
            for (int d = 0; d < 4; ++d)
                uin[d] = (S)U[t * 4 + d];
            clamp_u(M, uin, u);
            vjp_step(M, x, u, g, gu);
            for (int d = 0; d < 4; ++d)
            {
                const S lo = d == 0 ? M.tmin : -M.rmax, hi = d == 0 ? M.tmax : M.rmax;
                dJdU[t * 4 + d] = (uin[d] >= lo && uin[d] <= hi) ? gu[d] : (S)0;
            }
            if (t >= 1)
                stage_grad(M, Q, K, tr, x, opp, n_opp, (S)(M.dt * (S)t), dec[t - 1], g.p);
        }
    }
    return J;
}

} // namespace

extern "C" {

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    W4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
    const W4 o = philox(c, key[0], key[1]);
    memcpy(out, o.v, sizeof(o.v));
}

void or_normals4(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, float out[4])
{
    normals4(k0, k1, c0, c1, c2, c3, out);
}

float or_log32(float x) { return log32(x); }
float or_exp32(float x) { return exp32(x); }
void or_sincos_turns(uint32_t k, float* s, float* c) { sincos_turns(k, s, c); }

float or_inv_sqrt32(float x) { return inv_sqrt<float>(x); }

// One clamped Euler step + gate update on a 13-vector (p, q, v, w).
void or_quad_step(int f64, const dr_quad_params* P, const dr_track* tr, float margin, double* x, int* gate, int* laps, const double* u_in)
{
    if (f64)
    {
        const Model<double> M(*P);
        St<double> s = arr_to_state<double>(x, *gate, *laps);
        double u[4];
        clamp_u(M, u_in, u);
        const double p0[3] = {s.p[0], s.p[1], s.p[2]};
        euler(M, s, u);
        if (tr)
            gate_update(*tr, (double)margin, p0, s);
        state_to_arr(s, x);
        *gate = s.gate;
        *laps = s.laps;
    }
    else
    {
        const Model<float> M(*P);
        St<float> s = arr_to_state<float>(x, *gate, *laps);
        float ui[4] = {(float)u_in[0], (float)u_in[1], (float)u_in[2], (float)u_in[3]}, u[4];
        clamp_u(M, ui, u);
        const float p0[3] = {s.p[0], s.p[1], s.p[2]};
        euler(M, s, u);
        if (tr)
            gate_update(*tr, margin, p0, s);
        state_to_arr(s, x);
        *gate = s.gate;
        *laps = s.laps;
    }
}

double or_progress(int f64, const dr_track* tr, const double* x, int gate, int laps)
{
    if (f64)
        return progress(*tr, arr_to_state<double>(x, gate, laps));
    return (double)progress(*tr, arr_to_state<float>(x, gate, laps));
}

double or_running_cost(int f64, const dr_quad_params* P, const dr_cost_cfg* K, const dr_track* tr, const double* x, int gate, int laps,
                       const dr_opponent* opp, int n_opp, double tau, int* term)
{
    if (f64)
        return running_cost(Model<double>(*P), CostK<double>(*P, *K), *K, *tr, arr_to_state<double>(x, gate, laps), opp, n_opp, tau,
                            term);
    return (double)running_cost(Model<float>(*P), CostK<float>(*P, *K), *K, *tr, arr_to_state<float>(x, gate, laps), opp, n_opp,
                                (float)tau, term);
}

// K sampled rollouts of one controller.  costs[K] (FP32 values widened in
// FP32 mode); noise_out [k][t][4] (may be NULL); traj_cost = noise-free
// rollout of the nominal (may be NULL).
int or_rollout(int f64, const dr_quad_params* P, const dr_cost_cfg* K, const dr_track* tr, const dr_drone_state* ego, const dr_opponent* opp,
               int n_opp, const float* nominal, int n_samples, int T, uint32_t key0, uint32_t key1, uint32_t step, const float* sigma,
               double* costs, float* noise_out, double* traj_cost, int n_threads)
{
    if (!P || !K || !tr || !ego || !nominal || !costs || !sigma || n_samples < 1 || T < 1 || T > DR_MAX_T || n_opp < 0 ||
        (n_opp > 0 && !opp))
        return 1;
    if (f64)
        rollout_all<double>(*P, *K, *tr, *ego, opp, n_opp, nominal, n_samples, T, key0, key1, step, sigma, costs, noise_out, traj_cost,
                            n_threads);
    else
        rollout_all<float>(*P, *K, *tr, *ego, opp, n_opp, nominal, n_samples, T, key0, key1, step, sigma, costs, noise_out, traj_cost,
                           n_threads);
    return 0;
}

int or_mppi_update(int f64, const dr_quad_params* P, const double* costs, const float* noise, const float* nominal, int n_samples, int T,
                   float lambda, const float* fill, double* updated, double* action, double* shifted, double* min_cost, double* nu,
                   double* weights)
{
    if (!P || !costs || !noise || !nominal || !fill || !updated || !action || !shifted || !min_cost || !nu || n_samples < 1 || T < 1)
        return 1;
    if (f64)
        update<double>(*P, costs, noise, nominal, n_samples, T, lambda, fill, updated, action, shifted, min_cost, nu, weights);
    else
        update<float>(*P, costs, noise, nominal, n_samples, T, lambda, fill, updated, action, shifted, min_cost, nu, weights);
    return 0;
}

// Closed-loop run of a batch on the CPU in FP32 (the CPU baseline and the
// multi-step parity check): per control step, every instance rolls out K
// samples against constant-velocity opponents snapshotted at the step start,
// updates, clamps, takes its action and shifts; then every controlled drone
// is integrated with its action.  nominal holds sum_i T_i*4 floats in
// instance order; steps and lambdas are per instance and updated in place.
//
// n_grad_steps > 0 adds the adjoint refinement after each update (the
// engine's dr_engine_set_refine): projected gradient steps
// U <- clamp(U - grad_step dJ/dU) from the clamped update, checkpoints every
// ck steps, J of every iterate, the lowest kept (the update itself on ties),
// then action and shift.  The FP32 adjoint here sums in the one-lane order,
// so the refined plans agree with the engine's to rounding, not bit for bit.
int or_world_run(const dr_quad_params* P, const dr_cost_cfg* K, const dr_track* tracks, const dr_instance* inst, int n_inst,
                 const int32_t* drone_world, dr_drone_state* drones, int n_drones, float* nominal, uint32_t* steps, float* lambdas,
                 int n_steps, int n_threads, float* actions_out, float* traj_costs_out, int n_grad_steps, float grad_step, int ck)
{
    if (!P || !K || !tracks || !inst || !drone_world || !drones || !nominal || !steps || !lambdas || n_inst < 1 || n_steps < 0 ||
        n_grad_steps < 0 || (n_grad_steps > 0 && ck < 1))
        return 1;
    const Model<float> M(*P);
    std::vector<int> nom_off(n_inst);
    int off = 0;
    for (int i = 0; i < n_inst; ++i)
    {
        nom_off[i] = off;
        off += inst[i].n_timesteps * 4;
    }
    std::vector<float> act((size_t)n_inst * 4);
    for (int step = 0; step < n_steps; ++step)
    {
        std::vector<dr_drone_state> snap(drones, drones + n_drones);
        for (int i = 0; i < n_inst; ++i)
        {
            const dr_instance& I = inst[i];
            std::vector<dr_opponent> opp;
            for (int d = 0; d < n_drones; ++d)
                if (d != I.drone && drone_world[d] == drone_world[I.drone])
                {
                    dr_opponent o;
                    for (int c = 0; c < 3; ++c)
                    {
                        o.p[c] = snap[d].p[c];
                        o.v[c] = snap[d].v[c];
                    }
                    opp.push_back(o);
                }
            const int Kn = I.n_samples, T = I.n_timesteps;
            std::vector<double> costs(Kn);
            std::vector<float> noise((size_t)T * Kn * 4);
            double traj = 0.0;
            rollout_all<float>(*P, *K, tracks[I.track], snap[I.drone], opp.data(), (int)opp.size(), nominal + nom_off[i], Kn, T,
                               (uint32_t)(I.seed & 0xffffffffu), (uint32_t)(I.seed >> 32), steps[i], I.sigma, costs.data(), noise.data(),
                               &traj, n_threads);
            std::vector<double> upd((size_t)T * 4), sh((size_t)T * 4);
            double a[4], mc, nu;
            update<float>(*P, costs.data(), noise.data(), nominal + nom_off[i], Kn, T, lambdas[i], I.shift_fill, upd.data(), a, sh.data(),
                          &mc, &nu, nullptr);
            if (n_grad_steps > 0)
            {
                // refine the clamped update, keep the lowest-J iterate, then act and shift
                std::vector<float> U((size_t)T * 4), best;
                for (int j = 0; j < T * 4; ++j)
                    U[j] = (float)upd[j];
                std::vector<float> G((size_t)T * 4);
                float j_best = 0.0f;
                for (int it = 0; it <= n_grad_steps; ++it)
                {
                    const bool last = it == n_grad_steps;
                    const float J = cost_grad<float>(*P, *K, tracks[I.track], snap[I.drone], opp.data(), (int)opp.size(), U.data(), T, ck,
                                                     last ? nullptr : G.data(), nullptr);
                    if (it == 0 || J < j_best)
                    {
                        j_best = J;
                        best = U;
                    }
                    if (last)
                        break;
                    for (int t = 0; t < T; ++t)
                    {
                        float in[4], out[4];
                        for (int d = 0; d < 4; ++d)
                            in[d] = fmaf(-grad_step, G[(size_t)t * 4 + d], U[(size_t)t * 4 + d]);
                        clamp_u(M, in, out);
                        for (int d = 0; d < 4; ++d)
                            U[(size_t)t * 4 + d] = out[d];
                    }
                }
                for (int d = 0; d < 4; ++d)
                    a[d] = best[d];
                for (int j = 0; j < T * 4; ++j)
                    sh[j] = (j + 4 < T * 4) ? best[j + 4] : I.shift_fill[j & 3];
            }
            for (int j = 0; j < T * 4; ++j)
                nominal[nom_off[i] + j] = (float)sh[j];
            for (int d = 0; d < 4; ++d)
                act[(size_t)i * 4 + d] = (float)a[d];
            if (I.adapt_lambda)
            {
                const float nuf = (float)nu;
                if (nuf > 0.001f * (float)Kn)
                    lambdas[i] = lambdas[i] * 0.9f;
                else if (nuf < 0.0005f * (float)Kn)
                    lambdas[i] = lambdas[i] * 1.2f;
            }
            if (traj_costs_out)
                traj_costs_out[i] = (float)traj;
            steps[i] = steps[i] + 1u;
        }
        for (int d = 0; d < n_drones; ++d)
        {
            int ctrl = -1;
            for (int i = 0; i < n_inst; ++i)
                if (inst[i].drone == d)
                    ctrl = i;
            if (ctrl < 0)
            {
                for (int c = 0; c < 3; ++c)
                    drones[d].p[c] = fmaf(M.dt, drones[d].v[c], drones[d].p[c]);
                continue;
            }
            St<float> s = load_state<float>(drones[d]);
            float u[4];
            clamp_u(M, &act[(size_t)ctrl * 4], u);
            const float p0[3] = {s.p[0], s.p[1], s.p[2]};
            euler(M, s, u);
            gate_update(tracks[inst[ctrl].track], K->gate_traversal_margin, p0, s);
            for (int c = 0; c < 3; ++c)
            {
                drones[d].p[c] = s.p[c];
                drones[d].v[c] = s.v[c];
                drones[d].w[c] = s.w[c];
            }
            for (int c = 0; c < 4; ++c)
                drones[d].q[c] = s.q[c];
            drones[d].gate = s.gate;
            drones[d].laps = s.laps;
        }
    }
    if (actions_out)
        memcpy(actions_out, act.data(), act.size() * sizeof(float));
    return 0;
}

// Algorithmic FP work of the rollouts of n_samples samples (samples
// 0..n_samples-1 of the given key/step): flops spent in dynamics, gate logic
// and costs (op-counting scalar, data-dependent branches included) and the
// number of active sample-steps.  The per-step noise (Box-Muller + scaling,
// a fixed count read off the code) is not included here.
int or_count_rollout_ops(const dr_quad_params* P, const dr_cost_cfg* K, const dr_track* tr, const dr_drone_state* ego,
                         const dr_opponent* opp, int n_opp, const float* nominal, int n_samples, int T, uint32_t key0, uint32_t key1,
                         uint32_t step, const float* sigma, long long* flops_out)
{
    if (!P || !K || !tr || !ego || !nominal || !sigma || !flops_out || n_samples < 1 || T < 1)
        return 1;
    const Model<Cn> M(*P);
    std::vector<float> eps((size_t)T * 4);
    g_tally.flops = 0;
    for (int k = 0; k < n_samples; ++k)
    {
        for (int t = 0; t < T; ++t)
        {
            float z[4];
            normals4(key0, key1, (uint32_t)k, (uint32_t)t, step, 0u, z);
            for (int d = 0; d < 4; ++d)
                eps[(size_t)t * 4 + d] = sigma[d] * z[d];
        }
        (void)rollout_one<Cn>(*P, M, *K, *tr, *ego, opp, n_opp, nominal, T, eps.data(), 4);
    }
    *flops_out = g_tally.flops;
    return 0;
}

// Algorithmic FP work of the adjoint of one control sequence U (T x 4):
// flops of the forward pass alone (J) and of forward + reverse (J and dJ/dU),
// op-counting scalar, checkpoints every ck steps (the recomputed segments
// counted too).
int or_count_cost_grad_ops(const dr_quad_params* P, const dr_cost_cfg* K, const dr_track* tr, const dr_drone_state* ego,
                           const dr_opponent* opp, int n_opp, const float* U, int T, int ck, long long* fwd_flops, long long* full_flops)
{
    if (!P || !K || !tr || !ego || !U || !fwd_flops || !full_flops || T < 1 || ck < 1 || (n_opp > 0 && !opp))
        return 1;
    std::vector<Cn> g((size_t)T * 4);
    g_tally.flops = 0;
    (void)cost_grad<Cn>(*P, *K, *tr, *ego, opp, n_opp, U, T, ck, nullptr, nullptr);
    *fwd_flops = g_tally.flops;
    g_tally.flops = 0;
    (void)cost_grad<Cn>(*P, *K, *tr, *ego, opp, n_opp, U, T, ck, g.data(), nullptr);
    *full_flops = g_tally.flops;
    return 0;
}

// J and dJ/dU for n_seq control sequences U[n_seq][T][4] (noise-free,
// checkpoints every ck steps); sig (may be NULL) gets the discrete-decision
// signature of each sequence.
int or_cost_grad(int f64, const dr_quad_params* P, const dr_cost_cfg* K, const dr_track* tr, const dr_drone_state* ego,
                 const dr_opponent* opp, int n_opp, const float* U, int n_seq, int T, int ck, double* J, double* dJdU, uint64_t* sig,
                 int n_threads)
{
    if (!P || !K || !tr || !ego || !U || !J || n_seq < 1 || T < 1 || ck < 1 || (n_opp > 0 && !opp))
        return 1;
#pragma omp parallel for schedule(dynamic) num_threads(n_threads > 0 ? n_threads : 1)
    for (int i = 0; i < n_seq; ++i)
    {
        if (f64)
        {
            std::vector<double> g(dJdU ? (size_t)T * 4 : 0);
            J[i] = cost_grad<double>(*P, *K, *tr, *ego, opp, n_opp, U + (size_t)i * T * 4, T, ck, dJdU ? g.data() : nullptr,
                                     sig ? sig + i : nullptr);
            if (dJdU)
                for (int j = 0; j < T * 4; ++j)
                    dJdU[(size_t)i * T * 4 + j] = g[j];
        }
        else
        {
            std::vector<float> g(dJdU ? (size_t)T * 4 : 0);
            J[i] = (double)cost_grad<float>(*P, *K, *tr, *ego, opp, n_opp, U + (size_t)i * T * 4, T, ck, dJdU ? g.data() : nullptr,
                                            sig ? sig + i : nullptr);
            if (dJdU)
                for (int j = 0; j < T * 4; ++j)
                    dJdU[(size_t)i * T * 4 + j] = (double)g[j];
        }
    }
    return 0;
}

} // extern "C"
