// Reverse-mode (adjoint) rollout of the Quad13 cost, J(U) = sum_t decay_t
// c(x_{t+1}) + c_f(x_end), for one control sequence U (T x 4) from a given
// state, noise-free.  The forward pass is dr_quad.h's; this file holds the
// vector-Jacobian products of its pieces:
//   clamp        1 inside [lo, hi] (bounds included), 0 outside
//   Euler step   quad_step's expressions (an fma differentiates as a*b + c;
//                the inverse square root as the exact 1/sqrt)
//   running cost opponent proximity, near-boundary term (gradient of the
//                active side of bd = min(d_opp, d_ground)); collision and win
//                terms are constants
//   terminal     -w * progress; d progress/dp = 0.8 (c_next - p)/(s |c_next - c_gate|^2),
//                s = |c_next - p| / |c_next - c_gate|
// Gate crossings and the terminal stop are discrete: piecewise constant.
// The reference has no differentiable path; this follows the forward code
// restated from include/costs.cuh:7-39 and env_dronerace.h:374-394.
#pragma once

#include "dr_quad.h"

namespace dr {

// dJ/dp of the running cost at state s (post-step, opponents at this step).
template <class OppAt>
DR_HD void stage_cost_grad_at(const QuadConst& C, const CostConst& Q, const dr_cost_cfg& K, const dr_track& tr, const QState& s,
                              int n_opp, OppAt opp_at, float weight, float g[3])
{
    float min_sq = __builtin_inff();
    int jmin = -1;
    for (int j = 0; j < n_opp; ++j)
    {
        const float4 o = opp_at(j);
        const float dx = s.p[0] - o.x, dy = s.p[1] - o.y, dz = s.p[2] - o.z;
        const float d2 = fma_(dx, dx, fma_(dy, dy, dz * dz));
        if (d2 < min_sq)
        {
            min_sq = d2;
            jmin = j;
        }
        const bool near = d2 < Q.r2_prox && d2 > 0.0f;
        if (any_lane(near))
        {
            keep_branch();
            if (near)
            {
                const float d = __builtin_sqrtf(d2);
                const float x = fma_(-d, Q.inv_prox, 1.0f);
                // d/dp W x^2 = 2 W x * (-inv_prox) * (p - o) / d
                const float f = weight * (2.0f * K.opp_prox_cost * x) * (-Q.inv_prox) / d;
                g[0] = fma_(f, dx, g[0]);
                g[1] = fma_(f, dy, g[1]);
                g[2] = fma_(f, dz, g[2]);
            }
        }
    }
    float d_opp = __builtin_inff();
    if (min_sq < Q.far2)
        d_opp = __builtin_sqrtf(min_sq) - C.drone_radius;
    const float d_ground = s.p[2] - tr.ground_z;
    const bool opp_side = d_opp < d_ground;
    const float bd = opp_side ? d_opp : d_ground;
    if (bd < Q.danger)
    {
        const float x = fma_(-bd, Q.inv_danger, 1.0f);
        const float dbd = weight * (2.0f * K.boundary_cost * x) * (-Q.inv_danger);
        if (opp_side && jmin >= 0)
        {
            const float4 o = opp_at(jmin);
            const float d = __builtin_sqrtf(min_sq);
            if (d > 0.0f)
            {
                const float f = dbd / d;
                g[0] = fma_(f, s.p[0] - o.x, g[0]);
                g[1] = fma_(f, s.p[1] - o.y, g[1]);
                g[2] = fma_(f, s.p[2] - o.z, g[2]);
            }
        }
        else
            g[2] = g[2] + dbd;
    }
}

DR_HD void stage_cost_grad(const QuadConst& C, const CostConst& Q, const dr_cost_cfg& K, const dr_track& tr, const QState& s,
                           const float4* opp, int n_opp, float weight, float g[3])
{
    stage_cost_grad_at(C, Q, K, tr, s, n_opp, [opp](int j) { return opp[j]; }, weight, g);
}

// dJ/dp of the terminal cost -final_adv_weight * progress.
DR_HD void final_cost_grad(const dr_cost_cfg& K, const dr_track& tr, const QState& s, float g[3])
{
    int next = s.gate + 1;
    if (next == tr.n_gates)
        next = 0;
    const dr_gate& a = tr.gates[s.gate];
    const dr_gate& b = tr.gates[next];
    const float gx = b.center[0] - s.p[0], gy = b.center[1] - s.p[1], gz = b.center[2] - s.p[2];
    const float cx = b.center[0] - a.center[0], cy = b.center[1] - a.center[1], cz = b.center[2] - a.center[2];
    const float sq_gate = gx * gx + gy * gy + gz * gz;
    const float sq_cons = cx * cx + cy * cy + cz * cz;
    const float sr = __builtin_sqrtf(sq_gate / sq_cons);
    if (sr > 0.0f)
    {
        const float f = -K.final_adv_weight * 0.8f / (sr * sq_cons);
        g[0] = fma_(f, gx, g[0]);
        g[1] = fma_(f, gy, g[1]);
        g[2] = fma_(f, gz, g[2]);
    }
}

// Adjoint state: gradient with respect to (p, q, v, w).
struct QGrad
{
    float p[3], q[4], v[3], w[3];
};

// Given x = state before the step, u = clamped control and gx = dJ/dx_{t+1},
// overwrite gx with dJ/dx_t (through this step only) and write dJ/du.
DR_HD void quad_step_vjp(const QuadConst& C, const QState& x, const float u[4], QGrad& gx, float gu[4])
{
    const float qw = x.q[0], qx = x.q[1], qy = x.q[2], qz = x.q[3];
    const float wx = x.w[0], wy = x.w[1], wz = x.w[2];
    const float b3x = 2.0f * fma_(qx, qz, qw * qy);
    const float b3y = 2.0f * fma_(qy, qz, -(qw * qx));
    const float b3z = fma_(-2.0f, fma_(qx, qx, qy * qy), 1.0f);
    const float at = u[0] * C.inv_mass;

    // recompute the attitude step's pieces: increment (h w, rs), product n, correction c
    const float h = C.half_dt;
    const float hx = h * wx, hy = h * wy, hz = h * wz;
    const float rs = rsqrt_nr(fma_(hx, hx, fma_(hy, hy, fma_(hz, hz, 1.0f))));
    const float n0 = fma_(-qx, hx, fma_(-qy, hy, fma_(-qz, hz, qw)));
    const float n1 = fma_(qw, hx, fma_(qy, hz, fma_(-qz, hy, qx)));
    const float n2 = fma_(qw, hy, fma_(qz, hx, fma_(-qx, hz, qy)));
    const float n3 = fma_(qw, hz, fma_(qx, hy, fma_(-qy, hx, qz)));
    const float q2 = fma_(qw, qw, fma_(qx, qx, fma_(qy, qy, qz * qz)));
    const float c = fma_(-0.5f, q2, 1.5f);
    const float sc = rs * c;

    // w' = w + k (u - w)
    QGrad g;
    const float kk = C.rate_k;
    gu[1] = kk * gx.w[0];
    gu[2] = kk * gx.w[1];
    gu[3] = kk * gx.w[2];
    g.w[0] = (1.0f - kk) * gx.w[0];
    g.w[1] = (1.0f - kk) * gx.w[1];
    g.w[2] = (1.0f - kk) * gx.w[2];

    // q' = n sc, sc = rs c: through n, then through sc to rs (r2 = 1 + |h w|^2) and c (c = 3/2 - |q|^2/2)
    const float gn0 = sc * gx.q[0], gn1 = sc * gx.q[1], gn2 = sc * gx.q[2], gn3 = sc * gx.q[3];
    const float gsc = fma_(n0, gx.q[0], fma_(n1, gx.q[1], fma_(n2, gx.q[2], n3 * gx.q[3])));
    const float gc = rs * gsc;
    const float gr2 = (-0.5f * (rs * rs * rs)) * (c * gsc);
    // n = q (x) (1, h w)
    g.q[0] = (gn0 + hx * gn1 + hy * gn2 + hz * gn3) - qw * gc;
    g.q[1] = (-hx * gn0 + gn1 - hz * gn2 + hy * gn3) - qx * gc;
    g.q[2] = (-hy * gn0 + hz * gn1 + gn2 - hx * gn3) - qy * gc;
    g.q[3] = (-hz * gn0 - hy * gn1 + hx * gn2 + gn3) - qz * gc;
    const float ghx = (-qx * gn0 + qw * gn1 + qz * gn2 - qy * gn3) + 2.0f * hx * gr2;
    const float ghy = (-qy * gn0 - qz * gn1 + qw * gn2 + qx * gn3) + 2.0f * hy * gr2;
    const float ghz = (-qz * gn0 + qy * gn1 - qx * gn2 + qw * gn3) + 2.0f * hz * gr2;
    g.w[0] = g.w[0] + h * ghx;
    g.w[1] = g.w[1] + h * ghy;
    g.w[2] = g.w[2] + h * ghz;

    // v' = v + dt a, a = at b3 - drag v - g e3; p' = p + dt v
    const float ga0 = C.dt * gx.v[0], ga1 = C.dt * gx.v[1], ga2 = C.dt * gx.v[2];
    g.p[0] = gx.p[0];
    g.p[1] = gx.p[1];
    g.p[2] = gx.p[2];
    g.v[0] = gx.v[0] + C.dt * gx.p[0] - C.drag * ga0;
    g.v[1] = gx.v[1] + C.dt * gx.p[1] - C.drag * ga1;
    g.v[2] = gx.v[2] + C.dt * gx.p[2] - C.drag * ga2;
    const float gat = ga0 * b3x + ga1 * b3y + ga2 * b3z;
    gu[0] = gat * C.inv_mass;
    const float gb0 = at * ga0, gb1 = at * ga1, gb2 = at * ga2;
    // b3 = (2(qx qz + qw qy), 2(qy qz - qw qx), 1 - 2(qx^2 + qy^2))
    g.q[0] = g.q[0] + 2.0f * (qy * gb0 - qx * gb1);
    g.q[1] = g.q[1] + 2.0f * (qz * gb0 - qw * gb1) - 4.0f * qx * gb2;
    g.q[2] = g.q[2] + 2.0f * (qw * gb0 + qz * gb1) - 4.0f * qy * gb2;
    g.q[3] = g.q[3] + 2.0f * (qx * gb0 + qy * gb1);
    gx = g;
}

// The same linearisation as quad_step_vjp, as the blocks of the step's
// Jacobian that depend on the state, so that many adjoint vectors can be
// carried through one step with no recomputation of the forward pieces.
// With lam = (P, V, Q, W) = dJ/d(p, v, q, w) after the step, the adjoint
// before it and dJ/du are
//   P' = P,  V' = (1 - drag dt) V + dt P,
//   Q'_i = sum_j Mq[i][j] Q_j + sum_a Mv[i][a] V_a,
//   W'_k = (1 - k) W_k + sum_j Mw[k][j] Q_j,
//   dJ/du = (bu . V, k W),
// where, with n = q (x) (1, h w), c = (3 - |q|^2)/2, rs = 1/|(1, h w)|,
// sc = rs c (the pieces quad_step_vjp recomputes):
//   Mq[i][j] = sc B[i][j] - rs q_i n_j      (B: the product's matrix in h w)
//   Mw[k][j] = h (sc Cq[k][j] - h_k rs^3 c n_j)   (Cq: the product's matrix in q)
//   Mv[i][a] = a_T dt dR(q)e3_a/dq_i,  bu = dt R(q)e3 / m.
// Layout of J: Mq row-major [0, 16), Mw [16, 28), Mv [28, 40), bu [40, 43).
constexpr int QJAC = 43;

DR_HD void quad_step_jacobian(const QuadConst& C, const QState& x, const float u[4], float J[QJAC])
{
    const float qw = x.q[0], qx = x.q[1], qy = x.q[2], qz = x.q[3];
    const float h = C.half_dt;
    const float hx = h * x.w[0], hy = h * x.w[1], hz = h * x.w[2];
    const float rs = rsqrt_nr(fma_(hx, hx, fma_(hy, hy, fma_(hz, hz, 1.0f))));
    const float n[4] = {fma_(-qx, hx, fma_(-qy, hy, fma_(-qz, hz, qw))), fma_(qw, hx, fma_(qy, hz, fma_(-qz, hy, qx))),
                        fma_(qw, hy, fma_(qz, hx, fma_(-qx, hz, qy))), fma_(qw, hz, fma_(qx, hy, fma_(-qy, hx, qz)))};
    const float q2 = fma_(qw, qw, fma_(qx, qx, fma_(qy, qy, qz * qz)));
    const float c = fma_(-0.5f, q2, 1.5f);
    const float sc = rs * c;
    const float B[4][4] = {{1.0f, hx, hy, hz}, {-hx, 1.0f, -hz, hy}, {-hy, hz, 1.0f, -hx}, {-hz, -hy, hx, 1.0f}};
    const float Cq[3][4] = {{-qx, qw, qz, -qy}, {-qy, -qz, qw, qx}, {-qz, qy, -qx, qw}};
    const float q[4] = {qw, qx, qy, qz}, hv[3] = {hx, hy, hz};
    const float r3c = (rs * rs * rs) * c;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            J[i * 4 + j] = fma_(sc, B[i][j], -(rs * q[i]) * n[j]);
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 4; ++j)
            J[16 + k * 4 + j] = h * fma_(sc, Cq[k][j], -(hv[k] * r3c) * n[j]);
    const float atd = (u[0] * C.inv_mass) * C.dt;
    const float D[4][3] = {{2.0f * qy, -2.0f * qx, 0.0f}, {2.0f * qz, -2.0f * qw, -4.0f * qx},
                           {2.0f * qw, 2.0f * qz, -4.0f * qy}, {2.0f * qx, 2.0f * qy, 0.0f}};
    for (int i = 0; i < 4; ++i)
        for (int a = 0; a < 3; ++a)
            J[28 + i * 3 + a] = atd * D[i][a];
    float b3[3];
    thrust_dir(x.q, b3);
    const float dm = C.dt * C.inv_mass;
    J[40] = dm * b3[0];
    J[41] = dm * b3[1];
    J[42] = dm * b3[2];
}

// One step of the adjoint through J (quad_step_jacobian): lam (13 floats:
// p, q, v, w, as grad_to_flat) after the step -> before it; gu = dJ/du.
DR_HD void quad_step_jvp_t(const QuadConst& C, const float* J, float lam[13], float gu[4])
{
    const float P[3] = {lam[0], lam[1], lam[2]};
    const float Qv[4] = {lam[3], lam[4], lam[5], lam[6]};
    const float V[3] = {lam[7], lam[8], lam[9]};
    const float W[3] = {lam[10], lam[11], lam[12]};
    gu[0] = fma_(J[40], V[0], fma_(J[41], V[1], J[42] * V[2]));
    const float kk = C.rate_k, dv = fma_(-C.drag, C.dt, 1.0f);
    for (int k = 0; k < 3; ++k)
    {
        gu[1 + k] = kk * W[k];
        float a = (1.0f - kk) * W[k];
        for (int j = 0; j < 4; ++j)
            a = fma_(J[16 + k * 4 + j], Qv[j], a);
        lam[10 + k] = a;
    }
    for (int i = 0; i < 4; ++i)
    {
        float a = 0.0f;
        for (int j = 0; j < 4; ++j)
            a = fma_(J[i * 4 + j], Qv[j], a);
        for (int b = 0; b < 3; ++b)
            a = fma_(J[28 + i * 3 + b], V[b], a);
        lam[3 + i] = a;
    }
    for (int b = 0; b < 3; ++b)
        lam[7 + b] = fma_(C.dt, P[b], dv * V[b]);
}

DR_HD float clamp_grad(float x, float lo, float hi)
{
    return (x >= lo && x <= hi) ? 1.0f : 0.0f;
}

} // namespace dr
