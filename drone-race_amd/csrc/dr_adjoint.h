// Reverse-mode (adjoint) rollout of the Quad13 cost, J(U) = sum_t decay_t
// c(x_{t+1}) + c_f(x_end), for one control sequence U (T x 4) from a given
// state, noise-free.  The forward pass is dr_quad.h's; this file holds the
// vector-Jacobian products of its pieces:
//   clamp        1 inside [lo, hi] (bounds included), 0 outside
//   Euler step   quad_step's expressions (an fma differentiates as a*b + c)
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
        if (d2 < Q.r2_prox && d2 > 0.0f)
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

    // recompute the normalisation of the attitude update
    const float dqw = -fma_(qx, wx, fma_(qy, wy, qz * wz));
    const float dqx = fma_(qw, wx, fma_(qy, wz, -(qz * wy)));
    const float dqy = fma_(qw, wy, fma_(qz, wx, -(qx * wz)));
    const float dqz = fma_(qw, wz, fma_(qx, wy, -(qy * wx)));
    const float nw = fma_(C.half_dt, dqw, qw);
    const float nx = fma_(C.half_dt, dqx, qx);
    const float ny = fma_(C.half_dt, dqy, qy);
    const float nz = fma_(C.half_dt, dqz, qz);
    const float inv = rsqrt_nr(fma_(nw, nw, fma_(nx, nx, fma_(ny, ny, nz * nz))));
    const float o0 = nw * inv, o1 = nx * inv, o2 = ny * inv, o3 = nz * inv;

    // w' = w + k (u - w)
    QGrad g;
    const float kk = C.rate_k;
    gu[1] = kk * gx.w[0];
    gu[2] = kk * gx.w[1];
    gu[3] = kk * gx.w[2];
    g.w[0] = (1.0f - kk) * gx.w[0];
    g.w[1] = (1.0f - kk) * gx.w[1];
    g.w[2] = (1.0f - kk) * gx.w[2];

    // q' = n / |n|: dn = (dq' - q' (q'.dq')) / |n|
    const float dot = fma_(o0, gx.q[0], fma_(o1, gx.q[1], fma_(o2, gx.q[2], o3 * gx.q[3])));
    const float gn0 = (gx.q[0] - o0 * dot) * inv;
    const float gn1 = (gx.q[1] - o1 * dot) * inv;
    const float gn2 = (gx.q[2] - o2 * dot) * inv;
    const float gn3 = (gx.q[3] - o3 * dot) * inv;
    // n = q + h dq(q, w)
    const float h = C.half_dt;
    const float hw = h * gn0, hx = h * gn1, hy = h * gn2, hz = h * gn3;
    g.q[0] = gn0 + (wx * hx + wy * hy + wz * hz);
    g.q[1] = gn1 + (-wx * hw - wz * hy + wy * hz);
    g.q[2] = gn2 + (-wy * hw + wz * hx - wx * hz);
    g.q[3] = gn3 + (-wz * hw - wy * hx + wx * hy);
    g.w[0] = g.w[0] + (-qx * hw + qw * hx + qz * hy - qy * hz);
    g.w[1] = g.w[1] + (-qy * hw - qz * hx + qw * hy + qx * hz);
    g.w[2] = g.w[2] + (-qz * hw + qy * hx - qx * hy + qw * hz);

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

DR_HD float clamp_grad(float x, float lo, float hi)
{
    return (x >= lo && x <= hi) ? 1.0f : 0.0f;
}

} // namespace dr
