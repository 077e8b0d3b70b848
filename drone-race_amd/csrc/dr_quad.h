// Quad13 environment policy: 13-state quaternion quadrotor, gate race,
// constant-velocity opponents.  Every expression is written in the order in
// which it must be evaluated: the engine is compiled with -ffp-contract=off
// and only the explicit fmaf calls fuse, so host and device agree bit for bit.
//
// The reference compiles one 2-D point-mass environment (EnvStratRace,
// include/environments/env_stratrace.h:711-894).  The quadrotor below is the
// environment the benchmark configurations call for; its gate bookkeeping
// follows the reference's gate logic written for any dimension:
//   gate crossing   include/environments/env_dronerace.h:412-470 (updateGates)
//   race progress   include/environments/env_dronerace.h:374-394 (getAnyAdvance)
// and its cost terms follow the structure of include/costs.cuh:7-39
// (stateCost / finalCost): a running cost evaluated on the post-step state,
// discounted by DECAY per step, plus a terminal progress cost.
#pragma once

#include <math.h>

#include "dr_rng.h"
#include "../../include/dr_api.h"

namespace dr {

// include/config.h:30
constexpr float DECAY = 0.97f;

enum Terminal : int
{
    TERM_NONE = 0,
    TERM_LOSE = 1,
    TERM_WIN = 2,
};

struct QState
{
    float p[3], q[4], v[3], w[3];
    int gate, laps;
};

// Constants derived once on the host from dr_quad_params (same float ops in
// the oracle): 1/m, dt/2 and dt*rate_gain.
struct QuadConst
{
    float dt, half_dt, rate_k, inv_mass, gravity, drag;
    float thrust_min, thrust_max, rate_max, drone_radius;
};

inline QuadConst make_quad_const(const dr_quad_params& P)
{
    QuadConst c;
    c.dt = P.dt;
    c.half_dt = 0.5f * P.dt;
    c.rate_k = P.dt * P.rate_gain;
    c.inv_mass = 1.0f / P.mass;
    c.gravity = P.gravity;
    c.drag = P.drag;
    c.thrust_min = P.thrust_min;
    c.thrust_max = P.thrust_max;
    c.rate_max = P.rate_max;
    c.drone_radius = P.drone_radius;
    return c;
}

// Cost constants derived once on the host (same float ops in the oracle):
// reciprocals replace the divisions of the running cost, far2 is the squared
// opponent distance beyond which the boundary distance cannot reach the
// penalty band (no square root needed there).
struct CostConst
{
    float r2_prox, inv_prox, danger, inv_danger, coll, far2;
};

inline CostConst make_cost_const(const dr_quad_params& P, const dr_cost_cfg& K)
{
    CostConst c;
    c.r2_prox = K.opp_prox_radius * K.opp_prox_radius;
    c.inv_prox = 1.0f / K.opp_prox_radius;
    c.danger = K.boundary_threshold_factor * P.drone_radius;
    c.inv_danger = 1.0f / c.danger;
    c.coll = P.drone_radius * K.coll_dist_factor;
    const float reach = (c.danger > c.coll ? c.danger : c.coll) + P.drone_radius;
    c.far2 = reach * reach;
    return c;
}

DR_HD float fma_(float a, float b, float c)
{
    return __builtin_fmaf(a, b, c);
}

DR_HD float clampf(float x, float lo, float hi)
{
    return fminf(fmaxf(x, lo), hi);
}

// Whether c holds on any active lane of the wave (on the host: c itself).  A
// branch on it is uniform, so a block that no lane needs is skipped rather
// than computed and discarded.
DR_HD bool any_lane(bool c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(c) != 0;
#else
    return c;
#endif
}

// Placed first in a block under any_lane: keeps the compiler from hoisting
// the block's work above the branch (computing it on every wave and
// selecting the result), which would undo the branch.
DR_HD void keep_branch()
{
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_sched_barrier(0);
#endif
}

// 1/sqrt(x) for normal x > 0 (the attitude renormalisation): a first
// estimate from the bit pattern with Lomont's constant 0x5f375a86 (C. Lomont,
// "Fast Inverse Square Root", Purdue University, 2003; relative error below
// 3.5e-2), then three Newton-Raphson steps for 1/y^2 = x written as a
// correction, y <- y + y (1/2 - (x/2) y^2) (error 1.8e-3, 4.6e-6, then FP32
// rounding): one product and two fused multiply-adds each, so the host forms
// the same bits.  The correction form keeps the last step's residual exact,
// so 1/sqrt(1) is exactly 1.  Twelve instructions in place of a correctly
// rounded square root and division.
DR_HD float rsqrt_nr(float x)
{
    const float hx = 0.5f * x;
    float y = u2f(0x5f375a86u - (f2u(x) >> 1));
    y = fma_(y, fma_(-hx, y * y, 0.5f), y);
    y = fma_(y, fma_(-hx, y * y, 0.5f), y);
    y = fma_(y, fma_(-hx, y * y, 0.5f), y);
    return y;
}

// Box clamp of a control (collective thrust, three body rates).  Plays the
// role of the acceleration clamp of env_stratrace.h:753-772 for this model.
DR_HD void clamp_control(const QuadConst& C, const float in[4], float out[4])
{
    out[0] = clampf(in[0], C.thrust_min, C.thrust_max);
    out[1] = clampf(in[1], -C.rate_max, C.rate_max);
    out[2] = clampf(in[2], -C.rate_max, C.rate_max);
    out[3] = clampf(in[3], -C.rate_max, C.rate_max);
}

// One explicit Euler step, in parts that read disjoint parts of the state:
//   translation  position from the old velocity, velocity from the old
//                attitude and velocity;
//   attitude     q <- q (x) (1, h w) / |(1, h w)| * (3 - |q|^2) / 2, h = dt/2:
//                the Euler step q + h q (x) (0, w) renormalised.  The norm of a
//                quaternion product is the product of the norms, so the
//                normaliser is 1/|(1, h w)| (rsqrt_nr) times the first-order
//                correction (3 - |q|^2)/2 of 1/|q| (exact to 1e-14 for the unit
//                attitude the recurrence keeps; a small initial error decays
//                quadratically).  The increment (h w, 1/|(1, h w)|) depends on
//                the body rates alone, so a caller may form it apart from the
//                attitude recurrence (attitude_increment, attitude_apply);
//   body rates   first-order tracking of the commanded rates.
// Products that feed a sum are fused (fmaf) in the order written.
// thrust direction R(q) e3
DR_HD void thrust_dir(const float q[4], float b3[3])
{
    const float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    b3[0] = 2.0f * fma_(qx, qz, qw * qy);
    b3[1] = 2.0f * fma_(qy, qz, -(qw * qx));
    b3[2] = fma_(-2.0f, fma_(qx, qx, qy * qy), 1.0f);
}

DR_HD void quad_translation_step(const QuadConst& C, const float q[4], float p[3], float v[3], const float u[4])
{
    float b3[3];
    thrust_dir(q, b3);
    const float at = u[0] * C.inv_mass;
    const float ax = fma_(at, b3[0], -(C.drag * v[0]));
    const float ay = fma_(at, b3[1], -(C.drag * v[1]));
    const float az = fma_(at, b3[2], -(C.drag * v[2])) - C.gravity;

    p[0] = fma_(C.dt, v[0], p[0]);
    p[1] = fma_(C.dt, v[1], p[1]);
    p[2] = fma_(C.dt, v[2], p[2]);

    v[0] = fma_(C.dt, ax, v[0]);
    v[1] = fma_(C.dt, ay, v[1]);
    v[2] = fma_(C.dt, az, v[2]);
}

// (h w, 1/|(1, h w)|) of the body rates w.
DR_HD float4 attitude_increment(const QuadConst& C, const float w[3])
{
    const float hx = C.half_dt * w[0], hy = C.half_dt * w[1], hz = C.half_dt * w[2];
    const float r2 = fma_(hx, hx, fma_(hy, hy, fma_(hz, hz, 1.0f)));
    return make_float4(hx, hy, hz, rsqrt_nr(r2));
}

// q <- q (x) (1, h w) * (rs (3 - |q|^2)/2) for the increment (h w, rs).
DR_HD void attitude_apply(float q[4], const float4 inc)
{
    const float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    const float nw = fma_(-qx, inc.x, fma_(-qy, inc.y, fma_(-qz, inc.z, qw)));
    const float nx = fma_(qw, inc.x, fma_(qy, inc.z, fma_(-qz, inc.y, qx)));
    const float ny = fma_(qw, inc.y, fma_(qz, inc.x, fma_(-qx, inc.z, qy)));
    const float nz = fma_(qw, inc.z, fma_(qx, inc.y, fma_(-qy, inc.x, qz)));
    const float q2 = fma_(qw, qw, fma_(qx, qx, fma_(qy, qy, qz * qz)));
    const float sc = inc.w * fma_(-0.5f, q2, 1.5f);
    q[0] = nw * sc;
    q[1] = nx * sc;
    q[2] = ny * sc;
    q[3] = nz * sc;
}

// first-order body-rate tracking
DR_HD void rates_step(const QuadConst& C, float w[3], const float u[4])
{
    w[0] = fma_(C.rate_k, u[1] - w[0], w[0]);
    w[1] = fma_(C.rate_k, u[2] - w[1], w[1]);
    w[2] = fma_(C.rate_k, u[3] - w[2], w[2]);
}

DR_HD void quad_attitude_step(const QuadConst& C, float q[4], float w[3], const float u[4])
{
    attitude_apply(q, attitude_increment(C, w));
    rates_step(C, w, u);
}

DR_HD void quad_step(const QuadConst& C, QState& s, const float u[4])
{
    quad_translation_step(C, s.q, s.p, s.v, u);
    quad_attitude_step(C, s.q, s.w, u);
}

// Gate crossing test on the segment p0 -> p1 (env_dronerace.h:412-470),
// written without the division: with num = n.(c - p0) and den = n.(p1 - p0),
// lambda = num/den lies in [0, 1] iff num and den have the same sign and
// |num| <= |den|, and den * (x - c) = den (p0 - c) + num (p1 - p0) for the
// crossing point x, so the radius test is |.|^2 <= (radius*margin)^2 den^2.
// Either crossing direction counts; |den| < 1e-10 (motion in the gate plane)
// never counts.
DR_HD bool gate_crossed(const dr_gate& g, float margin, const float p0[3], const float p1[3])
{
    const float ax = p1[0] - p0[0], ay = p1[1] - p0[1], az = p1[2] - p0[2];
    const float cx = g.center[0] - p0[0], cy = g.center[1] - p0[1], cz = g.center[2] - p0[2];
    const float num = fma_(g.normal[0], cx, fma_(g.normal[1], cy, g.normal[2] * cz));
    const float den = fma_(g.normal[0], ax, fma_(g.normal[1], ay, g.normal[2] * az));
    const bool cand = (den > 0.0f) ? (num >= 0.0f && num <= den) : (num <= 0.0f && num >= den);
    if (cand && fabsf(den) >= 1e-10f)
    {
        const float wx = fma_(num, ax, -(den * cx));
        const float wy = fma_(num, ay, -(den * cy));
        const float wz = fma_(num, az, -(den * cz));
        const float rm = g.radius * margin;
        return fma_(wx, wx, fma_(wy, wy, wz * wz)) <= (rm * rm) * (den * den);
    }
    return false;
}

// The gate counter's step: gate g (index next) crossed by the move p0 -> s.p.
DR_HD void cross_gate(const dr_gate& g, int next, float margin, const float p0[3], QState& s)
{
    if (gate_crossed(g, margin, p0, s.p))
    {
        s.gate = next;
        if (next == 0)
            s.laps = s.laps + 1;
    }
}

DR_HD void update_gate(const dr_gate* gates, int n_gates, float margin, const float p0[3], QState& s)
{
    int next = s.gate + 1;
    if (next == n_gates)
        next = 0;
    cross_gate(gates[next], next, margin, p0, s);
}

// The next gate held in registers for a single latency-bound rollout: the
// same test as update_gate, the gate reloaded only after a crossing.
struct NextGate
{
    dr_gate g;
    int next;
    DR_HD void load(const dr_gate* gates, int n_gates, int gate)
    {
        next = gate + 1 == n_gates ? 0 : gate + 1;
        g = gates[next];
    }
    DR_HD void update(const dr_gate* gates, int n_gates, float margin, const float p0[3], QState& s)
    {
        cross_gate(g, next, margin, p0, s);
        if (s.gate == next)
            load(gates, n_gates, s.gate);
    }
};

// Race progress (env_dronerace.h:374-394): gate + n_gates*laps +
// 0.8 * (1 - |p - c_next| / |c_next - c_gate|).
// Gate counter with the track's gates held in registers (center, normal,
// radius): update_gate's test for the move p0 -> s.p; the next gate is
// selected from the table, by selects over the whole of it so that it stays
// in registers, when one is crossed.  No memory access per step.
struct GateRegs
{
    float gv[DR_MAX_GATES][7];
    int n_gates, nx;
    dr_gate g;

    __device__ void pick()
    {
        float v[7];
#pragma unroll
        for (int i = 0; i < 7; ++i)
            v[i] = gv[0][i];
#pragma unroll
        for (int j = 1; j < DR_MAX_GATES; ++j)
#pragma unroll
            for (int i = 0; i < 7; ++i)
                v[i] = j == nx ? gv[j][i] : v[i];
        g.center[0] = v[0], g.center[1] = v[1], g.center[2] = v[2];
        g.normal[0] = v[3], g.normal[1] = v[4], g.normal[2] = v[5];
        g.radius = v[6];
        g._pad = 0.0f;
    }
    __device__ void init(const dr_track& tr, int gate)
    {
#pragma unroll
        for (int k = 0; k < DR_MAX_GATES; ++k)
        {
            const dr_gate& src = tr.gates[k];
            gv[k][0] = src.center[0], gv[k][1] = src.center[1], gv[k][2] = src.center[2];
            gv[k][3] = src.normal[0], gv[k][4] = src.normal[1], gv[k][5] = src.normal[2];
            gv[k][6] = src.radius;
        }
        n_gates = tr.n_gates;
        nx = gate + 1 == n_gates ? 0 : gate + 1;
        pick();
    }
    __device__ void step(float margin, const float p0[3], QState& s)
    {
        cross_gate(g, nx, margin, p0, s);
        if (s.gate == nx)
        {
            nx = s.gate + 1 == n_gates ? 0 : s.gate + 1;
            pick();
        }
    }
};

DR_HD float advance(const dr_gate* gates, int n_gates, const QState& s)
{
    int next = s.gate + 1;
    if (next == n_gates)
        next = 0;
    const dr_gate& a = gates[s.gate];
    const dr_gate& b = gates[next];
    const float gx = b.center[0] - s.p[0], gy = b.center[1] - s.p[1], gz = b.center[2] - s.p[2];
    const float cx = b.center[0] - a.center[0], cy = b.center[1] - a.center[1], cz = b.center[2] - a.center[2];
    const float sq_gate = gx * gx + gy * gy + gz * gz;
    const float sq_cons = cx * cx + cy * cy + cz * cz;
    return (float)s.gate + (float)(n_gates * s.laps) + 0.8f * (1.0f - __builtin_sqrtf(sq_gate / sq_cons));
}

// Running cost on the post-step state (costs.cuh:7-28 structure):
//   opponent proximity  opp_prox_cost * (1 - d/r_prox)^2 for each d < r_prox
//   boundary distance   bd = min(min_j d_j - drone_radius, z - ground_z)
//   near-boundary       boundary_cost * (1 - bd/danger)^2 for bd < danger
//   collision           outside_cost, terminal (TERM_LOSE), for bd < r*coll
//   winner              -win_cost, terminal (TERM_WIN), for laps >= win laps
// opp holds the opponents' predicted positions at this step (x, y, z, -).
// An opponent farther than sqrt(far2) cannot bring bd below the penalty band,
// so its distance is taken as +inf without a square root.
//
// stage_cost_body_at: every term but the winner test (*term is TERM_LOSE
// after a collision, else TERM_NONE; reads the position only);
// stage_cost_win: the winner test on its result; stage_cost_at: both.
template <class OppAt>
DR_HD float stage_cost_body_at(const QuadConst& C, const CostConst& Q, const dr_cost_cfg& K, const dr_track& tr, const QState& s,
                               int n_opp, OppAt opp_at, int* term)
{
    float cost = 0.0f;
    float min_sq = __builtin_inff();
    float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (n_opp > 0)
        o = opp_at(0);
    for (int j = 0; j < n_opp; ++j)
    {
        const float4 on = opp_at(j + 1 < n_opp ? j + 1 : j); // read ahead
        const float dx = s.p[0] - o.x;
        const float dy = s.p[1] - o.y;
        const float dz = s.p[2] - o.z;
        const float d2 = fma_(dx, dx, fma_(dy, dy, dz * dz));
        min_sq = fminf(min_sq, d2);
        const bool near = d2 < Q.r2_prox;
        if (any_lane(near))
        {
            keep_branch();
            if (near)
            {
                const float x = fma_(-__builtin_sqrtf(d2), Q.inv_prox, 1.0f);
                cost = fma_(K.opp_prox_cost, x * x, cost);
            }
        }
        o = on;
    }
    float d_opp = __builtin_inff();
    const bool reach = min_sq < Q.far2;
    if (any_lane(reach))
    {
        keep_branch();
        if (reach)
            d_opp = __builtin_sqrtf(min_sq) - C.drone_radius;
    }
    const float d_ground = s.p[2] - tr.ground_z;
    const float bd = fminf(d_opp, d_ground);
    if (bd < Q.danger)
    {
        const float x = fma_(-bd, Q.inv_danger, 1.0f);
        cost = fma_(K.boundary_cost, x * x, cost);
    }
    int t = TERM_NONE;
    if (bd < Q.coll)
    {
        cost = cost + K.outside_cost;
        t = TERM_LOSE;
    }
    *term = t;
    return cost;
}

DR_HD float stage_cost_win(const dr_cost_cfg& K, const dr_track& tr, int laps, float cost, int* term)
{
    if (*term == TERM_NONE && laps >= tr.n_win_laps)
    {
        cost = cost - K.win_cost;
        *term = TERM_WIN;
    }
    return cost;
}

template <class OppAt>
DR_HD float stage_cost_at(const QuadConst& C, const CostConst& Q, const dr_cost_cfg& K, const dr_track& tr, const QState& s, int n_opp,
                          OppAt opp_at, int* term)
{
    const float cost = stage_cost_body_at(C, Q, K, tr, s, n_opp, opp_at, term);
    return stage_cost_win(K, tr, s.laps, cost, term);
}

DR_HD float stage_cost(const QuadConst& C, const CostConst& Q, const dr_cost_cfg& K, const dr_track& tr, const QState& s,
                       const float4* opp, int n_opp, int* term)
{
    return stage_cost_at(C, Q, K, tr, s, n_opp, [opp](int j) { return opp[j]; }, term);
}

// PR-MPPI running cost and safety on the post-step state (costs.cuh:44-63
// and :78-86 structure): the terms of stage_cost without the collision
// cost (a collision still ends the rollout), and
//   safety = r*coll - bd   (> 0 iff the state is unsafe)
// with the exact boundary distance (the square root is always taken, since
// the safety value is used as a cost of its own).  opp_at(j) returns the
// predicted position of opponent j at this step.
template <class OppAt>
DR_HD float pr_stage_cost(const QuadConst& C, const CostConst& Q, const dr_cost_cfg& K, const dr_track& tr, const QState& s, int n_opp,
                          OppAt opp_at, float* safety, int* term)
{
    float cost = 0.0f;
    float min_sq = __builtin_inff();
#pragma unroll
    for (int j = 0; j < DR_MAX_WORLD_DRONES - 1; ++j)
    {
        if (j >= n_opp)
            break;
        const float4 o = opp_at(j);
        const float dx = s.p[0] - o.x;
        const float dy = s.p[1] - o.y;
        const float dz = s.p[2] - o.z;
        const float d2 = fma_(dx, dx, fma_(dy, dy, dz * dz));
        min_sq = fminf(min_sq, d2);
        const bool near = d2 < Q.r2_prox;
        if (any_lane(near))
        {
            keep_branch();
            if (near)
            {
                const float x = fma_(-__builtin_sqrtf(d2), Q.inv_prox, 1.0f);
                cost = fma_(K.opp_prox_cost, x * x, cost);
            }
        }
    }
    const float d_opp = __builtin_sqrtf(min_sq) - C.drone_radius;
    const float d_ground = s.p[2] - tr.ground_z;
    const float bd = fminf(d_opp, d_ground);
    if (bd < Q.danger)
    {
        const float x = fma_(-bd, Q.inv_danger, 1.0f);
        cost = fma_(K.boundary_cost, x * x, cost);
    }
    *safety = Q.coll - bd;
    int t = TERM_NONE;
    if (bd < Q.coll)
        t = TERM_LOSE;
    else if (s.laps >= tr.n_win_laps)
    {
        cost = cost - K.win_cost;
        t = TERM_WIN;
    }
    *term = t;
    return cost;
}

// Predicted opponent position after t+1 control steps: p + v * (dt * (t+1)).
DR_HD float4 opp_predict(const dr_drone_state& o, float tau)
{
    float4 r;
    r.x = fma_(o.v[0], tau, o.p[0]);
    r.y = fma_(o.v[1], tau, o.p[1]);
    r.z = fma_(o.v[2], tau, o.p[2]);
    r.w = 0.0f;
    return r;
}

// Terminal cost (costs.cuh:31-39): -final_adv_weight * advance.
DR_HD float final_cost(const dr_cost_cfg& K, const dr_track& tr, const QState& s)
{
    return -K.final_adv_weight * advance(tr.gates, tr.n_gates, s);
}

DR_HD QState state_from_abi(const dr_drone_state& d)
{
    QState s;
    for (int i = 0; i < 3; ++i)
    {
        s.p[i] = d.p[i];
        s.v[i] = d.v[i];
        s.w[i] = d.w[i];
    }
    for (int i = 0; i < 4; ++i)
        s.q[i] = d.q[i];
    s.gate = d.gate;
    s.laps = d.laps;
    return s;
}

DR_HD void state_to_abi(const QState& s, dr_drone_state& d)
{
    for (int i = 0; i < 3; ++i)
    {
        d.p[i] = s.p[i];
        d.v[i] = s.v[i];
        d.w[i] = s.w[i];
    }
    for (int i = 0; i < 4; ++i)
        d.q[i] = s.q[i];
    d.gate = s.gate;
    d.laps = s.laps;
}

// Closed-loop step of a controlled drone: clamp its action, one Euler step,
// gate update (SimulationEngine::dynStep, engine.cpp:286-312).
DR_HD void advance_state(QState& s, const float* act, const dr_track& tr, const QuadConst& C, float margin)
{
    float u[4];
    clamp_control(C, act, u);
    const float p0[3] = {s.p[0], s.p[1], s.p[2]};
    quad_step(C, s, u);
    update_gate(tr.gates, tr.n_gates, margin, p0, s);
}

DR_HD void advance_drone(dr_drone_state& st, const float* act, const dr_track& tr, const QuadConst& C, float margin)
{
    QState s = state_from_abi(st);
    advance_state(s, act, tr, C, margin);
    state_to_abi(s, st);
}

} // namespace dr
