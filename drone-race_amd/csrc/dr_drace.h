// DroneRace environment (host and device): two agents race through gates
// on the plane; the opponent is a PD controller following one of several
// sampled racelines, model theta being its PD configuration.  Restated from
// include/environments/env_dronerace.h of pablo-gr-rey/drone-race:
//   agentSpeed, sampleCenterline               :43-66
//   projectOnTrack, findMinAlongDirection,
//   fastProjectOnTrack                         :69-209 (CHECK_PROJECTION off)
//   updateBelief                               :212-259
//   trackBoundaryDist, isWinner, isOppWinner   :262-370
//   getAnyAdvance, getMaxAccel                 :374-409
//   updateGates (no margin)                    :412-470
//   computePIDAction                           :473-541
//   environmentStep                            :546-680
// and the running / terminal costs of include/costs.cuh:7-39.  The counts the
// reference fixes at compile time (gates, obstacles, racelines, models) are
// fields of dr_drace_env.
//
// Arithmetic follows the reference expression by expression with no
// contraction.  Departures, each with the same value wherever the reference is
// defined: the Gaussian likelihood of updateBelief is formed in the log domain
// with the build's log and exp (as for StratRace2D); powf(x, y) of the
// repulsion is exp(y log x) with the build's functions; sampleCenterline holds
// its sample index at 511 where s - floor(s) rounds to 1 (the reference reads
// one sample past the raceline); the opponent's nominal action is formed for
// the followed model only where no belief is updated.  Environment noise is
// drawn from the same Philox sequence as the StratRace2D environment, in the
// reference's draw order.
#pragma once

#include "dr_sr.h"

namespace dr {
namespace drace {

constexpr int NS = DR_DRACE_SAMPLES;
constexpr int NR = DR_DRACE_MAX_RACELINES;

struct St
{
    float pos[4], vel[4], S[2][NR];
    int laps[2], gates[2];
};

// Register-resident access by a run-time agent or raceline index (selects,
// no indexed private memory).
DR_HD float agent_x(const float* a4, int agent, int d)
{
    return agent ? (d ? a4[3] : a4[2]) : (d ? a4[1] : a4[0]);
}

DR_HD float progress_of(const St& s, int agent, int r)
{
    float v = agent ? s.S[1][0] : s.S[0][0];
#pragma unroll
    for (int q = 1; q < NR; ++q)
        if (r == q)
            v = agent ? s.S[1][q] : s.S[0][q];
    return v;
}

DR_HD int int_of(const int* a2, int agent)
{
    return agent ? a2[1] : a2[0];
}

DR_HD St state_from_abi(const dr_drace_state& a)
{
    St s;
    for (int i = 0; i < 4; ++i)
    {
        s.pos[i] = a.pos[i];
        s.vel[i] = a.vel[i];
    }
    for (int g = 0; g < 2; ++g)
    {
        for (int r = 0; r < NR; ++r)
            s.S[g][r] = a.s[g][r];
        s.laps[g] = a.laps[g];
        s.gates[g] = a.gates[g];
    }
    return s;
}

DR_HD void state_to_abi(const St& s, dr_drace_state& a)
{
    for (int i = 0; i < 4; ++i)
    {
        a.pos[i] = s.pos[i];
        a.vel[i] = s.vel[i];
    }
    for (int g = 0; g < 2; ++g)
    {
        for (int r = 0; r < NR; ++r)
            a.s[g][r] = s.S[g][r];
        a.laps[g] = s.laps[g];
        a.gates[g] = s.gates[g];
    }
}

// the racelines [n_racelines][NS][2], in LDS or in the configuration
struct Lines
{
    const float* p;
};

DR_HD float sq_to_sample(const float* tp, int i, const float* pos)
{
    const float dx = tp[i * 2] - pos[0];
    const float dy = tp[i * 2 + 1] - pos[1];
    return dx * dx + dy * dy;
}

// projectOnTrack: the first sample of least squared distance; s = i / NS
DR_HD float project_full(const float* tp, const float* pos, float* bestDist)
{
    float bestS = 0.0f, bestD2 = 3.402823466e+38f;
    for (int i = 0; i < NS; ++i)
    {
        const float d2 = sq_to_sample(tp, i, pos);
        if (d2 < bestD2)
        {
            bestD2 = d2;
            bestS = (float)i * (1.0f / (float)NS);
        }
    }
    *bestDist = __builtin_sqrtf(bestD2);
    return bestS;
}

// findMinAlongDirection: walk while the distance decreases, then look
// `margin` = 10 more samples ahead (a strictly decreasing walk visits each
// sample at most once, so NS bounds it)
DR_HD int find_min_dir(const float* tp, int i0, const float* pos, int delta, float base, float* best)
{
    int bestI = i0;
    float bestD = base;
    int cur = (i0 + delta + NS) % NS;
    float curD = sq_to_sample(tp, cur, pos);
    for (int guard = 0; curD < bestD && guard < NS; ++guard)
    {
        bestI = cur;
        bestD = curD;
        cur = (cur + delta + NS) % NS;
        curD = sq_to_sample(tp, cur, pos);
    }
    for (int i = 0; i < 10; ++i)
    {
        curD = sq_to_sample(tp, cur, pos);
        if (curD < bestD)
        {
            bestI = cur;
            bestD = curD;
        }
        cur = (cur + delta + NS) % NS;
    }
    *best = bestD;
    return bestI;
}

// fastProjectOnTrack from the previous progress prevS (< 0: full scan)
DR_HD float fast_project(const float* tp, const float* pos, float prevS)
{
    float dist;
    if (prevS < 0.0f)
        return project_full(tp, pos, &dist);
    const int baseI = (int)(prevS * (float)NS + 0.5f) % NS;
    const float base = sq_to_sample(tp, baseI, pos);
    float bD, fD;
    const int bI = find_min_dir(tp, baseI, pos, -1, base, &bD);
    const int fI = find_min_dir(tp, baseI, pos, +1, base, &fD);
    return fD < bD ? (float)fI / (float)NS : (float)bI / (float)NS;
}

// sampleCenterline: linear interpolation at s wrapped to [0, 1)
DR_HD void sample_line(const float* tp, float s, float out[2])
{
    s = s - floorf(s);
    const float idx = s * (float)NS;
    int i0 = (int)idx;
    if (i0 > NS - 1)
        i0 = NS - 1;
    const int i1 = (i0 + 1) % NS;
    const float t = idx - (float)i0;
    for (int d = 0; d < 2; ++d)
        out[d] = (1.0f - t) * tp[i0 * 2 + d] + t * tp[i1 * 2 + d];
}

// computePIDAction of `agent` under PD configuration `pid`
DR_HD void pid_action(const Lines& L, const dr_drace_env& E, const St& s, int agent, const dr_drace_pid& pid, float out[2])
{
    const float* tp = L.p + pid.raceline_index * NS * 2;
    float target[2];
    sample_line(tp, progress_of(s, agent, pid.raceline_index) + E.target_distance, target);
    const float p0 = agent_x(s.pos, agent, 0), p1 = agent_x(s.pos, agent, 1);
    const float v0 = agent_x(s.vel, agent, 0), v1 = agent_x(s.vel, agent, 1);
    const float e0 = target[0] - p0, e1 = target[1] - p1;
    const float sqErr = e0 * e0 + e1 * e1;
    const float invDist = 1.0f / __builtin_sqrtf(sqErr + 1e-5f);
    const float dir0 = (target[0] - p0) * invDist, dir1 = (target[1] - p1) * invDist;
    float vPar = 0.0f; // summed from zero, as written (the sign of a zero sum)
    vPar = vPar + v0 * dir0;
    vPar = vPar + v1 * dir1;
    const float err0 = (target[0] - p0) * invDist, err1 = (target[1] - p1) * invDist;
    const float lat0 = v0 - vPar * err0, lat1 = v1 - vPar * err1;
    out[0] = pid.kp * err0 + pid.kd * (-lat0);
    out[1] = pid.kp * err1 + pid.kd * (-lat1);
    if (pid.repulsion_factor != 0.0f)
    {
        const int other = 1 - agent;
        const float d0 = agent_x(s.pos, other, 0) - p0, d1 = agent_x(s.pos, other, 1) - p1;
        const float dist = __builtin_sqrtf(d0 * d0 + d1 * d1) + 1e-8f;
        if (dist < pid.repulsion_dist_fact * E.drone_radius * 2.0f)
        {
            const float x = dist * 0.5f / E.drone_radius;
            const float sc = pid.repulsion_factor / expf_ieee((pid.repulsion_power + 1.0f) * logf_ieee(x));
            out[0] = out[0] - sc * d0;
            out[1] = out[1] - sc * d1;
        }
    }
    const float maxA = agent ? E.max_accel[1] : E.max_accel[0];
    const float sqA = out[0] * out[0] + out[1] * out[1];
    if (sqA > maxA * maxA)
    {
        const float f = maxA / __builtin_sqrtf(sqA);
        out[0] = out[0] * f;
        out[1] = out[1] * f;
    }
}

// the Gaussian constants of updateBelief per model: variance 0.2^2 +
// actionNoise^2 and its log
DR_HD void belief_consts(const dr_drace_env& E, int th, float* sq_sigma, float* log_sq)
{
    const float sigmaEnv = 0.2f;
    const float an = E.opp_pid[th].action_noise;
    *sq_sigma = sigmaEnv * sigmaEnv + an * an;
    *log_sq = logf_ieee(*sq_sigma);
}

// updateBelief: the opponent's applied action against each model's nominal
// action (scaled to its max acceleration); product and normalisation
// (uniform below 1e-20).  bc as for StratRace2D (NULL: formed here).
DR_HD void update_belief(const dr_drace_env& E, float* belief, const float obs[2], const float (*nom)[2], float maxA, const float* bc)
{
    const float log2pi = logf_ieee(2.0f * 3.14159265358979323846f);
    float sum = 0.0f;
#pragma unroll
    for (int th = 0; th < DR_SR_MAX_MODELS; ++th)
    {
        if (th >= E.n_models)
            break;
        const float sqA = nom[th][0] * nom[th][0] + nom[th][1] * nom[th][1];
        float scale = 1.0f;
        if (sqA > maxA * maxA)
            scale = maxA / __builtin_sqrtf(sqA);
        const float dx = obs[0] - nom[th][0] * scale;
        const float dy = obs[1] - nom[th][1] * scale;
        const float sqDist = dx * dx + dy * dy;
        float sqSigma, logSq;
        if (bc)
        {
            sqSigma = bc[th];
            logSq = bc[DR_SR_MAX_MODELS + th];
        }
        else
            belief_consts(E, th, &sqSigma, &logSq);
        const float ll = -0.5f * sqDist / sqSigma - log2pi - logSq;
        belief[th] = belief[th] * expf_ieee(ll);
        sum = sum + belief[th];
    }
    if (sum < 1e-20f)
    {
        const float uniform = 1.0f / (float)E.n_models;
#pragma unroll
        for (int th = 0; th < DR_SR_MAX_MODELS; ++th)
            if (th < E.n_models)
                belief[th] = uniform;
    }
    else
#pragma unroll
        for (int th = 0; th < DR_SR_MAX_MODELS; ++th)
            if (th < E.n_models)
                belief[th] = belief[th] / sum;
}

// squared distance from (x, y) to rectangle o (0 inside)
DR_HD float sq_rect(const dr_drace_env& E, int o, float x, float y)
{
    const float dx = fmaxf(0.0f, fmaxf(E.obstacles[o][0] - x, x - E.obstacles[o][2]));
    const float dy = fmaxf(0.0f, fmaxf(E.obstacles[o][1] - y, y - E.obstacles[o][3]));
    return dx * dx + dy * dy;
}

// trackBoundaryDist: walls, rectangles, round obstacles and the other agent
// (less a drone radius) around the controlled agent; -1 when lose_on_opp_win
// and the opponent has won
DR_HD float boundary_dist(const dr_drace_env& E, const St& s, bool lose_on_opp_win)
{
    const int me = E.i_mppi;
    if (lose_on_opp_win && int_of(s.laps, 1 - me) >= E.n_win_laps)
        return -1.0f;
    const float x = agent_x(s.pos, me, 0), y = agent_x(s.pos, me, 1);
    float minDist = __builtin_inff();
    minDist = fminf(minDist, fminf(x - E.arena_min[0], E.arena_max[0] - x));
    minDist = fminf(minDist, fminf(y - E.arena_min[1], E.arena_max[1] - y));
    for (int o = 0; o < E.n_obstacles; ++o)
        minDist = fminf(minDist, __builtin_sqrtf(sq_rect(E, o, x, y)));
    for (int o = 0; o < E.n_round_obstacles; ++o)
    {
        const float dx = x - E.round_centers[o][0], dy = y - E.round_centers[o][1];
        minDist = fminf(minDist, __builtin_sqrtf(dx * dx + dy * dy) - E.round_radius[o]);
    }
    const float ax = s.pos[0] - s.pos[2], ay = s.pos[1] - s.pos[3];
    minDist = fminf(minDist, __builtin_sqrtf(ax * ax + ay * ay) - E.drone_radius);
    return minDist;
}

// isWinner: the controlled agent's laps, or the opponent outside (its wall
// test adds the radius on both sides, as the reference writes it)
DR_HD bool is_winner(const dr_drace_env& E, const St& s)
{
    const int me = E.i_mppi;
    if (int_of(s.laps, me) >= E.n_win_laps)
        return true;
    const float x = agent_x(s.pos, 1 - me, 0), y = agent_x(s.pos, 1 - me, 1);
    const float r = E.drone_radius;
    if (x < E.arena_min[0] + r || x > E.arena_max[0] + r)
        return true;
    if (y < E.arena_min[1] + r || y > E.arena_max[1] + r)
        return true;
    for (int o = 0; o < E.n_obstacles; ++o)
        if (sq_rect(E, o, x, y) < r * r)
            return true;
    for (int o = 0; o < E.n_round_obstacles; ++o)
    {
        const float dx = x - E.round_centers[o][0], dy = y - E.round_centers[o][1];
        const float rr = r + E.round_radius[o];
        if (dx * dx + dy * dy < rr * rr)
            return true;
    }
    return false;
}

DR_HD bool is_opp_winner(const dr_drace_env& E, const St& s)
{
    return int_of(s.laps, 1 - E.i_mppi) >= E.n_win_laps;
}

// getAnyAdvance: gates + n_gates * laps + 0.8 (1 - |p - next| / |next - gate|)
DR_HD float advance(const dr_drace_env& E, const St& s, int agent)
{
    const int g = int_of(s.gates, agent);
    const int next = (g + 1) % E.n_gates;
    const float p[2] = {agent_x(s.pos, agent, 0), agent_x(s.pos, agent, 1)};
    float sqGate = 0.0f, sqCons = 0.0f;
    for (int d = 0; d < 2; ++d)
    {
        sqGate = sqGate + (E.gate_centers[next][d] - p[d]) * (E.gate_centers[next][d] - p[d]);
        sqCons = sqCons + (E.gate_centers[next][d] - E.gate_centers[g][d]) * (E.gate_centers[next][d] - E.gate_centers[g][d]);
    }
    return (float)(g + E.n_gates * int_of(s.laps, agent)) + 0.8f * (1.0f - __builtin_sqrtf(sqGate / sqCons));
}

// updateGates<false>: progress on every raceline (while tracked), then the
// crossing of each agent's next gate within its radius
DR_HD void update_gates(const Lines& L, const dr_drace_env& E, St& s, const float prev[4])
{
#pragma unroll
    for (int a = 0; a < 2; ++a)
    {
        if (s.S[a][0] >= 0.0f)
        {
#pragma unroll
            for (int r = 0; r < NR; ++r)
                if (r < E.n_racelines)
                    s.S[a][r] = fast_project(L.p + r * NS * 2, s.pos + a * 2, s.S[a][r]);
        }
        const int next = (s.gates[a] + 1) % E.n_gates;
        float num = 0.0f, den = 0.0f;
        for (int d = 0; d < 2; ++d)
        {
            num = num + E.gate_vectors[next][d] * (E.gate_centers[next][d] - prev[a * 2 + d]);
            den = den + E.gate_vectors[next][d] * (s.pos[a * 2 + d] - prev[a * 2 + d]);
        }
        if (__builtin_fabsf(den) < 1e-10f)
            continue;
        const float lambda = num / den;
        if (lambda < 0.0f || lambda > 1.0f)
            continue;
        float sq = 0.0f;
        for (int d = 0; d < 2; ++d)
        {
            const float dx = (1.0f - lambda) * prev[a * 2 + d] + lambda * s.pos[a * 2 + d] - E.gate_centers[next][d];
            sq = sq + dx * dx;
        }
        if (sq <= E.gate_radius[next] * E.gate_radius[next])
        {
            s.gates[a] = s.gates[a] + 1;
            if (s.gates[a] == E.n_gates)
            {
                s.gates[a] = 0;
                s.laps[a] = s.laps[a] + 1;
            }
        }
    }
}

// environmentStep<BELIEF, false, false, lose_on_opp_win> with applyNoise =
// apply_noise; returns SR_LOSE / SR_WIN / SR_OPP_WIN / SR_NONE.  With BELIEF
// and `update`, the belief is updated from the opponent's applied action;
// bc: see update_belief.
template <bool BELIEF>
DR_HD int env_step_t(const Lines& L, const dr_drace_env& E, int theta, const float ego[2], St& s, sr::NormalSeq& rng, bool lose_on_opp_win,
                     float* belief, const float* bc = nullptr, bool update = true, bool apply_noise = true)
{
    const int me = E.i_mppi, opp = 1 - me;
    const float prev[4] = {s.pos[0], s.pos[1], s.pos[2], s.pos[3]};
    float noms[DR_SR_MAX_MODELS][2];
    float nom0 = 0.0f, nom1 = 0.0f, an = 0.0f;
    const bool all = BELIEF && update;
#pragma unroll
    for (int th = 0; th < DR_SR_MAX_MODELS; ++th)
        if (th < E.n_models && (all || th == theta))
        {
            pid_action(L, E, s, opp, E.opp_pid[th], noms[th]);
            if (th == theta)
            {
                nom0 = noms[th][0];
                nom1 = noms[th][1];
                an = E.opp_pid[th].action_noise;
            }
        }
    // the opponent's applied action: its model's PD action plus its noise
    float o0 = nom0, o1 = nom1;
    if (apply_noise && an != 0.0f)
    {
        o0 = o0 + an * rng.next();
        o1 = o1 + an * rng.next();
    }
    float act[4];
    act[0] = me == 0 ? ego[0] : o0;
    act[1] = me == 0 ? ego[1] : o1;
    act[2] = me == 0 ? o0 : ego[0];
    act[3] = me == 0 ? o1 : ego[1];
#pragma unroll
    for (int a = 0; a < 2; ++a)
    {
        const float sq = act[a * 2] * act[a * 2] + act[a * 2 + 1] * act[a * 2 + 1];
        float factor = 1.0f;
        if (sq > E.max_accel[a] * E.max_accel[a])
            factor = E.max_accel[a] / __builtin_sqrtf(sq);
#pragma unroll
        for (int d = 0; d < 2; ++d)
        {
            float v = act[a * 2 + d] * factor;
            if (E.action_noise != 0.0f && apply_noise)
                v = v + rng.next() * E.action_noise;
            act[a * 2 + d] = v;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
        s.pos[i] = s.pos[i] + E.dt * s.vel[i];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        s.vel[i] = s.vel[i] + E.dt * act[i];
#pragma unroll
    for (int a = 0; a < 2; ++a)
    {
        const float spd = __builtin_sqrtf(s.vel[a * 2] * s.vel[a * 2] + s.vel[a * 2 + 1] * s.vel[a * 2 + 1]);
        if (spd > E.max_speed[a])
        {
            const float sc = E.max_speed[a] / spd;
            s.vel[a * 2] = s.vel[a * 2] * sc;
            s.vel[a * 2 + 1] = s.vel[a * 2 + 1] * sc;
        }
    }
    if ((E.pos_noise != 0.0f || E.speed_noise != 0.0f) && apply_noise)
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            if (E.pos_noise != 0.0f)
                s.pos[i] = s.pos[i] + rng.next() * E.pos_noise;
            if (E.speed_noise != 0.0f)
                s.vel[i] = s.vel[i] + rng.next() * E.speed_noise;
        }
    update_gates(L, E, s, prev);
    if (all)
    {
        const float obs[2] = {me == 0 ? act[2] : act[0], me == 0 ? act[3] : act[1]};
        update_belief(E, belief, obs, noms, opp ? E.max_accel[1] : E.max_accel[0], bc);
    }
    if (boundary_dist(E, s, lose_on_opp_win) < E.drone_radius)
        return sr::SR_LOSE;
    if (is_winner(E, s))
        return sr::SR_WIN;
    if (!lose_on_opp_win && boundary_dist(E, s, true) < E.drone_radius)
        return sr::SR_OPP_WIN;
    return sr::SR_NONE;
}

// stateCost (costs.cuh:7-28)
DR_HD float state_cost(const dr_drace_env& E, const dr_sr_cost& K, const St& s)
{
    float cost = 0.0f;
    const float bd = boundary_dist(E, s, false);
    const float danger = K.boundary_threshold_factor * E.drone_radius;
    if (bd < danger)
    {
        const float q = bd / E.drone_radius;
        cost = cost + K.boundary_cost * (q * q);
    }
    if (bd < E.drone_radius * K.coll_dist_factor)
        cost = cost + K.outside_cost;
    if (is_opp_winner(E, s))
        cost = cost + K.opp_win_cost;
    if (is_winner(E, s))
        cost = cost - K.win_cost;
    return cost;
}

// finalCost (costs.cuh:31-39)
DR_HD float final_cost(const dr_drace_env& E, const dr_sr_cost& K, const St& s)
{
    float cost = -K.final_adv_weight * advance(E, s, E.i_mppi);
    if (K.final_opp_adv_weight != 0.0f)
        cost = cost + K.final_opp_adv_weight * advance(E, s, 1 - E.i_mppi);
    return cost;
}

} // namespace drace
} // namespace dr
