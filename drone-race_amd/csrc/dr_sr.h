// StratRace2D environment on the device: the reference's compiled 2-D race
// (one ego, one opponent following opponent model theta), restated from
// include/environments/env_stratrace.h of pablo-gr-rey/drone-race:
//   sampleTrackArray / sampleNormalizedSpeed   :53-76
//   fastProjectOnTrack                         :88-162
//   trackBoundaryDist, isWinner, isOppWinner   :226-275
//   getAnyAdvance                              :277-281
//   computeSdot                                :302-315
//   computeOppTargetLatDist (one opponent)     :318-461
//   computeOpponentAction                      :577-706
//   updateBelief                               :166-222
//   environmentStep                            :711-894
// and the running/terminal costs of include/costs.cuh:7-39 (and the PR-MPPI
// costs :44-86); findConfident of include/state.h:93-115.
//
// Arithmetic follows the reference expression by expression (compiled with
// -ffp-contract=off, so a*b + c is two roundings); the CPU oracle repeats
// the same sequence.  Departures, each with the same value wherever the
// reference is defined:
//   * expf is the build's exp (dr_rng.h expf_ieee), identical on host and device;
//   * powf(x, 2) is x * x (one rounding, the exact square rounded);
//   * a track index that would reach 512 (s / L rounding up to 1) is held at
//     511, where the reference reads one past its grid;
//   * the opponent's nominal action is computed for the followed model only
//     where no belief is updated (the other models' actions feed the belief
//     update alone); the model-independent part of the opponent controller
//     is formed once per step and shared;
//   * the tangent and curvature sampled and never used at env_stratrace.h:368
//     and :783-786 are skipped,
//     and computeSdot reuses the tangent its caller has just sampled at the
//     same s (same function of the same input).
// Environment noise draws come from Philox (counter (k, 4 t + block, step,
// STREAM_ENV)) in the reference's draw order; which draws exist depends on
// which noise levels are nonzero, as in the reference.
#pragma once

#include "dr_rng.h"
#include "../../include/dr_api.h"

namespace dr {
namespace sr {

constexpr int NS = DR_SR_TRACK_SAMPLES;
static_assert((NS & (NS - 1)) == 0, "track sample count must be a power of two");
constexpr int WINDOW = 10; // env_stratrace_defs.h:39

enum : int
{
    SR_NONE = 0,
    SR_LOSE = 1,
    SR_WIN = 2,
    SR_OPP_WIN = 3,
};

// Track tables (in LDS in the rollout kernel).
struct Tables
{
    const float* p;     // [NS][2]
    const float* t;     // [NS][2]
    const float* kappa; // [NS]
    float L;
};

struct St
{
    float pos[4], vel[4], S[2], lat[2];
    int laps[2];
};

DR_HD float sq_norm(float x, float y)
{
    return x * x + y * y;
}

// Index of the grid cell of s in [0, L): s / L wrapped to [0, 1), times NS.
DR_HD void wrap_index(float s, float L, int* i0, float* frac)
{
    s = s / L;
    s = s - __builtin_floorf(s);
    const float f = s * (float)NS;
    int i = 0;
    if (f >= 0.0f)
        i = f < (float)NS ? (int)f : NS - 1;
    *i0 = i;
    *frac = f - (float)i;
}

DR_HD void sample2(const float* arr, float s, float L, float out[2])
{
    int i0;
    float t;
    wrap_index(s, L, &i0, &t);
    const int i1 = (i0 + 1) & (NS - 1);
    out[0] = (1.0f - t) * arr[i0 * 2] + t * arr[i1 * 2];
    out[1] = (1.0f - t) * arr[i0 * 2 + 1] + t * arr[i1 * 2 + 1];
}

DR_HD float sample1(const float* arr, float s, float L)
{
    int i0;
    float t;
    wrap_index(s, L, &i0, &t);
    const int i1 = (i0 + 1) & (NS - 1);
    return (1.0f - t) * arr[i0] + t * arr[i1];
}

DR_HD void sample_tangent(const Tables& tb, float s, float out[2])
{
    sample2(tb.t, s, tb.L, out);
    const float inv = 1.0f / __builtin_sqrtf(sq_norm(out[0], out[1]));
    out[0] = out[0] * inv;
    out[1] = out[1] * inv;
}

DR_HD float sq_dist_to_sample(const float* p, int i, float x, float y)
{
    const float dx = p[i * 2] - x, dy = p[i * 2 + 1] - y;
    return dx * dx + dy * dy;
}

// fastProjectOnTrack<false>: arc length of the closest point within +-WINDOW
// samples of prevS, refined on the two adjacent segments; lat = signed
// distance (positive left of the tangent).
DR_HD float project(const Tables& tb, float x, float y, float prevS, float* lat)
{
    // the window is centred on (int)(wrapped s * 512), which may be 512 (= 0 mod 512)
    float w = prevS / tb.L;
    w = w - __builtin_floorf(w);
    const float fw = w * (float)NS;
    const int idxPrev = (fw >= 0.0f && fw <= (float)NS) ? (int)fw : 0;
    int best = (idxPrev - WINDOW + NS) & (NS - 1); // operands >= 0, NS a power of two
    float bestSq = sq_dist_to_sample(tb.p, best, x, y);
    for (int off = NS - WINDOW + 1; off <= NS + WINDOW; ++off)
    {
        const int i = (idxPrev + off) & (NS - 1);
        const float d = sq_dist_to_sample(tb.p, i, x, y);
        if (d < bestSq)
        {
            bestSq = d;
            best = i;
        }
    }
    float segSq = __builtin_inff(), bestS = 0.0f, cx = 0.0f, cy = 0.0f;
    for (int seg = -1; seg <= 0; ++seg)
    {
        const int a = (best + seg + NS) & (NS - 1);
        const int b = (a + 1) & (NS - 1);
        const float ax = tb.p[a * 2], ay = tb.p[a * 2 + 1];
        const float abx = tb.p[b * 2] - ax;
        const float aby = tb.p[b * 2 + 1] - ay;
        const float ab2 = sq_norm(abx, aby);
        float t = 0.0f;
        if ((double)ab2 > 1e-12)
        {
            t = ((x - ax) * abx + (y - ay) * aby) / ab2;
            t = fmaxf(0.0f, fminf(t, 1.0f));
        }
        const float px = ax + t * abx;
        const float py = ay + t * aby;
        const float d = sq_norm(x - px, y - py);
        if (d < segSq)
        {
            segSq = d;
            bestS = ((float)a + t) / (float)NS * tb.L;
            cx = px;
            cy = py;
        }
    }
    float tg[2];
    sample2(tb.t, bestS, tb.L, tg);
    const float cross = -tg[1] * (x - cx) + tg[0] * (y - cy);
    *lat = __builtin_copysignf(__builtin_sqrtf(segSq), cross);
    return bestS;
}

// computeSdot with the unit tangent at S[i] already sampled.
DR_HD float sdot_with(const Tables& tb, const St& s, int i, const float tang[2])
{
    const float k = sample1(tb.kappa, s.S[i], tb.L);
    float metric = 1.0f - k * s.lat[i];
    if (__builtin_fabsf(metric) < 1e-12f)
        metric = __builtin_copysignf(1e-12f, metric);
    return (tang[0] * s.vel[2 * i] + tang[1] * s.vel[2 * i + 1]) / metric;
}

DR_HD float sdot(const Tables& tb, const St& s, int i)
{
    float tang[2];
    sample_tangent(tb, s.S[i], tang);
    return sdot_with(tb, s, i, tang);
}

// computeSdot with the unit tangent and the curvature at S[i] already sampled.
DR_HD float sdot_tk(const St& s, int i, const float tang[2], float k)
{
    float metric = 1.0f - k * s.lat[i];
    if (__builtin_fabsf(metric) < 1e-12f)
        metric = __builtin_copysignf(1e-12f, metric);
    return (tang[0] * s.vel[2 * i] + tang[1] * s.vel[2 * i + 1]) / metric;
}

// The model-independent part of computeOppTargetLatDist and
// computeOpponentAction at one state (single opponent, iOpp = 1): track
// samples at both agents, their progress rates, the opponent's track frame
// and lateral limits.  Every opponent model's target and action reuse it,
// and the environment step's progress update reuses the two rates (they read
// S, lat and vel only, which the position update leaves unchanged).
struct OppGeom
{
    float dAhead, dBehind, negQuarterL, sdot_i, sdot_j, tang_j[2], deltaE, blockTol, oppLongSpeed, maxLat;
    float tang[2], nx, ny, kappa, e_n, v_t, v_n, e_n_max, metric, a_cent, a_lat_max, d_pos, d_neg;
};

DR_HD OppGeom opp_geom(const Tables& tb, const dr_sr_env& E, const St& s)
{
    OppGeom g;
    const float L = tb.L;
    g.dAhead = s.S[0] - s.S[1];
    if (g.dAhead < 0.0f)
        g.dAhead = g.dAhead + L;
    g.dBehind = s.S[0] - s.S[1];
    if (g.dBehind > 0.0f)
        g.dBehind = g.dBehind - L;
    g.negQuarterL = -L * 0.25f;
    // opponent: tangent and curvature at S[1]
    const float s_now = s.S[1];
    float pc[2];
    sample2(tb.p, s_now, tb.L, pc);
    sample_tangent(tb, s_now, g.tang);
    g.kappa = sample1(tb.kappa, s_now, tb.L);
    g.sdot_i = sdot_tk(s, 1, g.tang, g.kappa);
    // ego: tangent and curvature at S[0]
    sample_tangent(tb, s.S[0], g.tang_j);
    g.sdot_j = sdot_tk(s, 0, g.tang_j, sample1(tb.kappa, s.S[0], tb.L));
    g.blockTol = E.max_speed[0] * 0.3f;
    g.deltaE = s.lat[0] - s.lat[1];
    g.oppLongSpeed = s.vel[0] * g.tang_j[0] + s.vel[1] * g.tang_j[1];
    g.maxLat = E.track_width - E.drone_radius * E.max_opp_lat_dist_fact;

    const float maxAccel = E.max_accel[1];
    const float px = s.pos[2], py = s.pos[3];
    const float vx = s.vel[2], vy = s.vel[3];
    g.nx = -g.tang[1];
    g.ny = g.tang[0];
    g.e_n = (px - pc[0]) * g.nx + (py - pc[1]) * g.ny;
    g.v_t = vx * g.tang[0] + vy * g.tang[1];
    g.v_n = vx * g.nx + vy * g.ny;
    g.e_n_max = E.track_width - E.drone_radius * 1.2f;
    g.metric = 1.0f - g.kappa * g.e_n;
    if (__builtin_fabsf(g.metric) < 1e-6f)
        g.metric = __builtin_copysignf(1e-6f, g.metric);
    g.a_cent = g.v_t * g.v_t * g.kappa / g.metric;
    g.a_lat_max = fmaxf(maxAccel - __builtin_fabsf(g.a_cent), maxAccel * 0.2f);
    g.d_pos = fmaxf(g.e_n_max - g.e_n, 0.001f);
    g.d_neg = fmaxf(g.e_n_max + g.e_n, 0.001f);
    return g;
}

// computeOppTargetLatDist (env_stratrace.h:318-461) of one opponent model:
// the only other agent ahead and behind is the ego (j = 0).
DR_HD void opp_target(const dr_sr_env& E, const dr_sr_opp_config& cfg, const St& s, const OppGeom& g, float* latDis_out,
                      float* speedGoal_out)
{
    float latDis = 0.0f;
    const float baseSpeed = cfg.speed_scale * E.max_speed[1];
    float speedGoal = baseSpeed;
#pragma unroll
    for (int k = 0; k <= 1; ++k)
    {
        const float deltaS = k ? g.dAhead : g.dBehind;
        const bool shouldBlock = (deltaS < 0.0f) && (deltaS > g.negQuarterL) && (g.sdot_j + g.blockTol > g.sdot_i);
        const bool shouldOvertake = (g.sdot_i > g.sdot_j) && (__builtin_fabsf(g.deltaE) < cfg.s1);
        const float factor = expf_ieee(-cfg.s2 * deltaS * deltaS);
        if (shouldBlock)
        {
            const float blockTerm = s.lat[0] * (1.0f - expf_ieee(-cfg.s3 * (g.sdot_j + g.blockTol - g.sdot_i))) * factor;
            latDis = latDis + blockTerm;
        }
        if (shouldOvertake)
        {
            const float overtakeTerm = (s.lat[0] - __builtin_copysignf(cfg.s1, g.deltaE)) * factor;
            latDis = latDis + overtakeTerm;
            if (deltaS < 3.0f * E.drone_radius)
                speedGoal = fminf(speedGoal, 0.7f * g.oppLongSpeed);
            else if (deltaS < 6.0f * E.drone_radius && (double)cfg.s1 > 1e-6)
            {
                const float latFactor = __builtin_fabsf(g.deltaE) / cfg.s1;
                speedGoal = fminf(speedGoal, latFactor * g.oppLongSpeed + (1.0f - latFactor) * baseSpeed);
            }
        }
    }
    *latDis_out = fmaxf(-g.maxLat, fminf(g.maxLat, latDis));
    *speedGoal_out = speedGoal;
}

// computeOpponentAction (env_stratrace.h:577-706, the active "second try"
// controller) of one opponent model, from the shared geometry.
DR_HD void opp_action_g(const dr_sr_env& E, const dr_sr_opp_config& cfg, const St& s, const OppGeom& g, float out[2])
{
    const float maxAccel = E.max_accel[1];
    const float dt = E.dt;
    float latDisRaw, speedGoal;
    opp_target(E, cfg, s, g, &latDisRaw, &speedGoal);

    const float a_lat_max = g.a_lat_max, v_n = g.v_n;
    const float latDis = fmaxf(-g.e_n_max, fminf(g.e_n_max, latDisRaw));
    const float e_err = latDis - g.e_n;

    float a_n;
    if (v_n > 0.0f && v_n * v_n > 1.8f * a_lat_max * g.d_pos)
        a_n = -a_lat_max;
    else if (v_n < 0.0f && v_n * v_n > 1.8f * a_lat_max * g.d_neg)
        a_n = a_lat_max;
    else if (__builtin_fabsf(e_err) < 0.02f)
    {
        const float kp = a_lat_max / 0.02f;
        a_n = kp * e_err - 2.0f * __builtin_sqrtf(kp) * v_n;
        a_n = fmaxf(-a_lat_max, fminf(a_lat_max, a_n));
    }
    else
    {
        const float stopping = v_n * v_n / (2.0f * a_lat_max);
        const bool toward = (e_err > 0.0f && v_n > 0.0f) || (e_err < 0.0f && v_n < 0.0f);
        const bool decel = toward && stopping >= __builtin_fabsf(e_err) * 0.95f;
        a_n = (e_err > 0.0f) ? a_lat_max : -a_lat_max;
        if (decel)
            a_n = -a_n;
        const float v_next = v_n + a_n * dt;
        if (v_next > 0.0f)
        {
            const float v_safe = __builtin_sqrtf(2.0f * a_lat_max * g.d_pos);
            if (v_next > v_safe)
                a_n = (v_safe - v_n) / dt;
        }
        else
        {
            const float v_safe = __builtin_sqrtf(2.0f * a_lat_max * g.d_neg);
            if (-v_next > v_safe)
                a_n = (-v_safe - v_n) / dt;
        }
    }

    const float kappa = g.kappa, v_t = g.v_t;
    const float a_n_tot = fmaxf(-maxAccel, fminf(maxAccel, a_n + g.a_cent));
    const float a_t_max = __builtin_sqrtf(fmaxf(maxAccel * maxAccel - a_n_tot * a_n_tot, 0.0f));
    const float kappa_eff = kappa / (1.0f - kappa * latDis + 1e-6f);
    float pathSpeed = fminf(speedGoal, __builtin_sqrtf(maxAccel / fmaxf(__builtin_fabsf(kappa_eff), 1e-12f)));
    if (__builtin_fabsf(a_n) > a_lat_max * 0.5f)
        pathSpeed = pathSpeed * fmaxf(0.4f, 1.5f - __builtin_fabsf(a_n) / a_lat_max);
    float a_t = E.kv * (pathSpeed - v_t) - v_n * kappa * v_t / g.metric;
    a_t = fmaxf(-a_t_max, fminf(a_t_max, a_t));
    out[0] = a_t * g.tang[0] + a_n_tot * g.nx;
    out[1] = a_t * g.tang[1] + a_n_tot * g.ny;
}

DR_HD void opp_action(const Tables& tb, const dr_sr_env& E, const dr_sr_opp_config& cfg, const St& s, float out[2])
{
    const OppGeom g = opp_geom(tb, E, s);
    opp_action_g(E, cfg, s, g, out);
}

// Sequential standard normals of one (sample, step): block b of four from
// Philox counter (k, 4 t + b, step, stream) (STREAM_ENV unless stated).
struct NormalSeq
{
    uint32_t key0, key1, k, c1, step, stream;
    int n;
    float z0, z1, z2, z3;
    DR_HD float next()
    {
        const int j = n & 3;
        if (j == 0)
        {
            float z[4];
            normal4(key0, key1, k, c1 + (uint32_t)(n >> 2), step, stream, z);
            z0 = z[0];
            z1 = z[1];
            z2 = z[2];
            z3 = z[3];
        }
        n = n + 1;
        return j == 0 ? z0 : (j == 1 ? z1 : (j == 2 ? z2 : z3));
    }
};

DR_HD NormalSeq normal_seq(uint32_t key0, uint32_t key1, uint32_t k, uint32_t t, uint32_t step, uint32_t stream = STREAM_ENV)
{
    NormalSeq q;
    q.key0 = key0;
    q.key1 = key1;
    q.k = k;
    q.c1 = 4u * t;
    q.step = step;
    q.stream = stream;
    q.n = 0;
    q.z0 = q.z1 = q.z2 = q.z3 = 0.0f;
    return q;
}

// trackBoundaryDist<lose>
DR_HD float boundary_dist(const dr_sr_env& E, const St& s, bool lose_on_opp_win)
{
    if (lose_on_opp_win && s.laps[1] >= E.n_win_laps)
        return -1.0f;
    const float minSq = fminf(__builtin_inff(), sq_norm(s.pos[2] - s.pos[0], s.pos[3] - s.pos[1]));
    return fminf(__builtin_sqrtf(minSq) - E.drone_radius, E.track_width - __builtin_fabsf(s.lat[0]));
}

DR_HD bool is_winner(const dr_sr_env& E, const St& s)
{
    return s.laps[0] >= E.n_win_laps || __builtin_fabsf(s.lat[1]) > E.track_width - E.drone_radius;
}

DR_HD bool is_opp_winner(const dr_sr_env& E, const St& s)
{
    return s.laps[1] >= E.n_win_laps;
}

// Per-model constants of updateBelief: sqSigma = 0.2^2 + actionNoise^2 and
// log(sqSigma).
DR_HD void belief_consts(const dr_sr_env& E, int th, float* sq_sigma, float* log_sq)
{
    const float sigmaEnv = 0.2f;
    const float an = E.opp[th].action_noise;
    *sq_sigma = sigmaEnv * sigmaEnv + an * an;
    *log_sq = logf_ieee(*sq_sigma);
}

// updateBelief (env_stratrace.h:166-222) for one opponent: the observed
// opponent action against each model's nominal action (scaled to the
// opponent's max acceleration), Gaussian likelihood with variance
// 0.2^2 + actionNoise^2, product and normalisation (uniform below 1e-20).
// log and exp are the build's (dr_rng.h).  bc = {sqSigma[4], log sqSigma[4]}
// from belief_consts, or NULL to form them here.
DR_HD void update_belief(const dr_sr_env& E, float* belief, const float obs[2], const float (*nom)[2], const float* bc = nullptr)
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
        if (sqA > E.max_accel[1] * E.max_accel[1])
            scale = E.max_accel[1] / __builtin_sqrtf(sqA);
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

// findConfident (state.h:93-115) for one model factor: theta + 1 for the
// first model whose belief exceeds the threshold, else 0.
DR_HD int find_confident(const float* belief, int n_models, float threshold)
{
    for (int th = 0; th < n_models; ++th)
        if (belief[th] > threshold)
            return th + 1;
    return 0;
}

// environmentStep<BELIEF, false, false, lose_on_opp_win> with applyNoise =
// apply_noise (every noise draw is skipped when false: the display
// predictions of SimulationEngine::sendState).  With BELIEF and `update`, the
// nominal opponent action of every model is formed and the belief is
// updated from the opponent's applied action; with `update` false the step
// is the plain one (a caller that no longer reads the belief skips its
// update).  bc: see update_belief.
template <bool BELIEF>
DR_HD int env_step_t(const Tables& tb, const dr_sr_env& E, int theta, const float ego[2], St& s, NormalSeq& rng, bool lose_on_opp_win,
                     float* belief, const float* bc = nullptr, bool update = true, bool apply_noise = true)
{
    const float prevS0 = s.S[0], prevS1 = s.S[1];
    const OppGeom g = opp_geom(tb, E, s);
    float act[4];
    act[0] = ego[0];
    act[1] = ego[1];
    float noms[DR_SR_MAX_MODELS][2];
    const bool upd = BELIEF && update;
    float nom[2];
    if (upd)
    {
        // constant indices only: the model tables stay in registers
#pragma unroll
        for (int th = 0; th < DR_SR_MAX_MODELS; ++th)
            if (th < E.n_models)
            {
                opp_action_g(E, E.opp[th], s, g, noms[th]);
                if (th == theta)
                {
                    nom[0] = noms[th][0];
                    nom[1] = noms[th][1];
                }
            }
    }
    else
        opp_action_g(E, E.opp[theta], s, g, nom);
    {
        const dr_sr_opp_config& cfg = E.opp[theta];
#pragma unroll
        for (int d = 0; d < 2; ++d)
        {
            float u = nom[d];
            if (apply_noise && cfg.action_noise != 0.0f)
                u = u + cfg.action_noise * rng.next();
            act[2 + d] = u;
        }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
    {
        const float sq = act[2 * a] * act[2 * a] + act[2 * a + 1] * act[2 * a + 1];
        float factor = 1.0f;
        const float maxSq = E.max_accel[a] * E.max_accel[a];
        if (sq > maxSq)
            factor = E.max_accel[a] / __builtin_sqrtf(sq);
        for (int d = 0; d < 2; ++d)
        {
            float v = act[2 * a + d] * factor;
            if (apply_noise && E.action_noise != 0.0f)
                v = v + rng.next() * E.action_noise;
            act[2 * a + d] = v;
        }
    }
    for (int i = 0; i < 4; ++i)
        s.pos[i] = s.pos[i] + E.dt * s.vel[i];
    // progress rates of the state before the position update (computeSdot
    // reads S, lat and vel only)
    s.S[0] = s.S[0] + E.dt * g.sdot_j;
    s.S[1] = s.S[1] + E.dt * g.sdot_i;
    for (int i = 0; i < 4; ++i)
        s.vel[i] = s.vel[i] + E.dt * act[i];
#pragma unroll
    for (int a = 0; a < 2; ++a)
    {
        const float sq = sq_norm(s.vel[2 * a], s.vel[2 * a + 1]);
        if (sq > E.max_speed[a] * E.max_speed[a])
        {
            const float sc = E.max_speed[a] / __builtin_sqrtf(sq);
            s.vel[2 * a] = s.vel[2 * a] * sc;
            s.vel[2 * a + 1] = s.vel[2 * a + 1] * sc;
        }
    }
    if (apply_noise && (E.pos_noise != 0.0f || E.speed_noise != 0.0f))
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            if (E.pos_noise != 0.0f)
                s.pos[i] = s.pos[i] + rng.next() * E.pos_noise;
            if (E.speed_noise != 0.0f)
                s.vel[i] = s.vel[i] + rng.next() * E.speed_noise;
        }
    {
        float lat0, lat1;
        s.S[0] = project(tb, s.pos[0], s.pos[1], s.S[0], &lat0);
        s.lat[0] = lat0;
        s.S[1] = project(tb, s.pos[2], s.pos[3], s.S[1], &lat1);
        s.lat[1] = lat1;
    }
    const float prev[2] = {prevS0, prevS1};
#pragma unroll
    for (int a = 0; a < 2; ++a)
    {
        if (s.S[a] < prev[a] - tb.L * 0.5f)
            s.laps[a] = s.laps[a] + 1;
        if (s.S[a] > prev[a] + tb.L * 0.5f)
            s.laps[a] = s.laps[a] - 1;
    }
    if (upd)
    {
        const float obs[2] = {act[2], act[3]};
        update_belief(E, belief, obs, noms, bc);
    }
    if (boundary_dist(E, s, lose_on_opp_win) < E.drone_radius)
        return SR_LOSE;
    if (is_winner(E, s))
        return SR_WIN;
    if (!lose_on_opp_win && boundary_dist(E, s, true) < E.drone_radius)
        return SR_OPP_WIN;
    return SR_NONE;
}


DR_HD int env_step(const Tables& tb, const dr_sr_env& E, int theta, const float ego[2], St& s, NormalSeq& rng, bool lose_on_opp_win)
{
    return env_step_t<false>(tb, E, theta, ego, s, rng, lose_on_opp_win, nullptr);
}

// stateCost (costs.cuh:7-28).
DR_HD float state_cost(const dr_sr_env& E, const dr_sr_cost& K, const St& s)
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

// finalCost (costs.cuh:31-39).
DR_HD float final_cost(const dr_sr_env& E, const dr_sr_cost& K, const St& s, float L)
{
    const float adv = s.S[0] / L + (float)s.laps[0];
    float cost = -K.final_adv_weight * adv;
    if (K.final_opp_adv_weight != 0.0f)
    {
        const float oppAdv = fmaxf(-__builtin_inff(), s.S[1] / L + (float)s.laps[1]);
        cost = cost + K.final_opp_adv_weight * oppAdv;
    }
    return cost;
}

// PRMPPIstateCost (costs.cuh:44-63): boundary band with the opponent-win
// rule of trackBoundaryDist, win and opponent-win terms; no collision cost.
DR_HD float pr_state_cost(const dr_sr_env& E, const dr_sr_cost& K, const St& s)
{
    float cost = 0.0f;
    const float bd = boundary_dist(E, s, true);
    const float danger = K.boundary_threshold_factor * E.drone_radius;
    if (bd < danger)
    {
        const float q = bd / E.drone_radius;
        cost = cost + K.boundary_cost * (q * q);
    }
    if (is_winner(E, s))
        cost = cost - K.win_cost;
    if (is_opp_winner(E, s))
        cost = cost + K.opp_win_cost;
    return cost;
}

// PRMPPIsafetyCost (costs.cuh:78-86): > 0 iff the state is unsafe.
DR_HD float pr_safety(const dr_sr_env& E, const dr_sr_cost& K, const St& s)
{
    return E.drone_radius * K.coll_dist_factor - boundary_dist(E, s, false);
}

DR_HD St state_from_abi(const dr_sr_state& a)
{
    St s;
    for (int i = 0; i < 4; ++i)
    {
        s.pos[i] = a.pos[i];
        s.vel[i] = a.vel[i];
    }
    for (int i = 0; i < 2; ++i)
    {
        s.S[i] = a.s[i];
        s.lat[i] = a.lat_dist[i];
        s.laps[i] = a.laps[i];
    }
    return s;
}

DR_HD void state_to_abi(const St& s, dr_sr_state& a)
{
    for (int i = 0; i < 4; ++i)
    {
        a.pos[i] = s.pos[i];
        a.vel[i] = s.vel[i];
    }
    for (int i = 0; i < 2; ++i)
    {
        a.s[i] = s.S[i];
        a.lat_dist[i] = s.lat[i];
        a.laps[i] = s.laps[i];
    }
}

} // namespace sr
} // namespace dr
