// ORACLE — test infrastructure only (see oracle.h).  CPU restatement of the
// reference's StratRace2D environment and of plain (dense, one committed
// opponent model) MPPI on it, templated on the scalar: float repeats the
// engine's FP32 operation sequence, double is the FP64 oracle.
//
// Reference logic restated (pablo-gr-rey/drone-race):
//   include/environments/env_stratrace.h:53-76     sampleTrackArray, sampleNormalizedSpeed
//   include/environments/env_stratrace.h:88-162    fastProjectOnTrack
//   include/environments/env_stratrace.h:226-281   trackBoundaryDist, isWinner, isOppWinner, advance
//   include/environments/env_stratrace.h:302-315   computeSdot
//   include/environments/env_stratrace.h:318-461   computeOppTargetLatDist (N_OPP = 1)
//   include/environments/env_stratrace.h:577-706   computeOpponentAction
//   include/environments/env_stratrace.h:711-894   environmentStep (no belief update / branching)
//   include/costs.cuh:7-39                          stateCost, finalCost
//   src/kernels.cu:100-155 (dense)                  rollout loop
//   src/kernels.cu:307-408, :434-456                weighted update, L2 clamp of the nominal
//   src/mppi_controller.cu:392-436                  action, temperature adaptation, zero-fill shift
//   include/costs.cuh:44-86                         PRMPPIstateCost, PRMPPIsafetyCost, PRMPPIfinalCost
//   src/prmppi_kernels.cu:55-393                    PR rollouts, averages, minima, candidates, safety
//   src/prmppi_controller.cu:227-346                choice, action, three temperatures
// The PR safety check reuses the candidate rollouts (same environment
// noise); the reference rolls the chosen plan out again with fresh draws.
// Where the engine departs from the reference (its own exp, x*x for
// powf(x, 2), the grid index held at 511), this file departs the same way;
// the departures change no value where the reference is defined.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "oracle.h"

using namespace orc;

namespace {

constexpr int NS = DR_SR_TRACK_SAMPLES;
constexpr int WIN = 10;

inline float ex(float x) { return exp32(x); }
inline double ex(double x) { return exp(x); }
inline float flo(float x) { return floorf(x); }
inline double flo(double x) { return floor(x); }
inline float cps(float a, float b) { return copysignf(a, b); }
inline double cps(double a, double b) { return copysign(a, b); }
inline float ab(float a) { return fabsf(a); }
inline double ab(double a) { return fabs(a); }

template <class S>
struct Env
{
    const dr_sr_env& e;
    S L, dt, r, tw, kv, lat_fact;
    S max_speed[2], max_accel[2];
    explicit Env(const dr_sr_env& env) : e(env)
    {
        L = (S)env.track_length;
        dt = (S)env.dt;
        r = (S)env.drone_radius;
        tw = (S)env.track_width;
        kv = (S)env.kv;
        lat_fact = (S)env.max_opp_lat_dist_fact;
        for (int a = 0; a < 2; ++a)
        {
            max_speed[a] = (S)env.max_speed[a];
            max_accel[a] = (S)env.max_accel[a];
        }
    }
    S p(int i, int d) const { return (S)e.p_grid[i][d]; }
    S t(int i, int d) const { return (S)e.t_grid[i][d]; }
    S k(int i) const { return (S)e.kappa_grid[i]; }
};

template <class S>
struct SState
{
    S pos[4], vel[4], Sa[2], lat[2];
    int laps[2];
};

template <class S>
SState<S> load(const dr_sr_state& a)
{
    SState<S> s;
    for (int i = 0; i < 4; ++i)
    {
        s.pos[i] = (S)a.pos[i];
        s.vel[i] = (S)a.vel[i];
    }
    for (int i = 0; i < 2; ++i)
    {
        s.Sa[i] = (S)a.s[i];
        s.lat[i] = (S)a.lat_dist[i];
        s.laps[i] = a.laps[i];
    }
    return s;
}

template <class S>
void store(const SState<S>& s, dr_sr_state& a)
{
    for (int i = 0; i < 4; ++i)
    {
        a.pos[i] = (float)s.pos[i];
        a.vel[i] = (float)s.vel[i];
    }
    for (int i = 0; i < 2; ++i)
    {
        a.s[i] = (float)s.Sa[i];
        a.lat_dist[i] = (float)s.lat[i];
        a.laps[i] = s.laps[i];
    }
}

// s / L wrapped to [0, 1) times 512: cell index (held at 511) and fraction.
template <class S>
void cell(const Env<S>& E, S s, int* i0, S* t)
{
    s = s / E.L;
    s = s - flo(s);
    const S f = s * (S)NS;
    int i = 0;
    if (f >= (S)0)
        i = f < (S)NS ? (int)f : NS - 1;
    *i0 = i;
    *t = f - (S)i;
}

// sampleTrackArray<true> on the tangent grid (tang = 1) or the centre line.
template <class S>
void sample_vec(const Env<S>& E, bool tang, S s, S out[2])
{
    int i0;
    S t;
    cell(E, s, &i0, &t);
    const int i1 = (i0 + 1) % NS;
    for (int d = 0; d < 2; ++d)
    {
        const S a = tang ? E.t(i0, d) : E.p(i0, d);
        const S b = tang ? E.t(i1, d) : E.p(i1, d);
        out[d] = ((S)1 - t) * a + t * b;
    }
}

template <class S>
S sample_kappa(const Env<S>& E, S s)
{
    int i0;
    S t;
    cell(E, s, &i0, &t);
    const int i1 = (i0 + 1) % NS;
    return ((S)1 - t) * E.k(i0) + t * E.k(i1);
}

template <class S>
void normalized_speed(const Env<S>& E, S s, S out[2])
{
    sample_vec(E, true, s, out);
    const S inv = (S)1 / sq_root(out[0] * out[0] + out[1] * out[1]);
    out[0] = out[0] * inv;
    out[1] = out[1] * inv;
}

template <class S>
S project(const Env<S>& E, const S pos[2], S prevS, S* lat)
{
    prevS = prevS / E.L;
    prevS = prevS - flo(prevS);
    const S fw = prevS * (S)NS;
    const int idxPrev = (fw >= (S)0 && fw <= (S)NS) ? (int)fw : 0;
    auto sqd = [&](int i) {
        const S dx = E.p(i, 0) - pos[0], dy = E.p(i, 1) - pos[1];
        return dx * dx + dy * dy;
    };
    int bestInd = (idxPrev - WIN + NS) % NS;
    S bestSq = sqd(bestInd);
    for (int off = NS - WIN + 1; off <= NS + WIN; ++off)
    {
        const S d = sqd((idxPrev + off) % NS);
        if (d < bestSq)
        {
            bestSq = d;
            bestInd = (idxPrev + off) % NS;
        }
    }
    S segSq = (S)INFINITY, bestS = 0, cx = 0, cy = 0;
    for (int seg = -1; seg <= 0; ++seg)
    {
        const int a = (bestInd + seg + NS) % NS;
        const int b = (a + 1) % NS;
        const S abx = E.p(b, 0) - E.p(a, 0);
        const S aby = E.p(b, 1) - E.p(a, 1);
        const S ab2 = abx * abx + aby * aby;
        S t = 0;
        if ((double)ab2 > 1e-12)
        {
            t = ((pos[0] - E.p(a, 0)) * abx + (pos[1] - E.p(a, 1)) * aby) / ab2;
            t = mx((S)0, mn(t, (S)1));
        }
        const S px = E.p(a, 0) + t * abx;
        const S py = E.p(a, 1) + t * aby;
        const S d = (pos[0] - px) * (pos[0] - px) + (pos[1] - py) * (pos[1] - py);
        if (d < segSq)
        {
            segSq = d;
            bestS = ((S)a + t) / (S)NS * E.L;
            cx = px;
            cy = py;
        }
    }
    S tang[2];
    sample_vec(E, true, bestS, tang);
    const S cross = -tang[1] * (pos[0] - cx) + tang[0] * (pos[1] - cy);
    *lat = cps(sq_root(segSq), cross);
    return bestS;
}

template <class S>
S sdot(const Env<S>& E, const SState<S>& s, int i)
{
    S tang[2];
    normalized_speed(E, s.Sa[i], tang);
    const S kap = sample_kappa(E, s.Sa[i]);
    S metric = (S)1 - kap * s.lat[i];
    if (ab(metric) < (S)1e-12f)
        metric = cps((S)1e-12f, metric);
    return (tang[0] * s.vel[2 * i] + tang[1] * s.vel[2 * i + 1]) / metric;
}

// computeOppTargetLatDist(iOpp = 1): with one opponent, the agent ahead and
// behind is the ego (index 0).
template <class S>
void opp_target(const Env<S>& E, const dr_sr_opp_config& c, const SState<S>& s, S* latOut, S* goalOut)
{
    const S L = E.L;
    const S s1 = (S)c.s1, s2 = (S)c.s2, s3 = (S)c.s3;
    int iAhead = 0;
    S dA = s.Sa[0] - s.Sa[1];
    if (dA < (S)0)
        dA = dA + L;
    int iBehind = 0;
    S dB = s.Sa[0] - s.Sa[1];
    if (dB > (S)0)
        dB = dB - L;
    S latDis = 0;
    const S sdot_i = sdot(E, s, 1);
    const S baseSpeed = (S)c.speed_scale * E.max_speed[1];
    S speedGoal = baseSpeed;
    for (int k = 0; k <= 1; ++k)
    {
        const int j = k ? iAhead : iBehind;
        const S deltaS = k ? dA : dB;
        const S deltaE = s.lat[j] - s.lat[1];
        S tang_j[2];
        normalized_speed(E, s.Sa[j], tang_j);
        const S sdot_j = sdot(E, s, j);
        const S blockTol = E.max_speed[j] * (S)0.3f;
        const bool block = (deltaS < (S)0) && (deltaS > -L * (S)0.25f) && (sdot_j + blockTol > sdot_i);
        const bool overtake = (sdot_i > sdot_j) && (ab(deltaE) < s1);
        const S factor = ex(-s2 * deltaS * deltaS);
        if (block)
            latDis = latDis + s.lat[j] * ((S)1 - ex(-s3 * (sdot_j + blockTol - sdot_i))) * factor;
        if (overtake)
        {
            latDis = latDis + (s.lat[j] - cps(s1, deltaE)) * factor;
            const S oppLong = s.vel[2 * j] * tang_j[0] + s.vel[2 * j + 1] * tang_j[1];
            if (deltaS < (S)3 * E.r)
                speedGoal = mn(speedGoal, (S)0.7f * oppLong);
            else if (deltaS < (S)6 * E.r && (double)c.s1 > 1e-6)
            {
                const S latFactor = ab(deltaE) / s1;
                speedGoal = mn(speedGoal, latFactor * oppLong + ((S)1 - latFactor) * baseSpeed);
            }
        }
    }
    const S maxLat = E.tw - E.r * E.lat_fact;
    *latOut = mx(-maxLat, mn(maxLat, latDis));
    *goalOut = speedGoal;
}

template <class S>
void opp_action(const Env<S>& E, const dr_sr_opp_config& c, const SState<S>& s, S out[2])
{
    const S maxA = E.max_accel[1];
    const S dt = E.dt;
    S latRaw, speedGoal;
    opp_target(E, c, s, &latRaw, &speedGoal);
    const S sn = s.Sa[1];
    const S pos_i[2] = {s.pos[2], s.pos[3]};
    const S vel_i[2] = {s.vel[2], s.vel[3]};
    S pc[2], tang[2];
    sample_vec(E, false, sn, pc);
    normalized_speed(E, sn, tang);
    const S kap = sample_kappa(E, sn);
    const S norm[2] = {-tang[1], tang[0]};
    const S e_n = (pos_i[0] - pc[0]) * norm[0] + (pos_i[1] - pc[1]) * norm[1];
    const S v_t = vel_i[0] * tang[0] + vel_i[1] * tang[1];
    const S v_n = vel_i[0] * norm[0] + vel_i[1] * norm[1];
    const S e_n_max = E.tw - E.r * (S)1.2f;
    S metric = (S)1 - kap * e_n;
    if (ab(metric) < (S)1e-6f)
        metric = cps((S)1e-6f, metric);
    const S a_c = v_t * v_t * kap / metric;
    const S a_lat = mx(maxA - ab(a_c), maxA * (S)0.2f);
    const S latDis = mx(-e_n_max, mn(e_n_max, latRaw));
    const S err = latDis - e_n;
    const S dpos = mx(e_n_max - e_n, (S)0.001f);
    const S dneg = mx(e_n_max + e_n, (S)0.001f);
    S a_n;
    if (v_n > (S)0 && v_n * v_n > (S)1.8f * a_lat * dpos)
        a_n = -a_lat;
    else if (v_n < (S)0 && v_n * v_n > (S)1.8f * a_lat * dneg)
        a_n = a_lat;
    else if (ab(err) < (S)0.02f)
    {
        const S kp = a_lat / (S)0.02f;
        a_n = kp * err - (S)2 * sq_root(kp) * v_n;
        a_n = mx(-a_lat, mn(a_lat, a_n));
    }
    else
    {
        const S stop = v_n * v_n / ((S)2 * a_lat);
        const bool toward = (err > (S)0 && v_n > (S)0) || (err < (S)0 && v_n < (S)0);
        const bool decel = toward && stop >= ab(err) * (S)0.95f;
        a_n = (err > (S)0) ? a_lat : -a_lat;
        if (decel)
            a_n = -a_n;
        const S vnext = v_n + a_n * dt;
        if (vnext > (S)0)
        {
            const S vs = sq_root((S)2 * a_lat * dpos);
            if (vnext > vs)
                a_n = (vs - v_n) / dt;
        }
        else
        {
            const S vs = sq_root((S)2 * a_lat * dneg);
            if (-vnext > vs)
                a_n = (-vs - v_n) / dt;
        }
    }
    const S a_tot = mx(-maxA, mn(maxA, a_n + a_c));
    const S a_t_max = sq_root(mx(maxA * maxA - a_tot * a_tot, (S)0));
    const S k_eff = kap / ((S)1 - kap * latDis + (S)1e-6f);
    S path = mn(speedGoal, sq_root(maxA / mx(ab(k_eff), (S)1e-12f)));
    if (ab(a_n) > a_lat * (S)0.5f)
        path = path * mx((S)0.4f, (S)1.5f - ab(a_n) / a_lat);
    S a_t = E.kv * (path - v_t) - v_n * kap * v_t / metric;
    a_t = mx(-a_t_max, mn(a_t_max, a_t));
    out[0] = a_t * tang[0] + a_tot * norm[0];
    out[1] = a_t * tang[1] + a_tot * norm[1];
}

// Sequential normals of one (sample, step) from Philox blocks (k, 4 t + b, step, 4).
struct Normals
{
    uint32_t k0, k1, k, c1, step;
    int n = 0;
    float z[4] = {0, 0, 0, 0};
    float next()
    {
        if ((n & 3) == 0)
            normals4(k0, k1, k, c1 + (uint32_t)(n >> 2), step, 4u, z);
        const float v = z[n & 3];
        n = n + 1;
        return v;
    }
};

template <class S>
S bdist(const Env<S>& E, const SState<S>& s, bool lose)
{
    if (lose && s.laps[1] >= E.e.n_win_laps)
        return (S)-1;
    S minSq = (S)INFINITY;
    const S dx = s.pos[2] - s.pos[0], dy = s.pos[3] - s.pos[1];
    minSq = mn(minSq, dx * dx + dy * dy);
    return mn(sq_root(minSq) - E.r, E.tw - ab(s.lat[0]));
}

template <class S>
bool winner(const Env<S>& E, const SState<S>& s)
{
    if (s.laps[0] >= E.e.n_win_laps)
        return true;
    return ab(s.lat[1]) > E.tw - E.r;
}

// environmentStep<false, false, false, lose>(applyNoise = true); returns
// 0 none, 1 lose, 2 win, 3 opponent wins.
template <class S>
int env_step(const Env<S>& E, int theta, const S ego[2], SState<S>& s, Normals& rng, bool lose)
{
    const S prevS[2] = {s.Sa[0], s.Sa[1]};
    S act[4] = {ego[0], ego[1], 0, 0};
    const dr_sr_opp_config& c = E.e.opp[theta];
    S nom[2];
    opp_action(E, c, s, nom);
    for (int d = 0; d < 2; ++d)
    {
        S u = nom[d];
        if (c.action_noise != 0.0f)
            u = u + (S)c.action_noise * (S)rng.next();
        act[2 + d] = u;
    }
    for (int a = 0; a < 2; ++a)
    {
        S sq = 0;
        for (int d = 0; d < 2; ++d)
            sq = sq + act[2 * a + d] * act[2 * a + d];
        S factor = 1;
        const S maxSq = E.max_accel[a] * E.max_accel[a];
        if (sq > maxSq)
            factor = E.max_accel[a] / sq_root(sq);
        for (int d = 0; d < 2; ++d)
        {
            S v = act[2 * a + d] * factor;
            if (E.e.action_noise != 0.0f)
                v = v + (S)rng.next() * (S)E.e.action_noise;
            act[2 * a + d] = v;
        }
    }
    for (int i = 0; i < 4; ++i)
        s.pos[i] = s.pos[i] + E.dt * s.vel[i];
    for (int a = 0; a < 2; ++a)
    {
        const S sd = sdot(E, s, a);
        s.Sa[a] = s.Sa[a] + E.dt * sd;
    }
    for (int i = 0; i < 4; ++i)
        s.vel[i] = s.vel[i] + E.dt * act[i];
    for (int a = 0; a < 2; ++a)
    {
        const S sq = s.vel[2 * a] * s.vel[2 * a] + s.vel[2 * a + 1] * s.vel[2 * a + 1];
        if (sq > E.max_speed[a] * E.max_speed[a])
        {
            const S sc = E.max_speed[a] / sq_root(sq);
            s.vel[2 * a] = s.vel[2 * a] * sc;
            s.vel[2 * a + 1] = s.vel[2 * a + 1] * sc;
        }
    }
    if (E.e.pos_noise != 0.0f || E.e.speed_noise != 0.0f)
        for (int a = 0; a < 2; ++a)
            for (int d = 0; d < 2; ++d)
            {
                if (E.e.pos_noise != 0.0f)
                    s.pos[2 * a + d] = s.pos[2 * a + d] + (S)rng.next() * (S)E.e.pos_noise;
                if (E.e.speed_noise != 0.0f)
                    s.vel[2 * a + d] = s.vel[2 * a + d] + (S)rng.next() * (S)E.e.speed_noise;
            }
    for (int a = 0; a < 2; ++a)
        s.Sa[a] = project(E, &s.pos[2 * a], s.Sa[a], &s.lat[a]);
    for (int a = 0; a < 2; ++a)
    {
        if (s.Sa[a] < prevS[a] - E.L * (S)0.5f)
            s.laps[a] = s.laps[a] + 1;
        if (s.Sa[a] > prevS[a] + E.L * (S)0.5f)
            s.laps[a] = s.laps[a] - 1;
    }
    if (bdist(E, s, lose) < E.r)
        return 1;
    if (winner(E, s))
        return 2;
    if (!lose && bdist(E, s, true) < E.r)
        return 3;
    return 0;
}

template <class S>
S state_cost(const Env<S>& E, const dr_sr_cost& K, const SState<S>& s)
{
    S cost = 0;
    const S bd = bdist(E, s, false);
    const S danger = (S)K.boundary_threshold_factor * E.r;
    if (bd < danger)
    {
        const S q = bd / E.r;
        cost = cost + (S)K.boundary_cost * (q * q);
    }
    if (bd < E.r * (S)K.coll_dist_factor)
        cost = cost + (S)K.outside_cost;
    if (s.laps[1] >= E.e.n_win_laps)
        cost = cost + (S)K.opp_win_cost;
    if (winner(E, s))
        cost = cost - (S)K.win_cost;
    return cost;
}

template <class S>
S final_cost(const Env<S>& E, const dr_sr_cost& K, const SState<S>& s)
{
    S cost = -(S)K.final_adv_weight * (s.Sa[0] / E.L + (S)s.laps[0]);
    if (K.final_opp_adv_weight != 0.0f)
        cost = cost + (S)K.final_opp_adv_weight * mx(-(S)INFINITY, s.Sa[1] / E.L + (S)s.laps[1]);
    return cost;
}

// One sample: control noise sigma * (z0, z1) from the first Box-Muller pair
// of Philox (k, t, step, 0); environment noise from stream 4.
template <class S>
S rollout(const Env<S>& E, const dr_sr_cost& K, const dr_sr_state& x0, int theta, const float* nominal, int T, uint32_t key0, uint32_t key1,
          uint32_t step, float sigma, uint32_t k, float* noise_out)
{
    SState<S> s = load<S>(x0);
    S cost = 0, decay = 1;
    bool stop = false;
    for (int t = 0; t < T; ++t)
    {
        const W4 w = philox(W4{{k, (uint32_t)t, step, 0u}}, key0, key1);
        float z0, z1;
        normal_pair(w.v[0], w.v[1], &z0, &z1);
        const float e[2] = {sigma * z0, sigma * z1};
        if (noise_out)
        {
            noise_out[t * 2] = e[0];
            noise_out[t * 2 + 1] = e[1];
        }
        if (stop) // the noise of every step is drawn (the weighted average reads all of it)
            continue;
        S u[2];
        for (int d = 0; d < 2; ++d)
            u[d] = sizeof(S) == 4 ? (S)(nominal[t * 2 + d] + e[d]) : (S)nominal[t * 2 + d] + (S)e[d];
        Normals rng{key0, key1, k, 4u * (uint32_t)t, step};
        const int term = env_step(E, theta, u, s, rng, true);
        cost = fm(state_cost(E, K, s), decay, cost);
        decay = decay * (S)0.97f;
        stop = term != 0;
    }
    return cost + final_cost(E, K, s);
}

// ------------------------------------------------------------ PR-MPPI

// PRMPPIstateCost (costs.cuh:44-63): trackBoundaryDist with the opponent-win
// rule, band cost, win / opponent-win terms; no collision cost.
template <class S>
S pr_state_cost(const Env<S>& E, const dr_sr_cost& K, const SState<S>& s)
{
    S cost = 0;
    const S bd = bdist(E, s, true);
    const S danger = (S)K.boundary_threshold_factor * E.r;
    if (bd < danger)
    {
        const S q = bd / E.r;
        cost = cost + (S)K.boundary_cost * (q * q);
    }
    if (winner(E, s))
        cost = cost - (S)K.win_cost;
    if (s.laps[1] >= E.e.n_win_laps)
        cost = cost + (S)K.opp_win_cost;
    return cost;
}

// PRMPPIsafetyCost (costs.cuh:78-86): > 0 iff unsafe.
template <class S>
S pr_safety(const Env<S>& E, const dr_sr_cost& K, const SState<S>& s)
{
    return E.r * (S)K.coll_dist_factor - bdist(E, s, false);
}

// One PR rollout (prmppi_kernels.cu:98-127): plan U (T x 2, + eps when not
// NULL), environment noise from counter word kk.
template <class S, class UT>
void prsr_rollout(const Env<S>& E, const dr_sr_cost& K, const dr_sr_state& x0, int theta, const UT* U, const float* eps, int T,
                  uint32_t key0, uint32_t key1, uint32_t kk, uint32_t step, S* cost, S* safe)
{
    SState<S> s = load<S>(x0);
    S acc = 0, decay = 1, smax = -(S)INFINITY;
    for (int t = 0; t < T; ++t)
    {
        S u[2];
        for (int d = 0; d < 2; ++d)
        {
            const float e = eps ? eps[t * 2 + d] : 0.0f;
            u[d] = sizeof(S) == 4 ? (S)((float)U[t * 2 + d] + e) : (S)U[t * 2 + d] + (S)e;
        }
        Normals rng{key0, key1, kk, 4u * (uint32_t)t, step};
        const int term = env_step(E, theta, u, s, rng, true);
        acc = fm(pr_state_cost(E, K, s), decay, acc);
        smax = mx(smax, pr_safety(E, K, s));
        decay = decay * (S)0.97f;
        if (term)
            break;
    }
    *cost = acc + final_cost(E, K, s);
    *safe = smax;
}

int prsr_models(const dr_prsr_instance& I, const dr_sr_env& env) { return I.sampled ? I.n_draws : env.n_models; }

template <class S>
S prsr_avg(const dr_prsr_instance& I, const float* belief, const S* c, int M)
{
    S acc = 0;
    if (I.sampled)
    {
        for (int m = 0; m < M; ++m)
            acc = acc + c[m];
        return acc / (S)M;
    }
    for (int m = 0; m < M; ++m)
        acc = fm((S)belief[m], c[m], acc);
    return acc;
}

// Environment-noise counter word of rollout (model m, sample k) and of
// candidate lane l.
inline uint32_t sample_stream(int m, int K, int k) { return (uint32_t)m * (uint32_t)K + (uint32_t)k; }
inline uint32_t cand_stream(int lane) { return 0xF0000000u + (uint32_t)lane; }

template <class S>
void prsr_step(const dr_sr_env& env, const dr_sr_cost& K, const dr_sr_state& x0, const dr_prsr_instance& I, const float* belief,
               uint32_t step, const float lam[3], const float* nom, const float* rob, int n_threads, float* noise_out, double* pair_out,
               double* agg_out, int* models_out, double* cand_out, double* res_out)
{
    const int N = I.n_samples, T = I.n_timesteps, M = prsr_models(I, env);
    const uint32_t key0 = (uint32_t)(I.seed & 0xffffffffu), key1 = (uint32_t)(I.seed >> 32);
    const Env<S> E(env);
    std::vector<int> th(M);
    for (int m = 0; m < M; ++m)
        th[m] = I.sampled ? sample_branch(key0, key1, (uint32_t)m, step, belief, env.n_models) : m;
    if (models_out)
        for (int m = 0; m < M; ++m)
            models_out[m] = th[m];
    std::vector<float> eps_all((size_t)N * T * 2);
    std::vector<S> fn(N), fr(N), sr(N);
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads > 0 ? n_threads : 1)
    for (int k = 0; k < N; ++k)
    {
        float* eps = &eps_all[(size_t)k * T * 2];
        for (int t = 0; t < T; ++t)
        {
            const W4 w = philox(W4{{(uint32_t)k, (uint32_t)t, step, 0u}}, key0, key1);
            float z0, z1;
            normal_pair(w.v[0], w.v[1], &z0, &z1);
            eps[t * 2] = I.sigma * z0;
            eps[t * 2 + 1] = I.sigma * z1;
        }
        S c[2][DR_MAX_MODELS], sf[2][DR_MAX_MODELS];
        for (int plan = 0; plan < 2; ++plan)
            for (int m = 0; m < M; ++m)
            {
                prsr_rollout<S, float>(E, K, x0, th[m], plan ? rob : nom, eps, T, key0, key1, sample_stream(m, N, k), step, &c[plan][m],
                                       &sf[plan][m]);
                if (pair_out)
                {
                    double* o = pair_out + (((size_t)plan * M + m) * N + k) * 2;
                    o[0] = (double)c[plan][m];
                    o[1] = (double)sf[plan][m];
                }
            }
        S full[2], smax[2];
        for (int plan = 0; plan < 2; ++plan)
        {
            const S avg = prsr_avg<S>(I, belief, c[plan], M);
            S mxs = -(S)INFINITY;
            for (int m = 0; m < M; ++m)
                mxs = mx(mxs, sf[plan][m]);
            full[plan] = mxs > (S)0 ? avg + (S)I.safety_weight : avg;
            smax[plan] = mxs;
        }
        fn[k] = full[0];
        fr[k] = full[1];
        sr[k] = smax[1];
        if (agg_out)
        {
            agg_out[(size_t)k * 4 + 0] = (double)full[0];
            agg_out[(size_t)k * 4 + 1] = (double)smax[0];
            agg_out[(size_t)k * 4 + 2] = (double)full[1];
            agg_out[(size_t)k * 4 + 3] = (double)smax[1];
        }
    }
    if (noise_out)
        memcpy(noise_out, eps_all.data(), eps_all.size() * sizeof(float));
    const std::vector<S>* cs[3] = {&fn, &fr, &sr};
    for (int j = 0; j < 3; ++j)
    {
        const std::vector<S>& c = *cs[j];
        S mnc = (S)INFINITY;
        for (int k = 0; k < N; ++k)
            mnc = mn(mnc, c[k]);
        const float* base = j == 0 ? nom : rob;
        std::vector<double> num((size_t)T * 2, 0.0);
        double den = 0.0;
        for (int k = 0; k < N; ++k)
        {
            double w;
            if (sizeof(S) == 4)
                w = isfinite((double)mnc) ? (double)exp32(-((float)c[k] - (float)mnc) / lam[j]) : 0.0;
            else
                w = isfinite((double)mnc) ? exp(-((double)c[k] - (double)mnc) / (double)lam[j]) : 0.0;
            const float* eps = &eps_all[(size_t)k * T * 2];
            for (int i = 0; i < T * 2; ++i)
                num[i] += w * (sizeof(S) == 4 ? (double)(base[i] + eps[i]) : (double)base[i] + (double)eps[i]);
            den += w;
        }
        for (int i = 0; i < T * 2; ++i)
        {
            double v = (double)base[i];
            if (den > 1e-30)
                v = sizeof(S) == 4 ? (double)(float)(num[i] / den) : num[i] / den;
            cand_out[(size_t)j * T * 2 + i] = v;
        }
        res_out[j] = (double)mnc;
        res_out[3 + j] = sizeof(S) == 4 ? (double)(float)den : den;
    }
}

template <class S>
void prsr_choose(const dr_sr_env& env, const dr_sr_cost& K, const dr_sr_state& x0, const dr_prsr_instance& I, const float* belief,
                 const int* models, const double* cand, uint32_t step, const float lam[3], const double nu[3], double* cc_out, double* cs_out,
                 double* res_out)
{
    const int T = I.n_timesteps, M = prsr_models(I, env);
    const uint32_t key0 = (uint32_t)(I.seed & 0xffffffffu), key1 = (uint32_t)(I.seed >> 32);
    const Env<S> E(env);
    std::vector<S> U((size_t)3 * T * 2);
    for (size_t i = 0; i < U.size(); ++i)
        U[i] = sizeof(S) == 4 ? (S)(float)cand[i] : (S)cand[i];
    S cc[2][DR_MAX_MODELS], sf[2][DR_MAX_MODELS];
    for (int which = 0; which < 2; ++which)
        for (int m = 0; m < M; ++m)
        {
            S c;
            prsr_rollout<S, S>(E, K, x0, models[m], &U[(size_t)which * T * 2], nullptr, T, key0, key1, cand_stream(which * M + m), step, &c,
                               &sf[which][m]);
            cc[which][m] = sf[which][m] > (S)0 ? c + (S)I.safety_weight : c;
            cc_out[which * M + m] = (double)cc[which][m];
            cs_out[which * M + m] = (double)sf[which][m];
        }
    const S c1 = prsr_avg<S>(I, belief, cc[0], M);
    const S c2 = prsr_avg<S>(I, belief, cc[1], M);
    const int pick = c1 <= c2 ? 0 : 1;
    int use_nom = 1;
    for (int m = 0; m < M; ++m)
        if (sf[pick][m] > (S)0)
            use_nom = 0;
    res_out[0] = (double)c1;
    res_out[1] = (double)c2;
    res_out[2] = use_nom;
    res_out[3] = c2 < c1 ? 1 : 0;
    const S* act = use_nom ? &U[(size_t)pick * T * 2] : &U[(size_t)2 * T * 2];
    res_out[4] = (double)act[0];
    res_out[5] = (double)act[1];
    for (int j = 0; j < 3; ++j)
    {
        float l = lam[j];
        if (I.adapt_lambda)
        {
            const float v = (float)nu[j];
            if (v > 0.001f * (float)I.n_samples && l > 0.01f)
                l = l * 0.9f;
            else if (v < 0.0005f * (float)I.n_samples && l < 100.0f)
                l = l * 1.2f;
        }
        res_out[6 + j] = (double)l;
    }
}

} // namespace

extern "C" {

// One environment step of a state (the environment's noise from the stream
// of sample k, step t, control step `step`).  Returns the terminal code.
int or_sr_env_step(int f64, const dr_sr_env* env, int theta, const float* ego, dr_sr_state* st, uint64_t seed, uint32_t k, uint32_t t,
                   uint32_t step, int lose_on_opp_win)
{
    Normals rng{(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32), k, 4u * t, step};
    if (f64)
    {
        const Env<double> E(*env);
        SState<double> s = load<double>(*st);
        const double u[2] = {ego[0], ego[1]};
        const int r = env_step(E, theta, u, s, rng, lose_on_opp_win != 0);
        store(s, *st);
        return r;
    }
    const Env<float> E(*env);
    SState<float> s = load<float>(*st);
    const float u[2] = {ego[0], ego[1]};
    const int r = env_step(E, theta, u, s, rng, lose_on_opp_win != 0);
    store(s, *st);
    return r;
}

double or_sr_project(int f64, const dr_sr_env* env, double x, double y, double prevS, double* lat)
{
    if (f64)
    {
        const Env<double> E(*env);
        const double p[2] = {x, y};
        return project(E, p, prevS, lat);
    }
    const Env<float> E(*env);
    const float p[2] = {(float)x, (float)y};
    float l;
    const float s = project(E, p, (float)prevS, &l);
    *lat = l;
    return s;
}

void or_sr_opp_action(int f64, const dr_sr_env* env, int theta, const dr_sr_state* st, double* out)
{
    if (f64)
    {
        const Env<double> E(*env);
        opp_action(E, env->opp[theta], load<double>(*st), out);
        return;
    }
    const Env<float> E(*env);
    float o[2];
    opp_action(E, env->opp[theta], load<float>(*st), o);
    out[0] = o[0];
    out[1] = o[1];
}

double or_sr_state_cost(int f64, const dr_sr_env* env, const dr_sr_cost* K, const dr_sr_state* st)
{
    if (f64)
        return state_cost(Env<double>(*env), *K, load<double>(*st));
    return state_cost(Env<float>(*env), *K, load<float>(*st));
}

// K sampled rollouts: costs[K], noise [k][t][2] (may be NULL).
int or_sr_rollout(int f64, const dr_sr_env* env, const dr_sr_cost* K, const dr_sr_state* x0, int theta, const float* nominal, int n_samples,
                  int T, uint64_t seed, uint32_t step, float sigma, int n_threads, double* costs, float* noise_out)
{
    if (!env || !K || !x0 || !nominal || !costs || theta < 0 || theta >= env->n_models)
        return 1;
    const uint32_t key0 = (uint32_t)(seed & 0xffffffffu), key1 = (uint32_t)(seed >> 32);
    const Env<float> Ef(*env);
    const Env<double> Ed(*env);
#pragma omp parallel for schedule(dynamic, 64) num_threads(n_threads > 0 ? n_threads : 1)
    for (int k = 0; k < n_samples; ++k)
    {
        float* no = noise_out ? noise_out + (size_t)k * T * 2 : nullptr;
        costs[k] = f64 ? rollout<double>(Ed, *K, *x0, theta, nominal, T, key0, key1, step, sigma, (uint32_t)k, no)
                       : (double)rollout<float>(Ef, *K, *x0, theta, nominal, T, key0, key1, step, sigma, (uint32_t)k, no);
    }
    return 0;
}

// Dense MPPI update of kernels.cu:376-403 (FP32 weights, FP64 sums, FP32
// increment; all FP64 when f64), L2 clamp of every step to max_accel
// (kernels.cu:434-456), action, shift with fill.
int or_sr_update(int f64, const double* costs, const float* noise, const float* nominal, int n_samples, int T, float lambda, float max_accel,
                 const float* fill, double* updated, double* action, double* shifted, double* min_cost, double* nu)
{
    double mnc = INFINITY;
    for (int k = 0; k < n_samples; ++k)
        mnc = costs[k] < mnc ? costs[k] : mnc;
    std::vector<double> num((size_t)T * 2, 0.0);
    double den = 0.0;
    for (int k = 0; k < n_samples; ++k)
    {
        const double w = f64 ? exp(-(costs[k] - mnc) / (double)lambda) : (double)exp32(-((float)costs[k] - (float)mnc) / lambda);
        for (int i = 0; i < T * 2; ++i)
            num[i] += w * (double)noise[(size_t)k * T * 2 + i];
        den += w;
    }
    for (int t = 0; t < T; ++t)
    {
        double u[2];
        for (int d = 0; d < 2; ++d)
        {
            const int i = t * 2 + d;
            if (f64)
                u[d] = den > 1e-30 ? (double)nominal[i] + num[i] / den : (double)nominal[i];
            else
                u[d] = den > 1e-30 ? (double)(nominal[i] + (float)(num[i] / den)) : (double)nominal[i];
        }
        if (f64)
        {
            const double sq = u[0] * u[0] + u[1] * u[1];
            if (sq > (double)max_accel * max_accel)
            {
                const double sc = max_accel / sqrt(sq);
                u[0] *= sc;
                u[1] *= sc;
            }
        }
        else
        {
            const float a = (float)u[0], b = (float)u[1];
            const float sq = a * a + b * b;
            float ua = a, ub = b;
            if (sq > max_accel * max_accel)
            {
                const float sc = max_accel / sqrtf(sq);
                ua = a * sc;
                ub = b * sc;
            }
            u[0] = ua;
            u[1] = ub;
        }
        updated[t * 2] = u[0];
        updated[t * 2 + 1] = u[1];
    }
    action[0] = updated[0];
    action[1] = updated[1];
    for (int t = 0; t < T; ++t)
        for (int d = 0; d < 2; ++d)
            shifted[t * 2 + d] = (t + 1 < T) ? updated[(t + 1) * 2 + d] : (double)fill[d];
    *min_cost = mnc;
    *nu = f64 ? den : (double)(float)den;
    return 0;
}

// PR-MPPI step of one StratRace2D controller from the shifted plans nom, rob
// (T x 2): noise [k][t][2], pair costs [plan][m][k][2], aggregates [k][4],
// models [M], candidates [3][T][2], res = min[3], nu[3].
int or_prsr_step(int f64, const dr_sr_env* env, const dr_sr_cost* K, const dr_sr_state* x0, const dr_prsr_instance* inst, const float* belief,
                 uint32_t step, const float* lam, const float* nom, const float* rob, int n_threads, float* noise_out, double* pair_out,
                 double* agg_out, int* models_out, double* cand_out, double* res_out)
{
    if (!env || !K || !x0 || !inst || !belief || !lam || !nom || !rob || !cand_out || !res_out)
        return 1;
    const int M = prsr_models(*inst, *env);
    if (M < 1 || M > DR_MAX_MODELS)
        return 1;
    if (f64)
        prsr_step<double>(*env, *K, *x0, *inst, belief, step, lam, nom, rob, n_threads, noise_out, pair_out, agg_out, models_out, cand_out,
                          res_out);
    else
        prsr_step<float>(*env, *K, *x0, *inst, belief, step, lam, nom, rob, n_threads, noise_out, pair_out, agg_out, models_out, cand_out,
                         res_out);
    return 0;
}

// Candidate costs and choice for candidates [3][T][2]: cc/cs [which][m];
// res = cand cost 1, 2, use_nom, reset_nom, action[2], lambda[3].
int or_prsr_choose(int f64, const dr_sr_env* env, const dr_sr_cost* K, const dr_sr_state* x0, const dr_prsr_instance* inst, const float* belief,
                   const int* models, const double* cand, uint32_t step, const float* lam, const double* nu, double* cc_out, double* cs_out,
                   double* res_out)
{
    if (!env || !K || !x0 || !inst || !belief || !models || !cand || !lam || !nu || !cc_out || !cs_out || !res_out)
        return 1;
    if (f64)
        prsr_choose<double>(*env, *K, *x0, *inst, belief, models, cand, step, lam, nu, cc_out, cs_out, res_out);
    else
        prsr_choose<float>(*env, *K, *x0, *inst, belief, models, cand, step, lam, nu, cc_out, cs_out, res_out);
    return 0;
}

} // extern "C"
