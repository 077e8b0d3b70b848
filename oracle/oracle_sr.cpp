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
//   include/environments/env_stratrace.h:166-222   updateBelief
//   include/environments/env_stratrace.h:711-894   environmentStep (belief update; branching in the caller)
//   include/state.h:93-115, :154-228               findConfident, initBranchState, sampleContributed,
//                                                   splineSampleContributed
//   src/mppi_controller.cu:67-141                   buildSplineMatrix
//   src/kernels.cu:34-486, src/mppi_controller.cu:250-466
//                                                   branching spline MPPI step (no plan verification)
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
#ifdef _OPENMP
#include <omp.h>
#endif

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
// the op-counting scalar: exp, log and floor count one each; copysign and
// fabs are sign-bit operations and count none
inline Cn ex(Cn x) { ++g_tally.flops; return Cn(exp(x.v)); }
inline Cn flo(Cn x) { ++g_tally.flops; return Cn(floor(x.v)); }
inline Cn cps(Cn a, Cn b) { return Cn(copysign(a.v, b.v)); }
inline Cn ab(Cn a) { return Cn(fabs(a.v)); }

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

// Sequential normals of one (sample, step) from Philox blocks (k, 4 t + b, step, stream).
struct Normals
{
    uint32_t k0, k1, k, c1, step;
    uint32_t stream = 4u; // 5: plan verification
    int n = 0;
    float z[4] = {0, 0, 0, 0};
    float next()
    {
        if ((n & 3) == 0)
            normals4(k0, k1, k, c1 + (uint32_t)(n >> 2), step, stream, z);
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

inline float lg(float x) { return log32(x); }
inline double lg(double x) { return log(x); }
inline Cn lg(Cn x) { ++g_tally.flops; return Cn(log(x.v)); }

// updateBelief (env_stratrace.h:166-222), one opponent: Gaussian likelihood
// of the applied opponent action `obs` under each model's nominal action
// (scaled to the opponent's max acceleration), variance 0.2^2 + noise^2;
// product, then normalisation (uniform when the sum is below 1e-20).
template <class S>
void update_belief(const Env<S>& E, S* belief, const S obs[2], const S (*nom)[2])
{
    const S sigmaEnv = (S)0.2f;
    const S log2pi = lg((S)2.0f * (S)3.14159265358979323846f);
    S sum = 0;
    const int NT = E.e.n_models;
    for (int th = 0; th < NT; ++th)
    {
        const S sqA = nom[th][0] * nom[th][0] + nom[th][1] * nom[th][1];
        S scale = 1;
        if (sqA > E.max_accel[1] * E.max_accel[1])
            scale = E.max_accel[1] / sq_root(sqA);
        const S dx = obs[0] - nom[th][0] * scale;
        const S dy = obs[1] - nom[th][1] * scale;
        const S sqDist = dx * dx + dy * dy;
        const S an = (S)E.e.opp[th].action_noise;
        const S sqSigma = sigmaEnv * sigmaEnv + an * an;
        const S ll = (S)-0.5f * sqDist / sqSigma - log2pi - lg(sqSigma);
        belief[th] = belief[th] * ex(ll);
        sum = sum + belief[th];
    }
    if (sum < (S)1e-20f)
    {
        const S uniform = (S)1 / (S)NT;
        for (int th = 0; th < NT; ++th)
            belief[th] = uniform;
    }
    else
        for (int th = 0; th < NT; ++th)
            belief[th] = belief[th] / sum;
}

// findConfident (state.h:93-115), one model factor: theta + 1 of the first
// model whose belief exceeds the threshold, else 0.
template <class S>
int find_confident(const S* belief, int NT, float threshold)
{
    for (int th = 0; th < NT; ++th)
        if (belief[th] > (S)threshold)
            return th + 1;
    return 0;
}

// environmentStep<belief != NULL, false, false, lose>(applyNoise = noise);
// returns 0 none, 1 lose, 2 win, 3 opponent wins.  With a belief, it is
// updated from the applied opponent action (env_stratrace.h:856-859).
template <class S>
int env_step(const Env<S>& E, int theta, const S ego[2], SState<S>& s, Normals& rng, bool lose, S* belief = nullptr, bool noise = true)
{
    const S prevS[2] = {s.Sa[0], s.Sa[1]};
    S act[4] = {ego[0], ego[1], 0, 0};
    const dr_sr_opp_config& c = E.e.opp[theta];
    S noms[DR_SR_MAX_MODELS][2];
    S nom[2];
    if (belief)
    {
        for (int th = 0; th < E.e.n_models; ++th)
            opp_action(E, E.e.opp[th], s, noms[th]);
        nom[0] = noms[theta][0];
        nom[1] = noms[theta][1];
    }
    else
        opp_action(E, c, s, nom);
    for (int d = 0; d < 2; ++d)
    {
        S u = nom[d];
        if (noise && c.action_noise != 0.0f)
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
            if (noise && E.e.action_noise != 0.0f)
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
    if (noise && (E.e.pos_noise != 0.0f || E.e.speed_noise != 0.0f))
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
    if (belief)
    {
        const S obs[2] = {act[2], act[3]};
        update_belief(E, belief, obs, noms);
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


// ------------------------------------------------------------ branching spline MPPI

// ------------------------------------------------------------ HiddenObs
// env_hiddenobs.h restated on the scalar S (every constant rounded to float
// first, as the engine holds it).

template <class S>
struct HoEnv
{
    const dr_ho_env& e;
    explicit HoEnv(const dr_ho_env& env) : e(env) {}
    S f(float v) const { return (S)v; }
};

template <class S>
struct HState
{
    S pos[2], vel[2];
    int laps, gates;
};

template <class S>
HState<S> ho_load(const dr_ho_state& a)
{
    HState<S> s;
    for (int d = 0; d < 2; ++d)
    {
        s.pos[d] = (S)a.pos[d];
        s.vel[d] = (S)a.vel[d];
    }
    s.laps = a.laps;
    s.gates = a.gates;
    return s;
}

template <class S>
void ho_store(const HState<S>& s, dr_ho_state& a)
{
    for (int d = 0; d < 2; ++d)
    {
        a.pos[d] = (float)s.pos[d];
        a.vel[d] = (float)s.vel[d];
    }
    a.laps = s.laps;
    a.gates = s.gates;
}

// sqDistToRectangle (env_hiddenobs.h:27-34): clamp the point into the box
template <class S>
S ho_sq_rect(S x, S y, S x0, S y0, S x1, S y1)
{
    const S dx = x - mn(mx(x, x0), x1);
    const S dy = y - mn(mx(y, y0), y1);
    return dx * dx + dy * dy;
}

// sqDistToHalfPlane (env_hiddenobs.h:36-60): region x0 <= x <= x1, x + y >= lam;
// outside it, the nearest of the diagonal edge and the two vertical rays
template <class S>
S ho_sq_hp(S x, S y, S x0, S x1, S lam)
{
    if (x0 <= x && x <= x1 && x + y >= lam)
        return (S)0;
    const S px = mx(mn((x - y + lam) * (S)0.5f, x1), x0);
    const S py = lam - px;
    S best = (x - px) * (x - px) + (y - py) * (y - py);
    const S yl = mx(y, lam - x0);
    best = mn(best, (x - x0) * (x - x0) + (y - yl) * (y - yl));
    const S yr = mx(y, lam - x1);
    best = mn(best, (x - x1) * (x - x1) + (y - yr) * (y - yr));
    return best;
}

// distToAnnulus (env_hiddenobs.h:62-86): top-left quadrant of the centre,
// obstacle where r <= rmin or r >= rmax; +inf outside the quadrant
template <class S>
S ho_annulus(S x, S y, S cx, S cy, S r0, S r1)
{
    const S ux = x - cx, uy = y - cy;
    const S rr = ux * ux + uy * uy;
    const bool quad = ux <= (S)0 && uy >= (S)0;
    if (quad && (rr <= r0 * r0 || rr >= r1 * r1))
        return (S)0;
    if (quad)
    {
        const S r = sq_root(rr);
        return mn(r - r0, r1 - r);
    }
    return (S)INFINITY;
}

// trackBoundaryDist (env_hiddenobs.h:93-132)
template <class S>
S ho_bdist(const HoEnv<S>& H, const HState<S>& s, int theta)
{
    const dr_ho_env& e = H.e;
    const S x = s.pos[0], y = s.pos[1];
    S sq = (S)INFINITY;
    for (int r = 0; r < 3; ++r)
        sq = mn(sq, ho_sq_rect(x, y, H.f(e.rect_obstacles[r][0]), H.f(e.rect_obstacles[r][1]), H.f(e.rect_obstacles[r][2]),
                               H.f(e.rect_obstacles[r][3])));
    const S a0 = H.f(e.dbl_hp_limits[0]), a1 = H.f(e.dbl_hp_limits[1]), b0 = H.f(e.dbl_hp_limits[2]), b1 = H.f(e.dbl_hp_limits[3]);
    // first double half-plane mirrored in x, second as is; the bottom halves mirrored in y
    const S h1 = mn(ho_sq_hp(-x, y, -a1, -a0, H.f(e.dbl_hp_lambdas[0])), ho_sq_hp(-x, -y, -a1, -a0, H.f(e.dbl_hp_lambdas[1])));
    const S h2 = mn(ho_sq_hp(x, y, b0, b1, H.f(e.dbl_hp_lambdas[2])), ho_sq_hp(x, -y, b0, b1, H.f(e.dbl_hp_lambdas[3])));
    sq = mn(sq, mn(h1, h2));
    S d = sq_root(sq);
    const S walls = mn(mn(x - H.f(e.arena_min[0]), H.f(e.arena_max[0]) - x), mn(y - H.f(e.arena_min[1]), H.f(e.arena_max[1]) - y));
    d = mn(d, walls);
    d = mn(d, ho_annulus(x, y, H.f(e.annulus_center[0]), H.f(e.annulus_center[1]), H.f(e.annulus_radius[0]), H.f(e.annulus_radius[1])));
    const S hx = x - H.f(e.hidden_centers[theta][0]), hy = y - H.f(e.hidden_centers[theta][1]);
    d = mn(d, sq_root(hx * hx + hy * hy) - H.f(e.hidden_radius[theta]));
    return d;
}

// segmentIntersectsRectangle (env_hiddenobs.h:147-201): Liang-Barsky
// clipping of the parameter interval [0, 1] against both slabs
template <class S>
bool ho_seg_rect(S px, S py, S qx, S qy, S x0, S y0, S x1, S y1)
{
    S lo = 0, hi = 1;
    const S p[2] = {px, py}, dq[2] = {qx - px, qy - py}, bmin[2] = {x0, y0}, bmax[2] = {x1, y1};
    for (int a = 0; a < 2; ++a)
    {
        if (ab(dq[a]) < (S)1e-12f)
        {
            if (p[a] < bmin[a] || p[a] > bmax[a])
                return false;
            continue;
        }
        S ta = (bmin[a] - p[a]) / dq[a], tb = (bmax[a] - p[a]) / dq[a];
        if (ta > tb)
        {
            const S t = ta;
            ta = tb;
            tb = t;
        }
        lo = mx(lo, ta);
        hi = mn(hi, tb);
        if (lo > hi)
            return false;
    }
    return true;
}

// segmentIntersectsAnnulus (env_hiddenobs.h:203-276): clip the segment to
// the quadrant, then its nearest point against rmin and its ends against
// rmax (|x(t)|^2 is convex in t)
template <class S>
bool ho_seg_annulus(S px, S py, S qx, S qy, S cx, S cy, S r0, S r1)
{
    px = px - cx;
    py = py - cy;
    qx = qx - cx;
    qy = qy - cy;
    S lo = 0, hi = 1;
    if (ab(px - qx) > (S)1e-12f)
    {
        const S c = px / (px - qx);
        if (qx > px)
            hi = mn(hi, c);
        else
            lo = mx(lo, c);
    }
    else if (px > (S)0)
        return false;
    if (hi < lo)
        return false;
    if (ab(py - qy) > (S)1e-12f)
    {
        const S c = py / (py - qy);
        if (qy > py)
            lo = mx(lo, c);
        else
            hi = mn(hi, c);
    }
    else if (py > (S)0)
        return false;
    if (hi < lo)
        return false;
    const S ddx = px - qx, ddy = py - qy;
    const S len2 = ddx * ddx + ddy * ddy;
    if (len2 > (S)1e-12f)
    {
        const S tc = mx(lo, mn(hi, (px * ddx + py * ddy) / len2));
        const S xm = px + tc * (qx - px), ym = py + tc * (qy - py);
        if (xm * xm + ym * ym < r0 * r0)
            return true;
    }
    for (int end = 0; end < 2; ++end)
    {
        const S t = end == 0 ? lo : hi;
        const S xe = px + t * (qx - px), ye = py + t * (qy - py);
        if (xe * xe + ye * ye > r1 * r1)
            return true;
    }
    return false;
}

// canSeeHiddenObs (env_hiddenobs.h:278-295)
template <class S>
bool ho_can_see(const HoEnv<S>& H, const HState<S>& s, int id)
{
    const dr_ho_env& e = H.e;
    const S qx = H.f(e.hidden_centers[id][0]), qy = H.f(e.hidden_centers[id][1]);
    for (int r = 0; r < 3; ++r)
        if (ho_seg_rect(s.pos[0], s.pos[1], qx, qy, H.f(e.rect_obstacles[r][0]), H.f(e.rect_obstacles[r][1]), H.f(e.rect_obstacles[r][2]),
                        H.f(e.rect_obstacles[r][3])))
            return false;
    return !ho_seg_annulus(s.pos[0], s.pos[1], qx, qy, H.f(e.annulus_center[0]), H.f(e.annulus_center[1]), H.f(e.annulus_radius[0]),
                           H.f(e.annulus_radius[1]));
}

// getAdvance (env_hiddenobs.h:297-316)
template <class S>
S ho_advance(const HoEnv<S>& H, const HState<S>& s)
{
    const dr_ho_env& e = H.e;
    const int nx = (s.gates + 1) % 3;
    S g2 = 0, c2 = 0;
    for (int d = 0; d < 2; ++d)
    {
        const S a = H.f(e.gate_centers[nx][d]) - s.pos[d];
        const S b = H.f(e.gate_centers[nx][d]) - H.f(e.gate_centers[s.gates][d]);
        g2 = g2 + a * a;
        c2 = c2 + b * b;
    }
    return (S)(float)(s.gates + 3 * s.laps) + (S)0.8f * ((S)1 - sq_root(g2 / c2));
}

// updateGates<false> (env_hiddenobs.h:330-376)
template <class S>
void ho_gates(const HoEnv<S>& H, HState<S>& s, const S prev[2])
{
    const dr_ho_env& e = H.e;
    const int nx = (s.gates + 1) % 3;
    S num = 0, den = 0;
    for (int d = 0; d < 2; ++d)
    {
        num = num + H.f(e.gate_vectors[nx][d]) * (H.f(e.gate_centers[nx][d]) - prev[d]);
        den = den + H.f(e.gate_vectors[nx][d]) * (s.pos[d] - prev[d]);
    }
    if (ab(den) < (S)1e-10f)
        return;
    const S lam = num / den;
    if (lam < (S)0 || lam > (S)1)
        return;
    S sq = 0;
    for (int d = 0; d < 2; ++d)
    {
        const S v = ((S)1 - lam) * prev[d] + lam * s.pos[d] - H.f(e.gate_centers[nx][d]);
        sq = sq + v * v;
    }
    if (sq <= H.f(e.gate_radius[nx]) * H.f(e.gate_radius[nx]))
    {
        s.gates = s.gates + 1;
        if (s.gates == 3)
        {
            s.gates = 0;
            s.laps = s.laps + 1;
        }
    }
}

// environmentStep<belief != NULL, false, false> (env_hiddenobs.h:378-483):
// returns 0 none, 1 lose, 2 win.  With a belief, it collapses on theta when
// either hidden obstacle is in sight while no model has 0.9.
template <class S>
int env_step(const HoEnv<S>& H, int theta, const S ego[2], HState<S>& s, Normals& rng, bool /*lose*/, S* belief = nullptr, bool noise = true)
{
    const dr_ho_env& e = H.e;
    const S prev[2] = {s.pos[0], s.pos[1]};
    const S ma = H.f(e.max_accel), vmax = H.f(e.max_speed), dt = H.f(e.dt);
    S u[2] = {ego[0], ego[1]};
    const S n2 = u[0] * u[0] + u[1] * u[1];
    const S scale = n2 > ma * ma ? ma / sq_root(n2) : (S)1;
    for (int d = 0; d < 2; ++d)
    {
        u[d] = u[d] * scale;
        if (noise && e.action_noise != 0.0f)
            u[d] = u[d] + (S)rng.next() * H.f(e.action_noise);
    }
    for (int d = 0; d < 2; ++d)
        s.pos[d] = s.pos[d] + dt * s.vel[d];
    for (int d = 0; d < 2; ++d)
        s.vel[d] = s.vel[d] + dt * u[d];
    const S v2 = s.vel[0] * s.vel[0] + s.vel[1] * s.vel[1];
    if (v2 > vmax * vmax)
    {
        const S c = vmax / sq_root(v2);
        for (int d = 0; d < 2; ++d)
            s.vel[d] = s.vel[d] * c;
    }
    if (noise && (e.pos_noise != 0.0f || e.speed_noise != 0.0f))
        for (int d = 0; d < 2; ++d)
        {
            if (e.pos_noise != 0.0f)
                s.pos[d] = s.pos[d] + (S)rng.next() * H.f(e.pos_noise);
            if (e.speed_noise != 0.0f)
                s.vel[d] = s.vel[d] + (S)rng.next() * H.f(e.speed_noise);
        }
    ho_gates(H, s, prev);
    if (belief && belief[0] < (S)0.9f && belief[1] < (S)0.9f && (ho_can_see(H, s, theta) || ho_can_see(H, s, 1 - theta)))
    {
        belief[theta] = (S)1;
        belief[1 - theta] = (S)0;
    }
    if (ho_bdist(H, s, theta) < H.f(e.drone_radius))
        return 1;
    if (s.laps >= e.n_win_laps)
        return 2;
    return 0;
}

// stateCost / finalCost (costs.cuh:7-39) on HiddenObs (no opponent)
template <class S>
S stage_cost(const HoEnv<S>& H, const dr_sr_cost& K, const HState<S>& s, int theta)
{
    const S r = H.f(H.e.drone_radius);
    const S bd = ho_bdist(H, s, theta);
    S cost = 0;
    if (bd < (S)K.boundary_threshold_factor * r)
    {
        const S q = bd / r;
        cost = cost + (S)K.boundary_cost * (q * q);
    }
    if (bd < r * (S)K.coll_dist_factor)
        cost = cost + (S)K.outside_cost;
    if (s.laps >= H.e.n_win_laps)
        cost = cost - (S)K.win_cost;
    return cost;
}

template <class S>
S final_cost(const HoEnv<S>& H, const dr_sr_cost& K, const HState<S>& s)
{
    S cost = -(S)K.final_adv_weight * ho_advance(H, s);
    if (K.final_opp_adv_weight != 0.0f)
        cost = cost + (S)K.final_opp_adv_weight * (S)0;
    return cost;
}

// The environment-specific pieces the branching MPPI step needs, for both
// environments.
template <class S>
S stage_cost(const Env<S>& E, const dr_sr_cost& K, const SState<S>& s, int)
{
    return state_cost(E, K, s);
}
template <class S>
int n_models(const Env<S>& E)
{
    return E.e.n_models;
}
template <class S>
int n_models(const HoEnv<S>&)
{
    return 2;
}
template <class S>
SState<S> load_st(const Env<S>&, const dr_sr_state& a)
{
    return load<S>(a);
}
template <class S>
HState<S> load_st(const HoEnv<S>&, const dr_ho_state& a)
{
    return ho_load<S>(a);
}
inline float max_accel_of(const dr_sr_env& e) { return e.max_accel[0]; }
inline float max_accel_of(const dr_ho_env& e) { return e.max_accel; }
template <class S>
Env<S> make_env(const dr_sr_env& e)
{
    return Env<S>(e);
}
template <class S>
HoEnv<S> make_env(const dr_ho_env& e)
{
    return HoEnv<S>(e);
}

// ------------------------------------------------------------ DroneRace
// env_dronerace.h restated on the scalar S (configuration values rounded to
// float first, as the engine holds them).

template <class S>
struct DrEnv
{
    const dr_drace_env& e;
    explicit DrEnv(const dr_drace_env& env) : e(env) {}
    S f(float v) const { return (S)v; }
    const float* line(int r) const { return &e.track_points[r][0][0]; }
};

template <class S>
struct DState
{
    S pos[4], vel[4], sp[2][DR_DRACE_MAX_RACELINES];
    int laps[2], gates[2];
};

template <class S>
DState<S> dr_load(const dr_drace_state& a)
{
    DState<S> s;
    for (int i = 0; i < 4; ++i)
    {
        s.pos[i] = (S)a.pos[i];
        s.vel[i] = (S)a.vel[i];
    }
    for (int g = 0; g < 2; ++g)
    {
        for (int r = 0; r < DR_DRACE_MAX_RACELINES; ++r)
            s.sp[g][r] = (S)a.s[g][r];
        s.laps[g] = a.laps[g];
        s.gates[g] = a.gates[g];
    }
    return s;
}

template <class S>
void dr_store(const DState<S>& s, dr_drace_state& a)
{
    for (int i = 0; i < 4; ++i)
    {
        a.pos[i] = (float)s.pos[i];
        a.vel[i] = (float)s.vel[i];
    }
    for (int g = 0; g < 2; ++g)
    {
        for (int r = 0; r < DR_DRACE_MAX_RACELINES; ++r)
            a.s[g][r] = (float)s.sp[g][r];
        a.laps[g] = s.laps[g];
        a.gates[g] = s.gates[g];
    }
}

// sqDistToSample (env_dronerace.h:96-102)
template <class S>
S dr_sq(const float* tp, int i, const S* pos)
{
    S sq = 0;
    for (int d = 0; d < 2; ++d)
    {
        const S dx = (S)tp[i * 2 + d] - pos[d];
        sq = sq + dx * dx;
    }
    return sq;
}

// projectOnTrack (:69-93): first sample of least distance, s = i / 512
template <class S>
S dr_project_full(const float* tp, const S* pos)
{
    S bestS = 0, bestD2 = (S)3.402823466e+38f;
    for (int i = 0; i < DR_DRACE_SAMPLES; ++i)
    {
        const S d2 = dr_sq(tp, i, pos);
        if (d2 < bestD2)
        {
            bestD2 = d2;
            bestS = (S)i * ((S)1 / (S)DR_DRACE_SAMPLES);
        }
    }
    return bestS;
}

// findMinAlongDirection (:105-139): greedy walk, then 10 samples of margin
template <class S>
int dr_find_min(const float* tp, int i0, const S* pos, int delta, S base, S* best)
{
    const int N = DR_DRACE_SAMPLES;
    int bestI = i0;
    S bestD = base;
    int cur = (i0 + delta + N) % N;
    S curD = dr_sq(tp, cur, pos);
    while (curD < bestD)
    {
        bestI = cur;
        bestD = curD;
        cur = (cur + delta + N) % N;
        curD = dr_sq(tp, cur, pos);
    }
    for (int i = 0; i < 10; ++i)
    {
        curD = dr_sq(tp, cur, pos);
        if (curD < bestD)
        {
            bestI = cur;
            bestD = curD;
        }
        cur = (cur + delta + N) % N;
    }
    *best = bestD;
    return bestI;
}

// fastProjectOnTrack (:143-209) from prevS (< 0: full scan)
template <class S>
S dr_fast_project(const float* tp, const S* pos, S prevS)
{
    const int N = DR_DRACE_SAMPLES;
    if (prevS < (S)0)
        return dr_project_full(tp, pos);
    const int baseI = (int)(prevS * (S)N + (S)0.5f) % N;
    const S base = dr_sq(tp, baseI, pos);
    S bD, fD;
    const int bI = dr_find_min(tp, baseI, pos, -1, base, &bD);
    const int fI = dr_find_min(tp, baseI, pos, +1, base, &fD);
    return fD < bD ? (S)fI / (S)N : (S)bI / (S)N;
}

// sampleCenterline (:55-66), index held at 511
template <class S>
void dr_sample_line(const float* tp, S s, S out[2])
{
    const int N = DR_DRACE_SAMPLES;
    s = s - flo(s);
    const S idx = s * (S)N;
    int i0 = (int)idx;
    if (i0 > N - 1)
        i0 = N - 1;
    const int i1 = (i0 + 1) % N;
    const S t = idx - (S)i0;
    for (int d = 0; d < 2; ++d)
        out[d] = ((S)1 - t) * (S)tp[i0 * 2 + d] + t * (S)tp[i1 * 2 + d];
}

// computePIDAction (:473-541); powf(x, y) = exp(y log x)
template <class S>
void dr_pid(const DrEnv<S>& E, const DState<S>& s, int agent, const dr_drace_pid& pid, S out[2])
{
    S target[2];
    dr_sample_line(E.line(pid.raceline_index), s.sp[agent][pid.raceline_index] + E.f(E.e.target_distance), target);
    const S* p = s.pos + agent * 2;
    const S* v = s.vel + agent * 2;
    S sqErr = 0;
    for (int d = 0; d < 2; ++d)
    {
        const S e = target[d] - p[d];
        sqErr = sqErr + e * e;
    }
    const S invDist = (S)1 / sq_root(sqErr + (S)1e-5f);
    S vPar = 0;
    for (int d = 0; d < 2; ++d)
        vPar = vPar + v[d] * ((target[d] - p[d]) * invDist);
    for (int d = 0; d < 2; ++d)
    {
        const S err = (target[d] - p[d]) * invDist;
        const S lat = v[d] - vPar * err;
        out[d] = E.f(pid.kp) * err + E.f(pid.kd) * (-lat);
    }
    if (pid.repulsion_factor != 0.0f)
    {
        const int o = 1 - agent;
        S diff[2], d2 = 0;
        for (int d = 0; d < 2; ++d)
        {
            diff[d] = s.pos[o * 2 + d] - p[d];
            d2 = d2 + diff[d] * diff[d];
        }
        const S dist = sq_root(d2) + (S)1e-8f;
        if (dist < E.f(pid.repulsion_dist_fact) * E.f(E.e.drone_radius) * (S)2)
        {
            const S x = dist * (S)0.5f / E.f(E.e.drone_radius);
            const S sc = E.f(pid.repulsion_factor) / ex((E.f(pid.repulsion_power) + (S)1) * lg(x));
            for (int d = 0; d < 2; ++d)
                out[d] = out[d] - sc * diff[d];
        }
    }
    S sqA = 0;
    for (int d = 0; d < 2; ++d)
        sqA = sqA + out[d] * out[d];
    const S ma = E.f(E.e.max_accel[agent]);
    if (sqA > ma * ma)
    {
        const S fct = ma / sq_root(sqA);
        for (int d = 0; d < 2; ++d)
            out[d] = out[d] * fct;
    }
}

// updateBelief (:212-259), the likelihood in the log domain
template <class S>
void dr_update_belief(const DrEnv<S>& E, S* belief, const S obs[2], const S (*nom)[2], S maxA)
{
    const S sigmaEnv = (S)0.2f;
    const S log2pi = lg((S)2.0f * (S)3.14159265358979323846f);
    const int NT = E.e.n_models;
    S sum = 0;
    for (int th = 0; th < NT; ++th)
    {
        const S sqA = nom[th][0] * nom[th][0] + nom[th][1] * nom[th][1];
        S scale = 1;
        if (sqA > maxA * maxA)
            scale = maxA / sq_root(sqA);
        const S dx = obs[0] - nom[th][0] * scale;
        const S dy = obs[1] - nom[th][1] * scale;
        const S sqDist = dx * dx + dy * dy;
        const S an = E.f(E.e.opp_pid[th].action_noise);
        const S sqSigma = sigmaEnv * sigmaEnv + an * an;
        const S ll = (S)-0.5f * sqDist / sqSigma - log2pi - lg(sqSigma);
        belief[th] = belief[th] * ex(ll);
        sum = sum + belief[th];
    }
    if (sum < (S)1e-20f)
        for (int th = 0; th < NT; ++th)
            belief[th] = (S)1 / (S)NT;
    else
        for (int th = 0; th < NT; ++th)
            belief[th] = belief[th] / sum;
}

// squared distance to rectangle o (:274-283)
template <class S>
S dr_sq_rect(const DrEnv<S>& E, int o, const S* c)
{
    S sq = 0;
    for (int d = 0; d < 2; ++d)
    {
        const S dist = mx((S)0, mx(E.f(E.e.obstacles[o][d]) - c[d], c[d] - E.f(E.e.obstacles[o][2 + d])));
        sq = sq + dist * dist;
    }
    return sq;
}

// trackBoundaryDist (:262-308)
template <class S>
S dr_bdist(const DrEnv<S>& E, const DState<S>& s, bool lose)
{
    const int me = E.e.i_mppi;
    if (lose && s.laps[1 - me] >= E.e.n_win_laps)
        return (S)-1;
    const S* c = s.pos + me * 2;
    S m = (S)INFINITY;
    for (int d = 0; d < 2; ++d)
        m = mn(m, mn(c[d] - E.f(E.e.arena_min[d]), E.f(E.e.arena_max[d]) - c[d]));
    for (int o = 0; o < E.e.n_obstacles; ++o)
        m = mn(m, sq_root(dr_sq_rect(E, o, c)));
    for (int o = 0; o < E.e.n_round_obstacles; ++o)
    {
        S sq = 0;
        for (int d = 0; d < 2; ++d)
        {
            const S dx = c[d] - E.f(E.e.round_centers[o][d]);
            sq = sq + dx * dx;
        }
        m = mn(m, sq_root(sq) - E.f(E.e.round_radius[o]));
    }
    S sq = 0;
    for (int d = 0; d < 2; ++d)
    {
        const S dx = s.pos[d] - s.pos[d + 2];
        sq = sq + dx * dx;
    }
    return mn(m, sq_root(sq) - E.f(E.e.drone_radius));
}

// isWinner (:316-359), the opponent's wall test as written (+ radius on both sides)
template <class S>
bool dr_winner(const DrEnv<S>& E, const DState<S>& s)
{
    const int me = E.e.i_mppi;
    if (s.laps[me] >= E.e.n_win_laps)
        return true;
    const S* c = s.pos + (1 - me) * 2;
    const S r = E.f(E.e.drone_radius);
    for (int d = 0; d < 2; ++d)
        if (c[d] < E.f(E.e.arena_min[d]) + r || c[d] > E.f(E.e.arena_max[d]) + r)
            return true;
    for (int o = 0; o < E.e.n_obstacles; ++o)
        if (dr_sq_rect(E, o, c) < r * r)
            return true;
    for (int o = 0; o < E.e.n_round_obstacles; ++o)
    {
        S sq = 0;
        for (int d = 0; d < 2; ++d)
        {
            const S dx = c[d] - E.f(E.e.round_centers[o][d]);
            sq = sq + dx * dx;
        }
        const S rr = r + E.f(E.e.round_radius[o]);
        if (sq < rr * rr)
            return true;
    }
    return false;
}

// getAnyAdvance (:374-394)
template <class S>
S dr_advance(const DrEnv<S>& E, const DState<S>& s, int agent)
{
    const int g = s.gates[agent], next = (g + 1) % E.e.n_gates;
    S g2 = 0, c2 = 0;
    for (int d = 0; d < 2; ++d)
    {
        const S a = E.f(E.e.gate_centers[next][d]) - s.pos[agent * 2 + d];
        const S b = E.f(E.e.gate_centers[next][d]) - E.f(E.e.gate_centers[g][d]);
        g2 = g2 + a * a;
        c2 = c2 + b * b;
    }
    return (S)(float)(g + E.e.n_gates * s.laps[agent]) + (S)0.8f * ((S)1 - sq_root(g2 / c2));
}

// updateGates<false> (:412-470)
template <class S>
void dr_gates(const DrEnv<S>& E, DState<S>& s, const S prev[4])
{
    for (int a = 0; a < 2; ++a)
    {
        if (s.sp[a][0] >= (S)0)
            for (int r = 0; r < E.e.n_racelines; ++r)
                s.sp[a][r] = dr_fast_project(E.line(r), s.pos + a * 2, s.sp[a][r]);
        const int next = (s.gates[a] + 1) % E.e.n_gates;
        S num = 0, den = 0;
        for (int d = 0; d < 2; ++d)
        {
            num = num + E.f(E.e.gate_vectors[next][d]) * (E.f(E.e.gate_centers[next][d]) - prev[a * 2 + d]);
            den = den + E.f(E.e.gate_vectors[next][d]) * (s.pos[a * 2 + d] - prev[a * 2 + d]);
        }
        if (ab(den) < (S)1e-10f)
            continue;
        const S lam = num / den;
        if (lam < (S)0 || lam > (S)1)
            continue;
        S sq = 0;
        for (int d = 0; d < 2; ++d)
        {
            const S v = ((S)1 - lam) * prev[a * 2 + d] + lam * s.pos[a * 2 + d] - E.f(E.e.gate_centers[next][d]);
            sq = sq + v * v;
        }
        if (sq <= E.f(E.e.gate_radius[next]) * E.f(E.e.gate_radius[next]))
        {
            s.gates[a] = s.gates[a] + 1;
            if (s.gates[a] == E.e.n_gates)
            {
                s.gates[a] = 0;
                s.laps[a] = s.laps[a] + 1;
            }
        }
    }
}

// environmentStep<belief != NULL, false, false, lose> (:546-680): returns
// 0 none, 1 lose, 2 win, 3 opponent wins (when !lose)
template <class S>
int env_step(const DrEnv<S>& E, int theta, const S ego[2], DState<S>& s, Normals& rng, bool lose, S* belief = nullptr, bool noise = true)
{
    const dr_drace_env& e = E.e;
    const int me = e.i_mppi, opp = 1 - me;
    S prev[4], act[4], noms[DR_SR_MAX_MODELS][2];
    for (int i = 0; i < 4; ++i)
        prev[i] = s.pos[i];
    for (int th = 0; th < e.n_models; ++th)
        if (belief || th == theta)
            dr_pid(E, s, opp, e.opp_pid[th], noms[th]);
    act[me * 2] = ego[0];
    act[me * 2 + 1] = ego[1];
    for (int d = 0; d < 2; ++d)
    {
        S u = noms[theta][d];
        if (noise && e.opp_pid[theta].action_noise != 0.0f)
            u = u + E.f(e.opp_pid[theta].action_noise) * (S)rng.next();
        act[opp * 2 + d] = u;
    }
    for (int a = 0; a < 2; ++a)
    {
        const S ma = E.f(e.max_accel[a]);
        S sq = 0;
        for (int d = 0; d < 2; ++d)
            sq = sq + act[a * 2 + d] * act[a * 2 + d];
        const S factor = sq > ma * ma ? ma / sq_root(sq) : (S)1;
        for (int d = 0; d < 2; ++d)
        {
            S v = act[a * 2 + d] * factor;
            if (e.action_noise != 0.0f && noise)
                v = v + (S)rng.next() * E.f(e.action_noise);
            act[a * 2 + d] = v;
        }
    }
    for (int i = 0; i < 4; ++i)
        s.pos[i] = s.pos[i] + E.f(e.dt) * s.vel[i];
    for (int i = 0; i < 4; ++i)
        s.vel[i] = s.vel[i] + E.f(e.dt) * act[i];
    for (int a = 0; a < 2; ++a)
    {
        const S spd = sq_root(s.vel[a * 2] * s.vel[a * 2] + s.vel[a * 2 + 1] * s.vel[a * 2 + 1]);
        if (spd > E.f(e.max_speed[a]))
        {
            const S sc = E.f(e.max_speed[a]) / spd;
            for (int d = 0; d < 2; ++d)
                s.vel[a * 2 + d] = s.vel[a * 2 + d] * sc;
        }
    }
    if ((e.pos_noise != 0.0f || e.speed_noise != 0.0f) && noise)
        for (int i = 0; i < 4; ++i)
        {
            if (e.pos_noise != 0.0f)
                s.pos[i] = s.pos[i] + (S)rng.next() * E.f(e.pos_noise);
            if (e.speed_noise != 0.0f)
                s.vel[i] = s.vel[i] + (S)rng.next() * E.f(e.speed_noise);
        }
    dr_gates(E, s, prev);
    if (belief)
    {
        const S obs[2] = {act[opp * 2], act[opp * 2 + 1]};
        dr_update_belief(E, belief, obs, noms, E.f(e.max_accel[opp]));
    }
    if (dr_bdist(E, s, lose) < E.f(e.drone_radius))
        return 1;
    if (dr_winner(E, s))
        return 2;
    if (!lose && dr_bdist(E, s, true) < E.f(e.drone_radius))
        return 3;
    return 0;
}

// stateCost / finalCost (costs.cuh:7-39) on DroneRace
template <class S>
S stage_cost(const DrEnv<S>& E, const dr_sr_cost& K, const DState<S>& s, int)
{
    const S r = E.f(E.e.drone_radius);
    const S bd = dr_bdist(E, s, false);
    S cost = 0;
    if (bd < (S)K.boundary_threshold_factor * r)
    {
        const S q = bd / r;
        cost = cost + (S)K.boundary_cost * (q * q);
    }
    if (bd < r * (S)K.coll_dist_factor)
        cost = cost + (S)K.outside_cost;
    if (s.laps[1 - E.e.i_mppi] >= E.e.n_win_laps)
        cost = cost + (S)K.opp_win_cost;
    if (dr_winner(E, s))
        cost = cost - (S)K.win_cost;
    return cost;
}

template <class S>
S final_cost(const DrEnv<S>& E, const dr_sr_cost& K, const DState<S>& s)
{
    S cost = -(S)K.final_adv_weight * dr_advance(E, s, E.e.i_mppi);
    if (K.final_opp_adv_weight != 0.0f)
        cost = cost + (S)K.final_opp_adv_weight * dr_advance(E, s, 1 - E.e.i_mppi);
    return cost;
}

template <class S>
int n_models(const DrEnv<S>& E)
{
    return E.e.n_models;
}
template <class S>
DState<S> load_st(const DrEnv<S>&, const dr_drace_state& a)
{
    return dr_load<S>(a);
}
inline float max_accel_of(const dr_drace_env& e) { return e.max_accel[e.i_mppi]; }
template <class S>
DrEnv<S> make_env(const dr_drace_env& e)
{
    return DrEnv<S>(e);
}

// Natural cubic spline basis B[t][i] (buildSplineMatrix,
// mppi_controller.cu:67-141): the spline through e_i at the knots tau
// (second derivatives by the Thomas algorithm, zero at both ends), read at
// t = 0..T-1 with the end values held outside [tau_0, tau_{M-1}].
template <class S>
void spline_basis(const int* tau, int M, int T, S* B)
{
    std::vector<S> h(M - 1);
    for (int j = 0; j < M - 1; ++j)
        h[j] = (S)(tau[j + 1] - tau[j]);
    for (int basis = 0; basis < M; ++basis)
    {
        std::vector<S> y(M, (S)0), m(M, (S)0);
        y[basis] = 1;
        if (M > 2)
        {
            const int N = M - 2;
            std::vector<S> a(N), b(N), c(N), d(N), x(N);
            for (int i = 0; i < N; ++i)
            {
                const int j = i + 1;
                a[i] = h[j - 1];
                b[i] = (S)2 * (h[j - 1] + h[j]);
                c[i] = h[j];
                d[i] = (S)6 * ((y[j + 1] - y[j]) / h[j] - (y[j] - y[j - 1]) / h[j - 1]);
            }
            for (int i = 1; i < N; ++i)
            {
                const S w = a[i] / b[i - 1];
                b[i] = b[i] - w * c[i - 1];
                d[i] = d[i] - w * d[i - 1];
            }
            x[N - 1] = d[N - 1] / b[N - 1];
            for (int i = N - 2; i >= 0; --i)
                x[i] = (d[i] - c[i] * x[i + 1]) / b[i];
            for (int i = 0; i < N; ++i)
                m[i + 1] = x[i];
        }
        int seg = 0;
        for (int t = 0; t < T; ++t)
        {
            S val;
            if (t <= tau[0])
                val = y[0];
            else if (t >= tau[M - 1])
                val = y[M - 1];
            else
            {
                while (!(tau[seg] <= t && t <= tau[seg + 1]))
                    ++seg;
                const S x0 = (S)tau[seg], x1 = (S)tau[seg + 1];
                const S hh = x1 - x0;
                const S A = (x1 - (S)t) / hh;
                const S Bq = ((S)t - x0) / hh;
                val = m[seg] * (A * A * A - A) * (hh * hh) / (S)6 + m[seg + 1] * (Bq * Bq * Bq - Bq) * (hh * hh) / (S)6 + y[seg] * A +
                      y[seg + 1] * Bq;
            }
            B[(size_t)t * M + basis] = val;
        }
    }
}

// Knot noise of sample k, plan b, knot m: sigma * (z0, z1), the first
// Box-Muller pair of Philox (k, b * DR_MAX_KNOTS + m, step, 0).
inline void bm_knot_noise(uint32_t key0, uint32_t key1, uint32_t k, int b, int m, uint32_t step, float sigma, float e[2])
{
    const W4 w = philox(W4{{k, (uint32_t)(b * DR_MAX_KNOTS + m), step, 0u}}, key0, key1);
    float z0, z1;
    normal_pair(w.v[0], w.v[1], &z0, &z1);
    e[0] = sigma * z0;
    e[1] = sigma * z1;
}

// sampleContributed / splineSampleContributed (state.h:203-228), one factor.
inline bool bm_sample_contributed(int b, int tLocal, int btime, int used, int T)
{
    const int origin = (b != 0 && btime > 0) ? btime : 0;
    const int tGlobal = tLocal + origin;
    if (tGlobal >= T)
        return false;
    if ((b != 0 && b != used) || (b == 0 && used != 0 && tGlobal >= btime))
        return false;
    return true;
}

inline bool bm_spline_contributed(int b, int m, int btime, int used, int T, int M, const int* tau)
{
    return (m > 0 && bm_sample_contributed(b, tau[m - 1], btime, used, T)) || bm_sample_contributed(b, tau[m], btime, used, T) ||
           (m < M - 1 && bm_sample_contributed(b, tau[m + 1], btime, used, T));
}

// Rollouts of sample k under every true model compatible with the initial
// commitment (fullRolloutKernel, kernels.cu:34-171): action = sum_m
// B[t - t_origin][m] (knot + noise) of the current branch; the environment
// step updates the belief; a branch commits when the belief is confident.
template <class S, class EnvT, class Abi>
void bm_rollout(const EnvT& E, const dr_sr_cost& K, const Abi& x0, const dr_bm_instance& I, const float* belief0f, const float* knots,
                const S* B, uint32_t step, uint32_t k, S* ctrue, int* used, int* btime, int* len = nullptr, long long* n_steps = nullptr)
{
    const int NT = n_models(E), T = I.n_timesteps, M = I.n_knots, NK = I.n_samples;
    const uint32_t key0 = (uint32_t)(I.seed & 0xffffffffu), key1 = (uint32_t)(I.seed >> 32);
    S belief0[DR_SR_MAX_MODELS];
    for (int th = 0; th < NT; ++th)
        belief0[th] = (S)belief0f[th];
    const int firstPred = find_confident(belief0, NT, I.min_confidence);
    const int firstTime = firstPred == 0 ? -1 : 0;
    for (int th = 0; th < NT; ++th)
    {
        ctrue[th] = (S)INFINITY;
        used[th] = firstPred;
        btime[th] = firstTime;
        if (len)
            len[th] = 0;
    }
    for (int th = 0; th < NT; ++th)
    {
        if (firstPred != 0 && firstPred != th + 1)
            continue;
        auto s = load_st(E, x0);
        S belief[DR_SR_MAX_MODELS];
        for (int q = 0; q < NT; ++q)
            belief[q] = belief0[q];
        int pred = firstPred, time = firstTime, origin = 0;
        S acc = 0, decay = 1;
        for (int t = 0; t < T; ++t)
        {
            const S* Brow = B + (size_t)(t - origin) * M;
            const float* kn = knots + (size_t)pred * M * 2;
            S u[2] = {0, 0};
            for (int m = 0; m < M; ++m)
            {
                float e[2];
                bm_knot_noise(key0, key1, k, pred, m, step, I.sigma, e);
                for (int d = 0; d < 2; ++d)
                {
                    const S v = sizeof(S) == 4 ? (S)(kn[m * 2 + d] + e[d]) : (S)kn[m * 2 + d] + (S)e[d];
                    u[d] = fm(Brow[m], v, u[d]);
                }
            }
            Normals rng{key0, key1, (uint32_t)th * (uint32_t)NK + k, 4u * (uint32_t)t, step};
            const int term = env_step(E, th, u, s, rng, true, belief);
            const int np = find_confident(belief, NT, I.min_confidence);
            if (pred == 0 && np != 0)
            {
                pred = np;
                time = t + 1;
                origin = t + 1;
            }
            acc = fm(stage_cost(E, K, s, th), decay, acc);
            decay = decay * (S)0.97f;
            if (n_steps)
                ++*n_steps;
            if (len)
                len[th] = t + 1;
            if (term != 0)
                break;
        }
        ctrue[th] = acc + final_cost(E, K, s);
        used[th] = pred;
        btime[th] = time;
    }
}

// verifyNominalFailureKernel (kernels.cu:488-551) for one plan set (0: this
// step's plans, 1: the previous step's): per verification sample, an
// opponent model drawn from the belief (Philox (s | set << 30, 0xFFFFFFFF,
// step, 5), inverse CDF), then verif_horizon world steps under the plan of
// the current branch at local time with belief update and branching, noise
// from stream 5, the opponent's win not a loss; a loss at step t adds one
// failure at t.
template <class S, class EnvT, class Abi>
void bm_verify(const EnvT& E, const Abi& x0, const dr_bm_instance& I, const float* belief0f, uint32_t step, int set, const S* plans,
               int n_threads, int* fail)
{
    const int NT = n_models(E), T = I.n_timesteps, H = I.verif_horizon;
    const uint32_t key0 = (uint32_t)(I.seed & 0xffffffffu), key1 = (uint32_t)(I.seed >> 32);
    for (int t = 0; t < T; ++t)
        fail[t] = 0;
    const int nt = n_threads > 0 ? n_threads : 1;
    std::vector<int> local((size_t)nt * T, 0);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int sidx = 0; sidx < I.n_verif_samples; ++sidx)
    {
#ifdef _OPENMP
        int* cnt = &local[(size_t)omp_get_thread_num() * T];
#else
        int* cnt = &local[0];
#endif
        const uint32_t word = (uint32_t)sidx | ((uint32_t)set << 30);
        S belief[DR_SR_MAX_MODELS];
        float bf[DR_SR_MAX_MODELS];
        for (int q = 0; q < NT; ++q)
        {
            belief[q] = (S)belief0f[q];
            bf[q] = belief0f[q];
        }
        int pred = find_confident(belief, NT, I.min_confidence);
        const W4 r = philox(W4{{word, 0xFFFFFFFFu, step, 5u}}, key0, key1);
        const float u = (float)((r.v[0] >> 8) + 1u) * 5.9604644775390625e-08f;
        int theta = NT - 1;
        float acc = 0.0f;
        for (int q = 0; q < NT; ++q)
        {
            acc = acc + bf[q];
            if (u <= acc)
            {
                theta = q;
                break;
            }
        }
        auto s = load_st(E, x0);
        int origin = 0;
        for (int t = 0; t < H; ++t)
        {
            const S* pl = plans + ((size_t)pred * T + (t - origin)) * 2;
            const S ego[2] = {pl[0], pl[1]};
            Normals rng{key0, key1, word, 4u * (uint32_t)t, step};
            rng.stream = 5u;
            const int term = env_step(E, theta, ego, s, rng, false, belief);
            if (pred == 0)
            {
                const int np = find_confident(belief, NT, I.min_confidence);
                if (np != 0)
                {
                    pred = np;
                    origin = t + 1;
                }
            }
            if (term == 1)
            {
                cnt[t] += 1;
                break;
            }
            if (term != 0)
                break;
        }
    }
    for (int q = 0; q < nt; ++q)
        for (int t = 0; t < T; ++t)
            fail[t] += local[(size_t)q * T + t];
}

// One control step of one branching spline MPPI controller (kernels.cu:175-486,
// mppi_controller.cu:250-466 without verification).  In FP32 the weights
// are FP32 and the weighted sums FP64, as in the engine.
template <class S, class Cfg, class Abi>
void bm_step(const Cfg& env, const dr_sr_cost& K, const Abi& x0, const dr_bm_instance& I, const float* belief, uint32_t step,
             float lambda, const float* knots, const float* prev_dense, int n_threads, float* noise_out, double* ctrue_out, int* used_out,
             int* btime_out, double* costs_out, double* mins_out, double* nu_out, double* knots_out, double* dense_out,
             double* dense_new_out, int* fail_out, int* len_out, double* res_out)
{
    const auto E = make_env<S>(env);
    const int NT = n_models(E), NB = NT + 1, T = I.n_timesteps, M = I.n_knots, NK = I.n_samples;
    const uint32_t key0 = (uint32_t)(I.seed & 0xffffffffu), key1 = (uint32_t)(I.seed >> 32);
    std::vector<S> B((size_t)T * M);
    spline_basis<S>(I.knots, M, T, B.data());
    std::vector<float> noise((size_t)NB * M * NK * 2);
    std::vector<S> ct((size_t)NT * NK);
    std::vector<int> used((size_t)NT * NK), bt((size_t)NT * NK), lens((size_t)NT * NK);
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads > 0 ? n_threads : 1)
    for (int k = 0; k < NK; ++k)
    {
        for (int b = 0; b < NB; ++b)
            for (int m = 0; m < M; ++m)
                bm_knot_noise(key0, key1, (uint32_t)k, b, m, step, I.sigma, &noise[(((size_t)b * M + m) * NK + k) * 2]);
        S c[DR_SR_MAX_MODELS];
        int u[DR_SR_MAX_MODELS], t[DR_SR_MAX_MODELS];
        int ln[DR_SR_MAX_MODELS];
        bm_rollout<S>(E, K, x0, I, belief, knots, B.data(), step, (uint32_t)k, c, u, t, ln);
        for (int th = 0; th < NT; ++th)
        {
            ct[(size_t)th * NK + k] = c[th];
            used[(size_t)th * NK + k] = u[th];
            bt[(size_t)th * NK + k] = t[th];
            lens[(size_t)th * NK + k] = ln[th];
        }
    }
    // aggregateBranchCostsKernel (kernels.cu:175-217)
    std::vector<S> costs((size_t)NB * NK);
    for (int b = 0; b < NB; ++b)
        for (int k = 0; k < NK; ++k)
        {
            S num = 0, den = 0;
            for (int th = 0; th < NT; ++th)
            {
                if (b != 0 && b != th + 1)
                    continue;
                const S c = ct[(size_t)th * NK + k];
                if (!std::isfinite((double)c))
                    continue;
                const S w = (S)belief[th];
                num = fm(w, c, num);
                den = den + w;
            }
            costs[(size_t)b * NK + k] = den > (S)1e-30f ? num / den : (S)INFINITY;
        }
    const int initPred = find_confident(belief, NT, I.min_confidence);
    auto coeff_of = [&](int b, int m, int k) {
        S coeff = 0;
        for (int th = 0; th < NT; ++th)
        {
            if (belief[th] <= 0.0f)
                continue;
            if (b != 0 && b != th + 1)
                continue;
            if (bm_spline_contributed(b, m, bt[(size_t)th * NK + k], used[(size_t)th * NK + k], T, M, I.knots))
                coeff = coeff + (S)belief[th];
        }
        return coeff;
    };
    // computeMaskedMinCostsKernel (kernels.cu:221-301)
    std::vector<S> mins((size_t)NB * M, (S)INFINITY);
    for (int b = 0; b < NB; ++b)
    {
        if (initPred != 0 && initPred != b)
            continue;
        for (int m = 0; m < M; ++m)
        {
            S v = (S)INFINITY;
            for (int k = 0; k < NK; ++k)
                if (coeff_of(b, m, k) > (S)0.01f)
                    v = mn(v, costs[(size_t)b * NK + k]);
            mins[(size_t)b * M + m] = v;
        }
    }
    // weightedAverageKernelUnified (kernels.cu:307-408)
    std::vector<S> kn((size_t)NB * M * 2);
    std::vector<S> nu(NB, (S)0);
    for (int b = 0; b < NB; ++b)
        for (int m = 0; m < M; ++m)
        {
            const S minC = mins[(size_t)b * M + m];
            double n0 = 0, n1 = 0, dn = 0;
            if (!std::isinf((double)minC))
                for (int k = 0; k < NK; ++k)
                {
                    const S coeff = coeff_of(b, m, k);
                    if (!(coeff > (S)0.01f))
                        continue;
                    const float* e = &noise[(((size_t)b * M + m) * NK + k) * 2];
                    double wc;
                    if (sizeof(S) == 4)
                    {
                        const float w = exp32(-((float)costs[(size_t)b * NK + k] - (float)minC) / lambda);
                        wc = (double)(w * (float)coeff);
                    }
                    else
                        wc = exp(-((double)costs[(size_t)b * NK + k] - (double)minC) / (double)lambda) * (double)coeff;
                    n0 += wc * (double)e[0];
                    n1 += wc * (double)e[1];
                    dn += wc;
                }
            const double nums[2] = {n0, n1};
            for (int d = 0; d < 2; ++d)
            {
                S v = (S)knots[((size_t)b * M + m) * 2 + d];
                if (dn > 1e-30)
                    v = sizeof(S) == 4 ? (S)((float)v + (float)(nums[d] / dn)) : v + (S)(nums[d] / dn);
                kn[((size_t)b * M + m) * 2 + d] = v;
            }
            if (m == 0)
                nu[b] = (S)dn;
        }
    // interpolateSplineKernel (kernels.cu:411-431) from the unclamped knots
    std::vector<S> dense((size_t)NB * T * 2);
    for (int b = 0; b < NB; ++b)
        for (int t = 0; t < T; ++t)
            for (int d = 0; d < 2; ++d)
            {
                S sum = 0;
                for (int m = 0; m < M; ++m)
                    sum = fm(B[(size_t)t * M + m], kn[((size_t)b * M + m) * 2 + d], sum);
                dense[((size_t)b * T + t) * 2 + d] = sum;
            }
    if (dense_new_out)
        for (size_t i = 0; i < dense.size(); ++i)
            dense_new_out[i] = (double)dense[i];
    if (fail_out && I.n_verif_samples > 0)
    {
        std::vector<S> prev(dense.size(), (S)0);
        if (prev_dense)
            for (size_t i = 0; i < prev.size(); ++i)
                prev[i] = (S)prev_dense[i];
        bm_verify<S>(E, x0, I, belief, step, 0, dense.data(), n_threads, fail_out);
        bm_verify<S>(E, x0, I, belief, step, 1, prev.data(), n_threads, fail_out + T);
    }
    // clampNominalKernel (kernels.cu:434-456) on the knots
    const S ma = (S)max_accel_of(env);
    for (int bm = 0; bm < NB * M; ++bm)
    {
        const S a = kn[bm * 2], c = kn[bm * 2 + 1];
        const S sq = a * a + c * c;
        if (sq > ma * ma)
        {
            const S sc = ma / sq_root(sq);
            kn[bm * 2] = a * sc;
            kn[bm * 2 + 1] = c * sc;
        }
    }
    // action, temperature, shifts (mppi_controller.cu:385-457)
    const int br = initPred;
    const S a0 = dense[(size_t)br * T * 2], a1 = dense[(size_t)br * T * 2 + 1];
    float lam = lambda;
    if (I.adapt_lambda)
    {
        const float usedNu = (float)nu[br];
        if (usedNu > 0.001f * (float)NK)
            lam = lam * 0.9f;
        else if (usedNu < 0.0005f * (float)NK)
            lam = lam * 1.2f;
    }
    std::vector<S> sh((size_t)M * 2);
    for (int m = 0; m < M; ++m)
    {
        int tEval = I.knots[m] + 1;
        if (tEval >= T)
            tEval = T - 1;
        for (int d = 0; d < 2; ++d)
        {
            S sum = 0;
            for (int j = 0; j < M; ++j)
                sum = fm(B[(size_t)tEval * M + j], kn[((size_t)br * M + j) * 2 + d], sum);
            sh[m * 2 + d] = sum;
        }
    }
    for (int m = 0; m < M * 2; ++m)
        kn[(size_t)br * M * 2 + m] = sh[m];
    for (int t = 0; t < T; ++t)
        for (int d = 0; d < 2; ++d)
            dense[((size_t)br * T + t) * 2 + d] = t + 1 < T ? dense[((size_t)br * T + t + 1) * 2 + d] : (S)0;

    if (noise_out)
        memcpy(noise_out, noise.data(), noise.size() * sizeof(float));
    for (size_t i = 0; i < ct.size(); ++i)
    {
        if (ctrue_out)
            ctrue_out[i] = (double)ct[i];
        if (used_out)
            used_out[i] = used[i];
        if (btime_out)
            btime_out[i] = bt[i];
        if (len_out)
            len_out[i] = lens[i];
    }
    if (costs_out)
        for (size_t i = 0; i < costs.size(); ++i)
            costs_out[i] = (double)costs[i];
    for (size_t i = 0; i < mins.size(); ++i)
        mins_out[i] = (double)mins[i];
    for (int b = 0; b < NB; ++b)
        nu_out[b] = (double)nu[b];
    for (size_t i = 0; i < kn.size(); ++i)
        knots_out[i] = (double)kn[i];
    for (size_t i = 0; i < dense.size(); ++i)
        dense_out[i] = (double)dense[i];
    res_out[0] = (double)a0;
    res_out[1] = (double)a1;
    res_out[2] = (double)lam;
    res_out[3] = (double)br;
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

// Natural cubic spline basis [T][M] of the knots tau.
int or_bm_basis(int f64, const int* tau, int M, int T, double* B)
{
    if (!tau || !B || M < 2 || T < 1)
        return 1;
    std::vector<float> bf((size_t)T * M);
    if (f64)
        spline_basis<double>(tau, M, T, B);
    else
    {
        spline_basis<float>(tau, M, T, bf.data());
        for (size_t i = 0; i < bf.size(); ++i)
            B[i] = bf[i];
    }
    return 0;
}

// Belief update of one observed opponent action against each model's
// nominal action noms [n_models][2] (belief updated in place).
int or_sr_update_belief(int f64, const dr_sr_env* env, double* belief, const double* obs, const double* noms)
{
    if (!env || !belief || !obs || !noms)
        return 1;
    const int NT = env->n_models;
    if (f64)
    {
        const Env<double> E(*env);
        double nm[DR_SR_MAX_MODELS][2];
        for (int th = 0; th < NT; ++th)
            for (int d = 0; d < 2; ++d)
                nm[th][d] = noms[th * 2 + d];
        update_belief(E, belief, obs, nm);
        return 0;
    }
    const Env<float> E(*env);
    float b[DR_SR_MAX_MODELS], o[2] = {(float)obs[0], (float)obs[1]}, nm[DR_SR_MAX_MODELS][2];
    for (int th = 0; th < NT; ++th)
    {
        b[th] = (float)belief[th];
        for (int d = 0; d < 2; ++d)
            nm[th][d] = (float)noms[th * 2 + d];
    }
    update_belief(E, b, o, nm);
    for (int th = 0; th < NT; ++th)
        belief[th] = b[th];
    return 0;
}

// One world step with belief update (SimulationEngine::dynStep,
// engine.cpp:286-312): the opponent follows theta; belief [n_models] in/out;
// environment noise from Philox stream `stream` (4; 5 for the verification).
int or_sr_env_step_belief(int f64, const dr_sr_env* env, int theta, const float* ego, dr_sr_state* st, float* belief, uint64_t seed,
                          uint32_t k, uint32_t t, uint32_t step, int lose_on_opp_win, uint32_t stream)
{
    Normals rng{(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32), k, 4u * t, step};
    rng.stream = stream;
    const int NT = env->n_models;
    if (f64)
    {
        const Env<double> E(*env);
        SState<double> s = load<double>(*st);
        const double u[2] = {ego[0], ego[1]};
        double b[DR_SR_MAX_MODELS];
        for (int th = 0; th < NT; ++th)
            b[th] = belief[th];
        const int r = env_step(E, theta, u, s, rng, lose_on_opp_win != 0, b);
        store(s, *st);
        for (int th = 0; th < NT; ++th)
            belief[th] = (float)b[th];
        return r;
    }
    const Env<float> E(*env);
    SState<float> s = load<float>(*st);
    const float u[2] = {ego[0], ego[1]};
    const int r = env_step(E, theta, u, s, rng, lose_on_opp_win != 0, belief);
    store(s, *st);
    return r;
}

// Branching spline MPPI step of one StratRace2D controller from its belief
// and knots [n_models + 1][M][2]: noise [n_models + 1][M][K][2], per-model
// costs / branch used / branching time [n_models][K], aggregated costs
// [n_models + 1][K], masked minima [n_models + 1][M], nu [n_models + 1],
// knots after update + clamp + shift, dense plans [n_models + 1][T][2]
// after the shift (the new plan executed), res = action[2], next lambda,
// chosen branch.  With verification: the unshifted new plans (dense_new_out)
// and the failure counts [2][T] of the new plans and of prev_dense (the
// previous step's executed plans, [n_models + 1][T][2]; NULL = zero).
// len_out [n_models][K]: environment steps each rollout ran.
int or_bm_step(int f64, const dr_sr_env* env, const dr_sr_cost* K, const dr_sr_state* x0, const dr_bm_instance* inst, const float* belief,
               uint32_t step, float lambda, const float* knots, const float* prev_dense, int n_threads, float* noise_out, double* ctrue_out,
               int* used_out, int* btime_out, double* costs_out, double* mins_out, double* nu_out, double* knots_out, double* dense_out,
               double* dense_new_out, int* fail_out, int* len_out, double* res_out)
{
    if (!env || !K || !x0 || !inst || !belief || !knots || !mins_out || !nu_out || !knots_out || !dense_out || !res_out)
        return 1;
    if (inst->n_knots < 2 || inst->n_knots > DR_MAX_KNOTS || env->n_models < 1 || env->n_models > DR_SR_MAX_MODELS)
        return 1;
    if (inst->n_verif_samples > 0 && (inst->verif_horizon < 1 || inst->verif_horizon > inst->n_timesteps))
        return 1;
    if (f64)
        bm_step<double>(*env, *K, *x0, *inst, belief, step, lambda, knots, prev_dense, n_threads, noise_out, ctrue_out, used_out, btime_out,
                        costs_out, mins_out, nu_out, knots_out, dense_out, dense_new_out, fail_out, len_out, res_out);
    else
        bm_step<float>(*env, *K, *x0, *inst, belief, step, lambda, knots, prev_dense, n_threads, noise_out, ctrue_out, used_out, btime_out,
                       costs_out, mins_out, nu_out, knots_out, dense_out, dense_new_out, fail_out, len_out, res_out);
    return 0;
}

// Display predictions of SimulationEngine::sendState (engine.cpp:96-166 for
// MPPI, :168-257 for PR-MPPI): the world from x0 under model theta, no noise,
// the opponent's win not a loss.  branching: MPPI (plans [NB][T][2], action
// of the current branch at local time t - t_origin, belief update and
// branching with min_conf); otherwise one plan [T][2] and no belief.
// full_pos [T][4] (the stop state repeated after a terminal step), actions
// [T][2] (zero after a terminal step); info = initial branch, branching time, final branch, stop reason
// (event code, 3 = truncated), stop time (-1 = none).
int or_sr_predict(const dr_sr_env* env, const dr_sr_state* x0, int theta, const float* plans, int T, int branching, const float* belief,
                  float min_conf, float* full_pos, float* actions, int* info)
{
    if (!env || !x0 || !plans || !full_pos || !actions || !info || T < 1)
        return 1;
    const Env<float> E(*env);
    const int NT = env->n_models;
    SState<float> s = load<float>(*x0);
    float b[DR_SR_MAX_MODELS] = {0, 0, 0, 0};
    int pred = 0, time = -1;
    if (branching)
    {
        for (int q = 0; q < NT; ++q)
            b[q] = belief[q];
        pred = find_confident(b, NT, min_conf);
        time = pred == 0 ? -1 : 0;
    }
    info[0] = pred;
    int origin = 0, stopReason = 3, stopTime = -1;
    Normals rng{0, 0, 0, 0, 0};
    int t = 0;
    for (; t < T; ++t)
    {
        const float* u = branching ? plans + ((size_t)pred * T + (t - origin)) * 2 : plans + (size_t)t * 2;
        actions[t * 2] = u[0];
        actions[t * 2 + 1] = u[1];
        const float ego[2] = {u[0], u[1]};
        const int term = env_step(E, theta, ego, s, rng, false, branching ? b : nullptr, false);
        if (branching && pred == 0)
        {
            const int np = find_confident(b, NT, min_conf);
            if (np != 0)
            {
                pred = np;
                time = t + 1;
                origin = t + 1;
            }
        }
        for (int i = 0; i < 4; ++i)
            full_pos[t * 4 + i] = s.pos[i];
        if (term != 0)
        {
            stopReason = term - 1;
            stopTime = t;
            break;
        }
    }
    for (int r = t + 1; r < T; ++r) // after a stop: the stop state, zero actions
    {
        for (int i = 0; i < 4; ++i)
            full_pos[r * 4 + i] = s.pos[i];
        actions[r * 2] = 0.0f;
        actions[r * 2 + 1] = 0.0f;
    }
    info[1] = time;
    info[2] = pred;
    info[3] = stopReason;
    info[4] = stopTime;
    return 0;
}

// Opponent target lateral offset and speed goal (computeOppTargetLatDist,
// env_stratrace.h:318-461) at st under theta.
double or_sr_opp_target_goal(int f64, const dr_sr_env* env, int theta, const dr_sr_state* st, double* goal)
{
    if (f64)
    {
        double lat, g;
        opp_target(Env<double>(*env), env->opp[theta], load<double>(*st), &lat, &g);
        *goal = g;
        return lat;
    }
    float lat, g;
    opp_target(Env<float>(*env), env->opp[theta], load<float>(*st), &lat, &g);
    *goal = g;
    return lat;
}

// Opponent target lateral offset (computeOppTargetLatDist, env_stratrace.h:318-461) at st under theta.
double or_sr_opp_target(int f64, const dr_sr_env* env, int theta, const dr_sr_state* st)
{
    if (f64)
    {
        double lat, goal;
        opp_target(Env<double>(*env), env->opp[theta], load<double>(*st), &lat, &goal);
        return lat;
    }
    float lat, goal;
    opp_target(Env<float>(*env), env->opp[theta], load<float>(*st), &lat, &goal);
    return lat;
}

// Algorithmic FP32 work of StratRace2D rollouts, counted with the
// op-counting scalar (add, sub, mul, div, sqrt, exp, log, floor, min/max and
// comparisons 1 each, fma 2) over n samples from x0: mode 0 the plain
// rollout (one model, dense plan, environment step + stateCost), mode 1 the
// PR-MPPI rollout (PR costs and safety), mode 2 the branching rollout of
// model theta (belief update of every model each step, spline action with
// the basis B [T][M] and knots [NB][M][2]).  Noise generation is not
// counted.  Returns the total count; steps_out = rollout-steps taken.
long long or_count_sr_ops(int mode, const dr_sr_env* env, const dr_sr_cost* K, const dr_sr_state* x0, int theta, const float* nominal,
                          int n, int T, uint64_t seed, const dr_bm_instance* bm, const float* belief, long long* steps_out)
{
    g_tally.flops = 0;
    const Env<Cn> E(*env);
    const uint32_t key0 = (uint32_t)(seed & 0xffffffffu), key1 = (uint32_t)(seed >> 32);
    long long steps = 0;
    if (mode == 2 && bm)
    {
        std::vector<double> Bd((size_t)bm->n_timesteps * bm->n_knots);
        spline_basis<double>(bm->knots, bm->n_knots, bm->n_timesteps, Bd.data());
        std::vector<Cn> B(Bd.begin(), Bd.end());
        g_tally.flops = 0;
        for (int k = 0; k < n; ++k)
        {
            Cn c[DR_SR_MAX_MODELS];
            int u[DR_SR_MAX_MODELS], t[DR_SR_MAX_MODELS];
            bm_rollout<Cn>(E, *K, *x0, *bm, belief, nominal, B.data(), 0, (uint32_t)k, c, u, t, nullptr, &steps);
        }
    }
    else
    {
        for (int k = 0; k < n; ++k)
        {
            SState<Cn> s = load<Cn>(*x0);
            Cn acc = 0, decay = 1, smax = -(Cn)INFINITY;
            for (int t = 0; t < T; ++t)
            {
                const Cn u[2] = {Cn(nominal[t * 2]), Cn(nominal[t * 2 + 1])};
                Normals rng{key0, key1, (uint32_t)k, 4u * (uint32_t)t, 0u};
                const int term = env_step(E, theta, u, s, rng, true);
                if (mode == 0)
                    acc = fm(state_cost(E, *K, s), decay, acc);
                else
                {
                    acc = fm(pr_state_cost(E, *K, s), decay, acc);
                    smax = mx(smax, pr_safety(E, *K, s));
                }
                decay = decay * Cn(0.97f);
                ++steps;
                if (term)
                    break;
            }
            acc = acc + final_cost(E, *K, s);
        }
    }
    if (steps_out)
        *steps_out = steps;
    return g_tally.flops;
}

// HiddenObs geometry primitives on the scalar of f64: which 0 squared
// distance to a rectangle (x, y, xmin, ymin, xmax, ymax), 1 squared distance
// to a half-plane strip (x, y, xmin, xmax, lambda), 2 distance to the
// annulus (x, y, cx, cy, rmin, rmax), 3 segment crosses a rectangle (px, py,
// qx, qy, xmin, ymin, xmax, ymax), 4 segment crosses the annulus (px, py, qx,
// qy, cx, cy, rmin, rmax); the predicates return 0 or 1.
double or_ho_prim(int f64, int which, const double* a)
{
    if (f64)
        switch (which)
        {
        case 0: return ho_sq_rect(a[0], a[1], a[2], a[3], a[4], a[5]);
        case 1: return ho_sq_hp(a[0], a[1], a[2], a[3], a[4]);
        case 2: return ho_annulus(a[0], a[1], a[2], a[3], a[4], a[5]);
        case 3: return ho_seg_rect(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]) ? 1.0 : 0.0;
        case 4: return ho_seg_annulus(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]) ? 1.0 : 0.0;
        default: return NAN;
        }
    float f[8];
    for (int i = 0; i < 8; ++i)
        f[i] = (float)a[i];
    switch (which)
    {
    case 0: return ho_sq_rect(f[0], f[1], f[2], f[3], f[4], f[5]);
    case 1: return ho_sq_hp(f[0], f[1], f[2], f[3], f[4]);
    case 2: return ho_annulus(f[0], f[1], f[2], f[3], f[4], f[5]);
    case 3: return ho_seg_rect(f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7]) ? 1.0 : 0.0;
    case 4: return ho_seg_annulus(f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7]) ? 1.0 : 0.0;
    default: return NAN;
    }
}

// HiddenObs state functions: what 0 boundary distance under model theta,
// 1 canSeeHiddenObs(theta), 2 getAdvance, 3 stateCost under theta (needs K),
// 4 finalCost (needs K).
double or_ho_eval(int f64, int what, const dr_ho_env* env, const dr_sr_cost* K, const dr_ho_state* st, int theta)
{
    if (!env || !st || theta < 0 || theta > 1 || ((what == 3 || what == 4) && !K))
        return NAN;
    if (f64)
    {
        const HoEnv<double> H(*env);
        const HState<double> s = ho_load<double>(*st);
        switch (what)
        {
        case 0: return ho_bdist(H, s, theta);
        case 1: return ho_can_see(H, s, theta) ? 1.0 : 0.0;
        case 2: return ho_advance(H, s);
        case 3: return stage_cost(H, *K, s, theta);
        case 4: return final_cost(H, *K, s);
        default: return NAN;
        }
    }
    const HoEnv<float> H(*env);
    const HState<float> s = ho_load<float>(*st);
    switch (what)
    {
    case 0: return ho_bdist(H, s, theta);
    case 1: return ho_can_see(H, s, theta) ? 1.0 : 0.0;
    case 2: return ho_advance(H, s);
    case 3: return stage_cost(H, *K, s, theta);
    case 4: return final_cost(H, *K, s);
    default: return NAN;
    }
}

// One HiddenObs environment step under model theta from *st (updated in
// place), noise from Philox (k, 4 t + block, step, stream); belief [2] is
// updated in place when not NULL.  Returns 0 none, 1 lose, 2 win.
int or_ho_env_step(int f64, const dr_ho_env* env, int theta, const float* ego, dr_ho_state* st, float* belief, uint64_t seed, uint32_t k,
                   uint32_t t, uint32_t step, uint32_t stream)
{
    if (!env || !ego || !st || theta < 0 || theta > 1)
        return -1;
    Normals rng{(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32), k, 4u * t, step};
    rng.stream = stream;
    if (f64)
    {
        const HoEnv<double> H(*env);
        HState<double> s = ho_load<double>(*st);
        const double u[2] = {ego[0], ego[1]};
        double b[2] = {belief ? belief[0] : 0.0, belief ? belief[1] : 0.0};
        const int r = env_step(H, theta, u, s, rng, true, belief ? b : nullptr);
        ho_store(s, *st);
        if (belief)
            for (int q = 0; q < 2; ++q)
                belief[q] = (float)b[q];
        return r;
    }
    const HoEnv<float> H(*env);
    HState<float> s = ho_load<float>(*st);
    const float u[2] = {ego[0], ego[1]};
    const int r = env_step(H, theta, u, s, rng, true, belief);
    ho_store(s, *st);
    return r;
}

// or_bm_step on HiddenObs (two models).
int or_bmho_step(int f64, const dr_ho_env* env, const dr_sr_cost* K, const dr_ho_state* x0, const dr_bm_instance* inst, const float* belief,
                 uint32_t step, float lambda, const float* knots, const float* prev_dense, int n_threads, float* noise_out, double* ctrue_out,
                 int* used_out, int* btime_out, double* costs_out, double* mins_out, double* nu_out, double* knots_out, double* dense_out,
                 double* dense_new_out, int* fail_out, int* len_out, double* res_out)
{
    if (!env || !K || !x0 || !inst || !belief || !knots || !mins_out || !nu_out || !knots_out || !dense_out || !res_out)
        return 1;
    if (inst->n_knots < 2 || inst->n_knots > DR_MAX_KNOTS)
        return 1;
    if (inst->n_verif_samples > 0 && (inst->verif_horizon < 1 || inst->verif_horizon > inst->n_timesteps))
        return 1;
    if (f64)
        bm_step<double>(*env, *K, *x0, *inst, belief, step, lambda, knots, prev_dense, n_threads, noise_out, ctrue_out, used_out, btime_out,
                        costs_out, mins_out, nu_out, knots_out, dense_out, dense_new_out, fail_out, len_out, res_out);
    else
        bm_step<float>(*env, *K, *x0, *inst, belief, step, lambda, knots, prev_dense, n_threads, noise_out, ctrue_out, used_out, btime_out,
                       costs_out, mins_out, nu_out, knots_out, dense_out, dense_new_out, fail_out, len_out, res_out);
    return 0;
}

// Algorithmic FP32 work of n branching HiddenObs rollouts (sample k, every
// model compatible with the belief) from x0, counted as or_count_sr_ops
// counts; steps_out = rollout-steps taken.
long long or_count_ho_ops(const dr_ho_env* env, const dr_sr_cost* K, const dr_ho_state* x0, const float* knots, int n,
                          const dr_bm_instance* bm, const float* belief, long long* steps_out)
{
    const HoEnv<Cn> H(*env);
    std::vector<double> Bd((size_t)bm->n_timesteps * bm->n_knots);
    spline_basis<double>(bm->knots, bm->n_knots, bm->n_timesteps, Bd.data());
    std::vector<Cn> B(Bd.begin(), Bd.end());
    g_tally.flops = 0;
    long long steps = 0;
    for (int k = 0; k < n; ++k)
    {
        Cn c[DR_SR_MAX_MODELS];
        int u[DR_SR_MAX_MODELS], t[DR_SR_MAX_MODELS];
        bm_rollout<Cn>(H, *K, *x0, *bm, belief, knots, B.data(), 0, (uint32_t)k, c, u, t, nullptr, &steps);
    }
    if (steps_out)
        *steps_out = steps;
    return g_tally.flops;
}

// DroneRace projection of (x, y) on the 512 samples tp [512][2]: the full
// scan when prevS < 0, else the greedy search from prevS.
double or_drace_project(int f64, const float* tp, double x, double y, double prevS)
{
    if (!tp)
        return NAN;
    if (f64)
    {
        const double p[2] = {x, y};
        return dr_fast_project(tp, p, prevS);
    }
    const float p[2] = {(float)x, (float)y};
    return dr_fast_project(tp, p, (float)prevS);
}

// DroneRace PD action of `agent` under opponent model `model`'s configuration.
int or_drace_pid(int f64, const dr_drace_env* env, const dr_drace_state* st, int agent, int model, double* out)
{
    if (!env || !st || !out || agent < 0 || agent > 1 || model < 0 || model >= env->n_models)
        return 1;
    if (f64)
    {
        const DrEnv<double> E(*env);
        double o[2];
        dr_pid(E, dr_load<double>(*st), agent, env->opp_pid[model], o);
        out[0] = o[0];
        out[1] = o[1];
        return 0;
    }
    const DrEnv<float> E(*env);
    float o[2];
    dr_pid(E, dr_load<float>(*st), agent, env->opp_pid[model], o);
    out[0] = o[0];
    out[1] = o[1];
    return 0;
}

// DroneRace state functions: what 0 boundary distance, 1 boundary distance
// with the opponent's win a loss, 2 isWinner, 3 isOppWinner, 4 advance of the
// controlled agent, 5 of the opponent, 6 stateCost (needs K), 7 finalCost.
double or_drace_eval(int f64, int what, const dr_drace_env* env, const dr_sr_cost* K, const dr_drace_state* st)
{
    if (!env || !st || ((what == 6 || what == 7) && !K))
        return NAN;
    auto run = [&](auto zero) -> double {
        using S = decltype(zero);
        const DrEnv<S> E(*env);
        const DState<S> s = dr_load<S>(*st);
        switch (what)
        {
        case 0: return (double)dr_bdist(E, s, false);
        case 1: return (double)dr_bdist(E, s, true);
        case 2: return dr_winner(E, s) ? 1.0 : 0.0;
        case 3: return s.laps[1 - env->i_mppi] >= env->n_win_laps ? 1.0 : 0.0;
        case 4: return (double)dr_advance(E, s, env->i_mppi);
        case 5: return (double)dr_advance(E, s, 1 - env->i_mppi);
        case 6: return (double)stage_cost(E, *K, s, 0);
        case 7: return (double)final_cost(E, *K, s);
        default: return NAN;
        }
    };
    return f64 ? run(0.0) : run(0.0f);
}

// One DroneRace environment step under opponent model theta from *st
// (updated in place), noise from Philox (k, 4 t + block, step, stream);
// belief [n_models] updated in place when not NULL.  Returns 0 none, 1 lose,
// 2 win, 3 opponent wins (lose_on_opp_win = 0).
int or_drace_env_step(int f64, const dr_drace_env* env, int theta, const float* ego, dr_drace_state* st, float* belief, uint64_t seed,
                      uint32_t k, uint32_t t, uint32_t step, uint32_t stream, int lose_on_opp_win)
{
    if (!env || !ego || !st || theta < 0 || theta >= env->n_models)
        return -1;
    Normals rng{(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32), k, 4u * t, step};
    rng.stream = stream;
    const int NT = env->n_models;
    if (f64)
    {
        const DrEnv<double> E(*env);
        DState<double> s = dr_load<double>(*st);
        const double u[2] = {ego[0], ego[1]};
        double b[DR_SR_MAX_MODELS] = {0, 0, 0, 0};
        if (belief)
            for (int q = 0; q < NT; ++q)
                b[q] = belief[q];
        const int r = env_step(E, theta, u, s, rng, lose_on_opp_win != 0, belief ? b : nullptr);
        dr_store(s, *st);
        if (belief)
            for (int q = 0; q < NT; ++q)
                belief[q] = (float)b[q];
        return r;
    }
    const DrEnv<float> E(*env);
    DState<float> s = dr_load<float>(*st);
    const float u[2] = {ego[0], ego[1]};
    const int r = env_step(E, theta, u, s, rng, lose_on_opp_win != 0, belief);
    dr_store(s, *st);
    return r;
}

// or_bm_step on DroneRace.
int or_bmdrace_step(int f64, const dr_drace_env* env, const dr_sr_cost* K, const dr_drace_state* x0, const dr_bm_instance* inst,
                    const float* belief, uint32_t step, float lambda, const float* knots, const float* prev_dense, int n_threads,
                    float* noise_out, double* ctrue_out, int* used_out, int* btime_out, double* costs_out, double* mins_out, double* nu_out,
                    double* knots_out, double* dense_out, double* dense_new_out, int* fail_out, int* len_out, double* res_out)
{
    if (!env || !K || !x0 || !inst || !belief || !knots || !mins_out || !nu_out || !knots_out || !dense_out || !res_out)
        return 1;
    if (inst->n_knots < 2 || inst->n_knots > DR_MAX_KNOTS || env->n_models < 1 || env->n_models > DR_SR_MAX_MODELS)
        return 1;
    if (inst->n_verif_samples > 0 && (inst->verif_horizon < 1 || inst->verif_horizon > inst->n_timesteps))
        return 1;
    if (f64)
        bm_step<double>(*env, *K, *x0, *inst, belief, step, lambda, knots, prev_dense, n_threads, noise_out, ctrue_out, used_out, btime_out,
                        costs_out, mins_out, nu_out, knots_out, dense_out, dense_new_out, fail_out, len_out, res_out);
    else
        bm_step<float>(*env, *K, *x0, *inst, belief, step, lambda, knots, prev_dense, n_threads, noise_out, ctrue_out, used_out, btime_out,
                       costs_out, mins_out, nu_out, knots_out, dense_out, dense_new_out, fail_out, len_out, res_out);
    return 0;
}

// Algorithmic FP32 work of n branching DroneRace rollouts, counted as
// or_count_sr_ops counts; steps_out = rollout-steps taken.
long long or_count_drace_ops(const dr_drace_env* env, const dr_sr_cost* K, const dr_drace_state* x0, const float* knots, int n,
                             const dr_bm_instance* bm, const float* belief, long long* steps_out)
{
    const DrEnv<Cn> E(*env);
    std::vector<double> Bd((size_t)bm->n_timesteps * bm->n_knots);
    spline_basis<double>(bm->knots, bm->n_knots, bm->n_timesteps, Bd.data());
    std::vector<Cn> B(Bd.begin(), Bd.end());
    g_tally.flops = 0;
    long long steps = 0;
    for (int k = 0; k < n; ++k)
    {
        Cn c[DR_SR_MAX_MODELS];
        int u[DR_SR_MAX_MODELS], t[DR_SR_MAX_MODELS];
        bm_rollout<Cn>(E, *K, *x0, *bm, belief, knots, B.data(), 0, (uint32_t)k, c, u, t, nullptr, &steps);
    }
    if (steps_out)
        *steps_out = steps;
    return g_tally.flops;
}

} // extern "C"
