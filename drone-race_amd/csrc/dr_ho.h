// HiddenObs environment (host and device): the reference's 2-D single-drone
// course with gates, fixed obstacles and one of two hidden circular
// obstacles, restated from include/environments/env_hiddenobs.h of
// pablo-gr-rey/drone-race:
//   sqDistToRectangle, sqDistToHalfPlane         :27-60
//   distToAnnulus, distToCircle                  :62-91
//   trackBoundaryDist, isOutside, isWinner       :93-145
//   segmentIntersectsRectangle / Annulus         :147-276
//   canSeeHiddenObs                              :278-295
//   getAdvance, getMaxAccel                      :297-328
//   updateGates (no margin)                      :330-376
//   environmentStep                              :378-483
// and the running / terminal costs of include/costs.cuh:7-39.  The model
// theta in {0, 1} is the hidden obstacle that exists.
//
// Arithmetic follows the reference expression by expression with no
// contraction (only the caller's discounted cost sum is a fused multiply-add);
// powf(x, 2) is x * x.  Environment noise is drawn from the same Philox
// sequence as the StratRace2D environment (counter (k, 4 t + block, step,
// stream)) in the reference's draw order.
#pragma once

#include "dr_sr.h"

namespace dr {
namespace ho {

constexpr int NG = 3; // gates
constexpr int NR = 3; // rectangle obstacles

struct St
{
    float pos[2], vel[2];
    int laps, gates;
};

DR_HD St state_from_abi(const dr_ho_state& a)
{
    St s;
    s.pos[0] = a.pos[0];
    s.pos[1] = a.pos[1];
    s.vel[0] = a.vel[0];
    s.vel[1] = a.vel[1];
    s.laps = a.laps;
    s.gates = a.gates;
    return s;
}

DR_HD void state_to_abi(const St& s, dr_ho_state& a)
{
    a.pos[0] = s.pos[0];
    a.pos[1] = s.pos[1];
    a.vel[0] = s.vel[0];
    a.vel[1] = s.vel[1];
    a.laps = s.laps;
    a.gates = s.gates;
}

// squared distance to the rectangle (0 inside)
DR_HD float sq_dist_rect(float x, float y, float xmin, float ymin, float xmax, float ymax)
{
    const float dx = x - fminf(fmaxf(x, xmin), xmax);
    const float dy = y - fminf(fmaxf(y, ymin), ymax);
    return dx * dx + dy * dy;
}

// squared distance to {xmin <= x <= xmax, x + y >= lambda}
DR_HD float sq_dist_half_plane(float x, float y, float xmin, float xmax, float lambda)
{
    if (xmin <= x && x <= xmax && x + y >= lambda)
        return 0.0f;
    const float projx = fmaxf(fminf((x - y + lambda) * 0.5f, xmax), xmin);
    const float projy = lambda - projx;
    float sq = (x - projx) * (x - projx) + (y - projy) * (y - projy);
    const float y_left = fmaxf(y, lambda - xmin);
    sq = fminf(sq, (x - xmin) * (x - xmin) + (y - y_left) * (y - y_left));
    const float y_right = fmaxf(y, lambda - xmax);
    sq = fminf(sq, (x - xmax) * (x - xmax) + (y - y_right) * (y - y_right));
    return sq;
}

// distance to the annulus obstacle of the top-left quadrant (+inf outside the quadrant)
DR_HD float dist_annulus(float x, float y, float xc, float yc, float rmin, float rmax)
{
    x = x - xc;
    y = y - yc;
    const float rSq = x * x + y * y;
    const float rminSq = rmin * rmin, rmaxSq = rmax * rmax;
    if (x <= 0.0f && y >= 0.0f && (rSq <= rminSq || rSq >= rmaxSq))
        return 0.0f;
    if (x <= 0.0f && y >= 0.0f)
    {
        const float r = __builtin_sqrtf(rSq);
        return fminf(r - rmin, rmax - r);
    }
    return __builtin_inff();
}

DR_HD float dist_circle(float x, float y, float xc, float yc, float radius)
{
    return __builtin_sqrtf((x - xc) * (x - xc) + (y - yc) * (y - yc)) - radius;
}

// trackBoundaryDist: distance to the closest obstacle, arena wall or the
// hidden obstacle of model theta
DR_HD float boundary_dist(const dr_ho_env& E, const St& s, int theta)
{
    float sqMin = __builtin_inff();
    const float x = s.pos[0], y = s.pos[1];
    for (int r = 0; r < NR; ++r)
        sqMin = fminf(sqMin, sq_dist_rect(x, y, E.rect_obstacles[r][0], E.rect_obstacles[r][1], E.rect_obstacles[r][2],
                                          E.rect_obstacles[r][3]));
    const float* L = E.dbl_hp_limits;
    const float* G = E.dbl_hp_lambdas;
    sqMin = fminf(sqMin, fminf(fminf(sq_dist_half_plane(-x, y, -L[1], -L[0], G[0]), sq_dist_half_plane(-x, -y, -L[1], -L[0], G[1])),
                               fminf(sq_dist_half_plane(x, y, L[2], L[3], G[2]), sq_dist_half_plane(x, -y, L[2], L[3], G[3]))));
    float d = __builtin_sqrtf(sqMin);
    d = fminf(d, fminf(fminf(x - E.arena_min[0], E.arena_max[0] - x), fminf(y - E.arena_min[1], E.arena_max[1] - y)));
    d = fminf(d, dist_annulus(x, y, E.annulus_center[0], E.annulus_center[1], E.annulus_radius[0], E.annulus_radius[1]));
    d = fminf(d, dist_circle(x, y, E.hidden_centers[theta][0], E.hidden_centers[theta][1], E.hidden_radius[theta]));
    return d;
}

DR_HD bool is_winner(const dr_ho_env& E, const St& s)
{
    return s.laps >= E.n_win_laps;
}

// Liang-Barsky test of the segment p -> q against the rectangle
DR_HD bool seg_rect(float px, float py, float qx, float qy, float xmin, float ymin, float xmax, float ymax)
{
    const float dx = qx - px, dy = qy - py;
    float tmin = 0.0f, tmax = 1.0f;
    if (__builtin_fabsf(dx) < 1e-12f)
    {
        if (px < xmin || px > xmax)
            return false;
    }
    else
    {
        float t1 = (xmin - px) / dx;
        float t2 = (xmax - px) / dx;
        if (t1 > t2)
        {
            const float tmp = t1;
            t1 = t2;
            t2 = tmp;
        }
        tmin = fmaxf(tmin, t1);
        tmax = fminf(tmax, t2);
        if (tmin > tmax)
            return false;
    }
    if (__builtin_fabsf(dy) < 1e-12f)
    {
        if (py < ymin || py > ymax)
            return false;
    }
    else
    {
        float t1 = (ymin - py) / dy;
        float t2 = (ymax - py) / dy;
        if (t1 > t2)
        {
            const float tmp = t1;
            t1 = t2;
            t2 = tmp;
        }
        tmin = fmaxf(tmin, t1);
        tmax = fminf(tmax, t2);
        if (tmin > tmax)
            return false;
    }
    return true;
}

// does the segment p -> q cross the annulus obstacle (top-left quadrant)?
DR_HD bool seg_annulus(float px, float py, float qx, float qy, float xc, float yc, float rmin, float rmax)
{
    px = px - xc;
    py = py - yc;
    qx = qx - xc;
    qy = qy - yc;
    float tmin = 0.0f, tmax = 1.0f;
    if (__builtin_fabsf(px - qx) > 1e-12f)
    {
        const float th = px / (px - qx);
        if (qx > px)
            tmax = fminf(tmax, th);
        else
            tmin = fmaxf(tmin, th);
    }
    else if (px > 0.0f)
        return false;
    if (tmax < tmin)
        return false;
    if (__builtin_fabsf(py - qy) > 1e-12f)
    {
        const float th = py / (py - qy);
        if (qy > py)
            tmin = fmaxf(tmin, th);
        else
            tmax = fminf(tmax, th);
    }
    else if (py > 0.0f)
        return false;
    if (tmax < tmin)
        return false;
    const float sqPQ = (px - qx) * (px - qx) + (py - qy) * (py - qy);
    if (sqPQ > 1e-12f)
    {
        float tm = (px * (px - qx) + py * (py - qy)) / sqPQ;
        tm = fmaxf(tmin, fminf(tmax, tm));
        const float xm = px + tm * (qx - px);
        const float ym = py + tm * (qy - py);
        if (xm * xm + ym * ym < rmin * rmin)
            return true;
    }
    float xt = px + tmin * (qx - px);
    float yt = py + tmin * (qy - py);
    if (xt * xt + yt * yt > rmax * rmax)
        return true;
    xt = px + tmax * (qx - px);
    yt = py + tmax * (qy - py);
    return xt * xt + yt * yt > rmax * rmax;
}

// canSeeHiddenObs: the line of sight to the obstacle's centre crosses no
// rectangle and not the annulus
DR_HD bool can_see(const dr_ho_env& E, const St& s, int id)
{
    const float px = s.pos[0], py = s.pos[1];
    const float qx = E.hidden_centers[id][0], qy = E.hidden_centers[id][1];
    for (int r = 0; r < NR; ++r)
        if (seg_rect(px, py, qx, qy, E.rect_obstacles[r][0], E.rect_obstacles[r][1], E.rect_obstacles[r][2], E.rect_obstacles[r][3]))
            return false;
    if (seg_annulus(px, py, qx, qy, E.annulus_center[0], E.annulus_center[1], E.annulus_radius[0], E.annulus_radius[1]))
        return false;
    return true;
}

// getAdvance: gates + NG * laps + 0.8 * (1 - |p - next gate| / |gate - next gate|)
DR_HD float advance(const dr_ho_env& E, const St& s)
{
    const int next = (s.gates + 1) % NG;
    const float* prevC = E.gate_centers[s.gates];
    const float* nextC = E.gate_centers[next];
    float sqGate = 0.0f, sqCons = 0.0f;
    for (int d = 0; d < 2; ++d)
    {
        sqGate = sqGate + (nextC[d] - s.pos[d]) * (nextC[d] - s.pos[d]);
        sqCons = sqCons + (nextC[d] - prevC[d]) * (nextC[d] - prevC[d]);
    }
    const float scale = 0.8f;
    return (float)(s.gates + NG * s.laps) + scale * (1.0f - __builtin_sqrtf(sqGate / sqCons));
}

// updateGates<false>: crossing the next gate's plane within its radius
DR_HD void update_gates(const dr_ho_env& E, St& s, const float prev[2])
{
    const int next = (s.gates + 1) % NG;
    float num = 0.0f, den = 0.0f;
    for (int d = 0; d < 2; ++d)
    {
        num = num + E.gate_vectors[next][d] * (E.gate_centers[next][d] - prev[d]);
        den = den + E.gate_vectors[next][d] * (s.pos[d] - prev[d]);
    }
    if (__builtin_fabsf(den) < 1e-10f)
        return;
    const float lambda = num / den;
    if (lambda < 0.0f || lambda > 1.0f)
        return;
    float sq = 0.0f;
    for (int d = 0; d < 2; ++d)
    {
        const float dx = (1.0f - lambda) * prev[d] + lambda * s.pos[d] - E.gate_centers[next][d];
        sq = sq + dx * dx;
    }
    const float rad = E.gate_radius[next] * E.gate_radius[next];
    if (sq <= rad)
    {
        s.gates = s.gates + 1;
        if (s.gates == NG)
        {
            s.gates = 0;
            s.laps = s.laps + 1;
        }
    }
}

// environmentStep<BELIEF, false, false>: clamp and noise of the action,
// Euler step, speed cap, state noise, gates; with BELIEF and `update`, the
// belief collapses on the true model once either hidden obstacle is in
// sight (while neither model has 0.9); terminal: 1 lose (too close to an
// obstacle), 2 win (laps).
template <bool BELIEF>
DR_HD int env_step_t(const dr_ho_env& E, int theta, const float ego[2], St& s, sr::NormalSeq& rng, float* belief, bool update = true,
                     bool apply_noise = true)
{
    const float prev[2] = {s.pos[0], s.pos[1]};
    float act[2] = {ego[0], ego[1]};
    const float sqN = act[0] * act[0] + act[1] * act[1];
    float factor = 1.0f;
    if (sqN > E.max_accel * E.max_accel)
        factor = E.max_accel / __builtin_sqrtf(sqN);
    for (int d = 0; d < 2; ++d)
    {
        act[d] = act[d] * factor;
        if (E.action_noise != 0.0f && apply_noise)
            act[d] = act[d] + rng.next() * E.action_noise;
    }
    for (int d = 0; d < 2; ++d)
        s.pos[d] = s.pos[d] + E.dt * s.vel[d];
    for (int d = 0; d < 2; ++d)
        s.vel[d] = s.vel[d] + E.dt * act[d];
    const float sqV = s.vel[0] * s.vel[0] + s.vel[1] * s.vel[1];
    if (sqV > E.max_speed * E.max_speed)
    {
        const float sc = E.max_speed / __builtin_sqrtf(sqV);
        s.vel[0] = s.vel[0] * sc;
        s.vel[1] = s.vel[1] * sc;
    }
    if (apply_noise && (E.pos_noise != 0.0f || E.speed_noise != 0.0f))
        for (int d = 0; d < 2; ++d)
        {
            if (E.pos_noise != 0.0f)
                s.pos[d] = s.pos[d] + rng.next() * E.pos_noise;
            if (E.speed_noise != 0.0f)
                s.vel[d] = s.vel[d] + rng.next() * E.speed_noise;
        }
    update_gates(E, s, prev);
    if (BELIEF && update)
    {
        if (belief[0] < 0.9f && belief[1] < 0.9f && (can_see(E, s, theta) || can_see(E, s, 1 - theta)))
        {
            belief[theta] = 1.0f;
            belief[1 - theta] = 0.0f;
        }
    }
    if (boundary_dist(E, s, theta) < E.drone_radius)
        return sr::SR_LOSE;
    if (is_winner(E, s))
        return sr::SR_WIN;
    return sr::SR_NONE;
}

// stateCost (costs.cuh:7-28): band cost near the boundary, collision cost,
// win reward (no opponent)
DR_HD float state_cost(const dr_ho_env& E, const dr_sr_cost& K, const St& s, int theta)
{
    float cost = 0.0f;
    const float bd = boundary_dist(E, s, theta);
    const float danger = K.boundary_threshold_factor * E.drone_radius;
    if (bd < danger)
    {
        const float q = bd / E.drone_radius;
        cost = cost + K.boundary_cost * (q * q);
    }
    if (bd < E.drone_radius * K.coll_dist_factor)
        cost = cost + K.outside_cost;
    if (is_winner(E, s))
        cost = cost - K.win_cost;
    return cost;
}

// finalCost (costs.cuh:31-39): -finalAdvWeight * getAdvance (the
// opponent's advance is 0 here)
DR_HD float final_cost(const dr_ho_env& E, const dr_sr_cost& K, const St& s)
{
    float cost = -K.final_adv_weight * advance(E, s);
    if (K.final_opp_adv_weight != 0.0f)
        cost = cost + K.final_opp_adv_weight * 0.0f;
    return cost;
}

} // namespace ho
} // namespace dr
