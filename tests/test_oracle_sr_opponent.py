"""Pins of the oracle's StratRace2D opponent controller (computeOppTargetLatDist
and computeOpponentAction, env_stratrace.h:318-530) by special cases whose
values follow in closed form, on the circle track of radius 6 (kappa = 1/6;
states on grid samples, where the tables are read without interpolation):

* isolated opponent (ego half a lap away, neither faster): target offset 0,
  speed goal = speedScale * maxSpeed[1]
* block (ego close behind, sdot_ego + 0.3 maxSpeed[0] > sdot_opp, no
  overtake): target = lat_ego (1 - exp(-s3 x)) exp(-s2 dS^2), x = sdot_ego +
  0.3 maxSpeed[0] - sdot_opp, dS = S_ego - S_opp, with sdot = (t . v) / (1 -
  kappa lat); s3 = 0 disables it
* overtake (ego close ahead and slower, |lat_ego - lat_opp| < s1): target =
  (lat_ego - sign(dE) s1) exp(-s2 dS^2); speed goal = 0.7 x the ego's speed
  along its tangent.  With one opponent the agent "behind" is the same ego at
  dS - L < 3r, so that cap holds at every distance (the reference's rule as
  written)
* clamp of the target to +-(trackWidth - r maxOppLatDistFact)
* action on the centre line at the goal speed, no lateral motion: only the
  centripetal term, v^2 kappa along the inward normal
* action at rest 0.5 m left of the centre line with target 0: full lateral
  acceleration maxAccel[1] back toward it (bang-bang branch), no tangential
  acceleration (the budget is used up)

A wrong sign of a wrapped distance, a swapped s2/s3, a dropped factor or a
wrong speed-goal branch fails one of these.  Both the FP64 oracle (2e-6) and
the FP32 oracle (1e-5) are checked; the FP64 tolerance covers the table
reads, which interpolate by ~1e-5 of a cell because the arc lengths of the
state are floats.  The reference's constants are the floats 0.3f and 0.7f.
"""
import math

import numpy as np
import pytest

F03 = float(np.float32(0.3))
F07 = float(np.float32(0.7))


@pytest.fixture(scope="module")
def env(pkg):
    return pkg.scenarios.sr_env(n_models=4)


def grid(env, i):
    """Centre-line point, unit tangent, inward normal and curvature at sample i."""
    p = np.array(env.p_grid[i][:], np.float64)
    t = np.array(env.t_grid[i][:], np.float64)
    t = t / np.linalg.norm(t)
    return p, t, np.array([-t[1], t[0]]), float(env.kappa_grid[i])


def s_of(env, i):
    return i * env.track_length / 512


def state(pkg, env, i_ego, i_opp, lat_ego, lat_opp, v_ego, v_opp):
    """Agents on grid samples, displaced laterally by lat, moving along their
    tangents at v_ego / v_opp."""
    st = pkg.abi.SrState()
    for a, (i, lat, v) in enumerate(((i_ego, lat_ego, v_ego), (i_opp, lat_opp, v_opp))):
        p, t, n, _ = grid(env, i)
        q = p + lat * n
        st.pos[2 * a], st.pos[2 * a + 1] = q
        st.vel[2 * a], st.vel[2 * a + 1] = v * t
        st.s[a] = s_of(env, i)
        st.lat_dist[a] = lat
    return st


def sdot(env, st, a, i):
    _, t, _, kap = grid(env, i)
    v = np.array(st.vel[2 * a:2 * a + 2], np.float64)
    return float(t @ v) / (1 - kap * st.lat_dist[a])


def check(oracle, env, theta, st, lat, goal=None):
    for f64, tol in ((1, 2e-6), (0, 1e-5)):
        got_lat, got_goal = oracle.sr_opp_target_goal(f64, env, theta, st)
        assert got_lat == pytest.approx(lat, abs=tol), (f64, got_lat, lat)
        if goal is not None:
            assert got_goal == pytest.approx(goal, abs=tol, rel=tol), (f64, got_goal, goal)


def test_isolated_opponent(oracle, env, pkg):
    for theta in range(4):
        st = state(pkg, env, 0, 256, 0.3, -0.2, 0.0, 0.0)
        c = env.opp[theta]
        check(oracle, env, theta, st, 0.0, c.speed_scale * env.max_speed[1])


@pytest.mark.parametrize("gap,x", [(12, 2.0), (12, 0.02), (30, 0.05), (3, 0.3)])
def test_block(oracle, env, pkg, gap, x):
    """Ego `gap` samples behind the opponent; speeds chosen so that
    sdot_ego + 0.3 maxSpeed[0] - sdot_opp = x."""
    i_opp = 100
    i_ego = i_opp - gap
    lat_ego, lat_opp = 0.9, -0.8        # |dE| = 1.7 > s1 of every model: no overtake
    v_opp = 1.5
    st = state(pkg, env, i_ego, i_opp, lat_ego, lat_opp, 0.0, v_opp)
    sd_opp = sdot(env, st, 1, i_opp)
    _, _, _, kap = grid(env, i_ego)
    v_ego = (sd_opp + x - F03 * env.max_speed[0]) * (1 - kap * lat_ego)
    st = state(pkg, env, i_ego, i_opp, lat_ego, lat_opp, v_ego, v_opp)
    xs = sdot(env, st, 0, i_ego) + F03 * env.max_speed[0] - sdot(env, st, 1, i_opp)
    dS = st.s[0] - st.s[1]
    assert -env.track_length / 4 < dS < 0
    assert abs(lat_ego - lat_opp) >= max(env.opp[m].s1 for m in range(4))
    for theta in range(4):
        c = env.opp[theta]
        want = lat_ego * (1 - math.exp(-c.s3 * xs)) * math.exp(-c.s2 * dS * dS)
        check(oracle, env, theta, st, want, c.speed_scale * env.max_speed[1])


@pytest.mark.parametrize("gap", [12, 25, 40])
def test_overtake(oracle, env, pkg, gap):
    """Ego `gap` samples ahead and slower, lateral gap 0.3 < s1 (models 2, 3)."""
    i_opp = 100
    i_ego = i_opp + gap
    lat_ego, lat_opp = 0.3, 0.0
    v_ego, v_opp = 0.5, 1.5
    st = state(pkg, env, i_ego, i_opp, lat_ego, lat_opp, v_ego, v_opp)
    dS = st.s[0] - st.s[1]
    assert dS > 0
    opp_long = v_ego  # ego's velocity along its own tangent
    for theta in (2, 3):
        c = env.opp[theta]
        dE = lat_ego - lat_opp
        assert abs(dE) < c.s1 and sdot(env, st, 1, i_opp) > sdot(env, st, 0, i_ego)
        want = (lat_ego - math.copysign(c.s1, dE)) * math.exp(-c.s2 * dS * dS)
        goal = min(c.speed_scale * env.max_speed[1], F07 * opp_long)
        check(oracle, env, theta, st, want, goal)
    # models 0, 1 (s1 = 0) never overtake: ego ahead -> no block either
    for theta in (0, 1):
        check(oracle, env, theta, st, 0.0, env.opp[theta].speed_scale * env.max_speed[1])


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_target_clamp(oracle, env, pkg, sign):
    i_opp = 100
    st = state(pkg, env, i_opp - 1, i_opp, sign * 1.9, 0.0, 2.0, 0.2)
    max_lat = env.track_width - env.drone_radius * env.max_opp_lat_dist_fact
    check(oracle, env, 0, st, sign * max_lat)


def test_action_centre_line_at_goal_speed(oracle, env, pkg):
    i_opp = 256
    goal = env.opp[0].speed_scale * env.max_speed[1]
    st = state(pkg, env, 0, i_opp, 0.0, 0.0, 0.0, goal)
    _, _, n, kap = grid(env, i_opp)
    want = goal * goal * kap * n
    for f64, tol in ((1, 2e-6), (0, 1e-5)):
        a = np.array(oracle.sr_opp_action(f64, env, 0, st))
        np.testing.assert_allclose(a, want, atol=tol)


def test_action_offset_at_rest(oracle, env, pkg):
    i_opp = 256
    st = state(pkg, env, 0, i_opp, 0.0, 0.5, 0.0, 0.0)
    _, _, n, _ = grid(env, i_opp)
    want = -env.max_accel[1] * n
    for f64, tol in ((1, 2e-6), (0, 1e-5)):
        a = np.array(oracle.sr_opp_action(f64, env, 0, st))
        np.testing.assert_allclose(a, want, atol=tol)
