"""Pins of the oracle's StratRace2D restatement (env_stratrace.h, costs.cuh).

* projection: equals a brute-force closest point on the centre-line polyline
  when the true nearest sample lies within the +-10 sample window; the signed
  lateral distance is positive left of the tangent
* environment step: an agent at rest keeps its position and takes v = dt * a
  (a clamped to max acceleration in L2); speeds never exceed max speed; the
  lap counter moves by +1 when the arc length wraps forward; the opponent's
  nominal action stays within max acceleration (the lateral/tangential split)
* costs: closed-form values of stateCost at constructed distances (boundary
  band, collision, opponent win, ego win) and of finalCost
* rollout: FP32 within float rounding of FP64 away from discrete events;
  dense update = brute-force FP64 softmax with the L2 clamp and zero-fill shift
* track tables: the circle of radius 6 has length 2 pi 6, unit tangents and
  curvature 1/6
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def race(pkg):
    return pkg.scenarios.sr_race(n_samples=64, n_timesteps=30)


def test_track_tables_circle(race):
    e = race.env
    assert e.track_length == pytest.approx(2 * math.pi * 6, rel=1e-4)
    t = np.array([e.t_grid[i][:] for i in range(512)])
    np.testing.assert_allclose(np.linalg.norm(t, axis=1), 1.0, atol=1e-6)
    k = np.array(e.kappa_grid[:])
    np.testing.assert_allclose(k, 1 / 6, rtol=1e-4)


def brute_project(p, x, y):
    best = (np.inf, 0.0, 0.0, 0.0)
    n = len(p)
    for a in range(n):
        b = (a + 1) % n
        ab = p[b] - p[a]
        t = np.clip(np.dot([x, y] - p[a], ab) / np.dot(ab, ab), 0, 1)
        q = p[a] + t * ab
        d = np.hypot(x - q[0], y - q[1])
        if d < best[0]:
            best = (d, a + t, q[0], q[1])
    return best


def test_projection_matches_brute_force(oracle, race):
    e = race.env
    p = np.array([e.p_grid[i][:] for i in range(512)], dtype=np.float64)
    L = e.track_length
    rng = np.random.default_rng(0)
    for _ in range(200):
        s_true = rng.uniform(0, L)
        i = int(s_true / L * 512) % 512
        off = rng.uniform(-1.5, 1.5)
        n = np.array([-e.t_grid[i][1], e.t_grid[i][0]])
        x, y = p[i] + off * n
        prev = s_true + rng.uniform(-0.5, 0.5)
        s, lat = oracle.sr_project(1, e, x, y, prev)
        d, idx, qx, qy = brute_project(p, x, y)
        s_b = idx / 512 * L
        assert abs(lat) == pytest.approx(d, abs=1e-9)
        ds = (s - s_b + L / 2) % L - L / 2
        assert abs(ds) < 1e-9
        # sign: positive on the left of the tangent (the inside of a counter-clockwise circle)
        assert np.sign(lat) == np.sign(off) or abs(off) < 1e-9


def test_step_from_rest_and_clamp(oracle, race, pkg):
    e = race.env
    st = race.states[0]
    for a in ([0.5, 0.2], [3.0, 4.0], [-10.0, 0.0]):
        s2, term = oracle.sr_env_step(1, e, 0, a, st, 1, 0, 0, 0)
        assert term == 0
        np.testing.assert_array_equal(np.array(s2.pos[:2]), np.array(st.pos[:2], dtype=np.float32))
        v = np.array(s2.vel[:2], dtype=np.float64)
        a = np.array(a, dtype=np.float64)
        na = np.linalg.norm(a)
        want = e.dt * (a if na <= e.max_accel[0] else a * e.max_accel[0] / na)
        np.testing.assert_allclose(v, want, rtol=1e-6, atol=1e-7)
        assert abs(s2.s[0] - st.s[0]) < 1e-5


def test_speed_cap_and_opponent_action_bound(oracle, race):
    e = race.env
    rng = np.random.default_rng(1)
    for trial in range(300):
        st = race.states[0].__class__.from_buffer_copy(race.states[0])
        for c in range(4):
            st.vel[c] = rng.uniform(-4, 4)
        st.lat_dist[0] = rng.uniform(-1, 1)
        st.lat_dist[1] = rng.uniform(-1, 1)
        theta = trial % e.n_models
        a = oracle.sr_opp_action(1, e, theta, st)
        assert np.linalg.norm(a) <= e.max_accel[1] * (1 + 1e-9)
        s2, _ = oracle.sr_env_step(0, e, theta, rng.uniform(-5, 5, 2), st, 7, trial, 0, 0)
        for ag in range(2):
            assert np.hypot(s2.vel[2 * ag], s2.vel[2 * ag + 1]) <= e.max_speed[ag] * (1 + 1e-6)


def test_lap_counter_wraps_forward(oracle, race, pkg):
    e = race.env
    L = e.track_length
    st = pkg.abi.SrState()
    i = 510
    st.pos[0], st.pos[1] = e.p_grid[i][:]
    st.s[0] = i / 512 * L
    # moving forward along the tangent at 2 m/s crosses s = L within one step
    st.vel[0], st.vel[1] = 2.5 * e.t_grid[i][0], 2.5 * e.t_grid[i][1]
    j = 300
    st.pos[2], st.pos[3] = e.p_grid[j][:]
    st.s[1] = j / 512 * L
    s2, term = oracle.sr_env_step(1, e, 0, [0.0, 0.0], st, 1, 0, 0, 0)
    assert s2.laps[0] == 1 and s2.s[0] < 1.0
    # the same lap completes the race: the ego wins
    assert term == 2


def test_state_cost_closed_forms(oracle, race, pkg):
    e, K = race.env, race.cost
    r = e.drone_radius

    def state(ego_lat, opp_dist, laps=(0, 0), opp_lat=0.0):
        st = pkg.abi.SrState()
        st.pos[0], st.pos[1] = 0.0, 0.0
        st.pos[2], st.pos[3] = opp_dist, 0.0
        st.lat_dist[0], st.lat_dist[1] = ego_lat, opp_lat
        st.laps[0], st.laps[1] = laps
        return st

    # far from everything: zero
    assert oracle.sr_state_cost(1, e, K, state(0.0, 10.0)) == 0.0
    # boundary band only: bd = track_width - |lat| in (coll, danger)
    bd = 0.6
    assert 1.3 * r < bd < 1.7 * r
    c = oracle.sr_state_cost(1, e, K, state(e.track_width - bd, 10.0))
    assert c == pytest.approx(K.boundary_cost * (bd / r) ** 2, rel=1e-6)  # inputs are FP32 fields
    # collision with the opponent: bd = d - r < r * coll -> band + outside cost
    d = 0.8
    bd = d - r
    c = oracle.sr_state_cost(1, e, K, state(0.0, d))
    assert c == pytest.approx(K.boundary_cost * (bd / r) ** 2 + K.outside_cost, rel=1e-9)
    # opponent has won / ego has won
    assert oracle.sr_state_cost(1, e, K, state(0.0, 10.0, laps=(0, 1))) == pytest.approx(K.opp_win_cost)
    assert oracle.sr_state_cost(1, e, K, state(0.0, 10.0, laps=(1, 0))) == pytest.approx(-K.win_cost)
    # opponent off the track counts as an ego win
    assert oracle.sr_state_cost(1, e, K, state(0.0, 10.0, opp_lat=e.track_width)) == pytest.approx(-K.win_cost)


def test_rollout_fp32_close_to_fp64(oracle, race):
    e, st = race.env, race.states[0]
    nom = np.zeros((30, 2), np.float32)
    c32, n32 = oracle.sr_rollout(0, e, race.cost, st, 0, nom, 128, 30, 99, 0, 1.5)
    c64, n64 = oracle.sr_rollout(1, e, race.cost, st, 0, nom, 128, 30, 99, 0, 1.5)
    np.testing.assert_array_equal(n32, n64)
    ok = (np.abs(c64) < 1e4) & (np.abs(c32) < 1e4)
    assert ok.sum() > 40
    np.testing.assert_allclose(c32[ok], c64[ok], rtol=1e-4, atol=1e-3)


def test_update_matches_brute_force(oracle, race):
    rng = np.random.default_rng(2)
    for K, T, lam in [(1, 1, 10.0), (50, 7, 10.0), (300, 20, 2.0)]:
        costs = rng.normal(0, 50, K)
        noise = rng.normal(0, 1.5, (K, T, 2)).astype(np.float32)
        nom = rng.normal(0, 1.5, (T, 2)).astype(np.float32)
        r = oracle.sr_update(1, costs, noise, nom, lam, 2.0)
        w = np.exp(-(costs - costs.min()) / float(np.float32(lam)))
        u = nom.astype(np.float64) + np.einsum("k,ktd->td", w, noise.astype(np.float64)) / w.sum()
        n = np.linalg.norm(u, axis=1, keepdims=True)
        u = np.where(n > 2.0, u * 2.0 / n, u)
        np.testing.assert_allclose(r["updated"], u, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(r["shifted"], np.vstack([u[1:], [[0.0, 0.0]]]), rtol=1e-12, atol=1e-12)
        assert r["nu"] == pytest.approx(w.sum(), rel=1e-12)
