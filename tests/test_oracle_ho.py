"""Pins of the oracle's HiddenObs environment (include/environments/env_hiddenobs.h)
and of the branching spline MPPI step on it.

* geometry: squared distances to rectangles and diagonal half-plane strips and the
  distance to the quarter annulus against brute force over densely sampled
  boundaries; the segment tests (Liang-Barsky rectangle, clipped annulus) against
  dense sampling of the segment; the FP32 and FP64 restatements agree
* the course of hiddenObsEnv (python-src/main.py:705-781): boundary distances,
  lines of sight and advance at points whose values follow from the geometry
* world step: the L2 action clamp and speed cap as invariants, gate crossings and
  the lap wrap, the draw order of the environment noise (Philox normals), the
  belief collapsing on the true model when a hidden obstacle comes into sight
* running and terminal costs in closed form
* the branching step: rollouts with a constant plan are replays through the world
  step (costs, branch used, branching time); with random knots the branching
  bookkeeping replays too; verification failure counts replay
"""
import numpy as np
import pytest


def course(pkg, **kw):
    return pkg.scenarios.ho_env(**kw)


def state(pkg, pos, vel=(0.0, 0.0), laps=0, gates=0):
    s = pkg.abi.HoState()
    s.pos[:] = pos
    s.vel[:] = vel
    s.laps, s.gates = laps, gates
    return s


# ------------------------------------------------------------------ geometry

def seg_points(a, b, n):
    t = np.linspace(0.0, 1.0, n)[:, None]
    return np.asarray(a, float)[None, :] * (1 - t) + np.asarray(b, float)[None, :] * t


def test_sq_rect_brute_force(oracle):
    rng = np.random.default_rng(1)
    for _ in range(40):
        x0, y0 = rng.uniform(-3, 3, 2)
        x1, y1 = x0 + rng.uniform(0.2, 4), y0 + rng.uniform(0.2, 4)
        corners = [(x0, y0), (x1, y0), (x1, y1), (x0, y1), (x0, y0)]
        edge = np.concatenate([seg_points(corners[i], corners[i + 1], 4001) for i in range(4)])
        for x, y in rng.uniform(-6, 6, (25, 2)):
            got = oracle.ho_prim(1, "sq_rect", x, y, x0, y0, x1, y1)
            inside = x0 <= x <= x1 and y0 <= y <= y1
            if inside:
                assert got == 0.0
                continue
            brute = np.min(np.hypot(edge[:, 0] - x, edge[:, 1] - y))
            assert abs(np.sqrt(got) - brute) < 2e-3
            assert oracle.ho_prim(0, "sq_rect", x, y, x0, y0, x1, y1) == pytest.approx(got, rel=1e-5, abs=1e-6)


def test_sq_half_plane_brute_force(oracle):
    """Region {x0 <= x <= x1, x + y >= lam}: a diagonal edge and two vertical rays."""
    rng = np.random.default_rng(2)
    for _ in range(40):
        x0 = rng.uniform(-3, 1)
        x1 = x0 + rng.uniform(0.3, 3)
        lam = rng.uniform(-2, 2)
        edge = np.concatenate([seg_points((x0, lam - x0), (x1, lam - x1), 4001),
                               seg_points((x0, lam - x0), (x0, lam - x0 + 30), 60001),
                               seg_points((x1, lam - x1), (x1, lam - x1 + 30), 60001)])
        for x, y in rng.uniform(-6, 6, (25, 2)):
            got = oracle.ho_prim(1, "sq_half_plane", x, y, x0, x1, lam)
            if x0 <= x <= x1 and x + y >= lam:
                assert got == 0.0
                continue
            brute = np.min(np.hypot(edge[:, 0] - x, edge[:, 1] - y))
            assert abs(np.sqrt(got) - brute) < 2e-3
            assert oracle.ho_prim(0, "sq_half_plane", x, y, x0, x1, lam) == pytest.approx(got, rel=1e-5, abs=1e-6)


def test_annulus_brute_force(oracle):
    """Quarter annulus (x <= cx, y >= cy): 0 in the obstacle (r <= rmin or r >= rmax), the
    distance to the nearer arc in the ring, +inf outside the quadrant."""
    rng = np.random.default_rng(3)
    cx, cy, r0, r1 = 5.0, 3.0, 3.0, 5.0
    ang = np.linspace(np.pi / 2, np.pi, 20001)
    arcs = np.concatenate([np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], 1) for r in (r0, r1)])
    n_ring = 0
    for x, y in rng.uniform(-1, 11, (400, 2)):
        got = oracle.ho_prim(1, "annulus", x, y, cx, cy, r0, r1)
        u, v = x - cx, y - cy
        r = np.hypot(u, v)
        if u > 0 or v < 0:
            assert got == np.inf
        elif r <= r0 or r >= r1:
            assert got == 0.0
        else:
            n_ring += 1
            assert abs(got - np.min(np.hypot(arcs[:, 0] - x, arcs[:, 1] - y))) < 1e-3
            assert oracle.ho_prim(0, "annulus", x, y, cx, cy, r0, r1) == pytest.approx(got, rel=1e-5, abs=1e-6)
    assert n_ring > 20


def random_segments(rng, n, lo, hi):
    segs = rng.uniform(lo, hi, (n, 4))
    segs[: n // 4, 2] = segs[: n // 4, 0]          # vertical
    segs[n // 4: n // 2, 3] = segs[n // 4: n // 2, 1]  # horizontal
    return segs


def test_seg_rect_sampling(oracle):
    rng = np.random.default_rng(4)
    decided = 0
    for _ in range(30):
        x0, y0 = rng.uniform(-2, 2, 2)
        x1, y1 = x0 + rng.uniform(0.3, 3), y0 + rng.uniform(0.3, 3)
        for px, py, qx, qy in random_segments(rng, 24, -5, 5):
            pts = seg_points((px, py), (qx, qy), 20001)
            depth = np.minimum.reduce([pts[:, 0] - x0, x1 - pts[:, 0], pts[:, 1] - y0, y1 - pts[:, 1]])
            got = oracle.ho_prim(1, "seg_rect", px, py, qx, qy, x0, y0, x1, y1)
            if depth.max() > 2e-3:
                assert got == 1.0
            elif depth.max() < -2e-3:
                assert got == 0.0
            else:
                continue
            decided += 1
            assert oracle.ho_prim(0, "seg_rect", px, py, qx, qy, x0, y0, x1, y1) == got
    assert decided > 600


def test_seg_annulus_sampling(oracle):
    """The segment meets {u <= 0, v >= 0, r < rmin or r > rmax} (centre-relative).

    A horizontal segment is the exception (env_hiddenobs.h:241-247): its y-test has the sign
    of the x-test, so one above the centre (v > 0) is never blocked, and one below it is
    clipped in x only.  The oracle and the engine keep that behaviour."""
    rng = np.random.default_rng(5)
    cx, cy, r0, r1 = 5.0, 3.0, 3.0, 5.0
    hits = misses = 0
    for px, py, qx, qy in random_segments(rng, 1200, -2, 12):
        pts = seg_points((px, py), (qx, qy), 20001) - np.array([cx, cy])
        r = np.hypot(pts[:, 0], pts[:, 1])
        got = oracle.ho_prim(1, "seg_annulus", px, py, qx, qy, cx, cy, r0, r1)
        horizontal = abs(py - qy) <= 1e-12
        if horizontal and py - cy > 0:
            assert got == 0.0
            continue
        parts = [-pts[:, 0], np.maximum(r0 - r, r - r1)] + ([] if horizontal else [pts[:, 1]])
        depth = np.minimum.reduce(parts)
        if depth.max() > 2e-3:
            assert got == 1.0, (px, py, qx, qy)
            hits += 1
        elif depth.max() < -2e-3:
            assert got == 0.0, (px, py, qx, qy)
            misses += 1
        else:
            continue
        assert oracle.ho_prim(0, "seg_annulus", px, py, qx, qy, cx, cy, r0, r1) == got
    assert hits > 200 and misses > 200
    # a segment of zero length inside the ring sees through
    assert oracle.ho_prim(1, "seg_annulus", 1.2, 6.8, 1.2, 6.8, cx, cy, r0, r1) == 0.0


# ------------------------------------------------------------------ the course

def test_course_closed_forms(oracle, pkg):
    e = course(pkg)
    # the start: 0.45 above the floor, both hidden obstacles out of sight
    s0 = pkg.scenarios.ho_start_state(e)
    for f64 in (0, 1):
        for th in (0, 1):
            assert oracle.ho_eval(f64, "bdist", e, s0, th) == pytest.approx(0.45, abs=1e-6)
            assert oracle.ho_eval(f64, "can_see", e, s0, th) == 0.0
    # advance: 0.8 (1 - |p - g1| / |g1 - g0|) with g0 = (12, 7), g1 = (1, 3)
    assert oracle.ho_eval(1, "advance", e, s0) == pytest.approx(0.8 * (1 - np.hypot(0, 2.55) / np.hypot(11, 4)), rel=1e-7)
    assert oracle.ho_eval(1, "advance", e, state(pkg, (1.0, 3.0))) == pytest.approx(0.8, rel=1e-7)
    assert oracle.ho_eval(1, "advance", e, state(pkg, (12.0, 7.0), laps=1, gates=2)) == pytest.approx(5.8, rel=1e-7)
    assert oracle.ho_eval(1, "advance", e, state(pkg, (5.0, 7.0), gates=2)) == pytest.approx(2.0, rel=1e-7)
    # at the bottom hidden obstacle's centre: inside it (-radius) under model 0; under
    # model 1, 1.0 from the rectangles above (y = 6.5) and below (y = 4.5)
    c0 = state(pkg, (9.0, 5.5))
    assert oracle.ho_eval(1, "bdist", e, c0, 0) == pytest.approx(-0.8, abs=1e-7)
    assert oracle.ho_eval(1, "bdist", e, c0, 1) == pytest.approx(1.0, abs=1e-7)
    # in the bottom lane the bottom obstacle is in sight, the top one behind the small rectangle
    lane = state(pkg, (7.5, 5.5))
    assert oracle.ho_eval(1, "can_see", e, lane, 0) == 1.0
    assert oracle.ho_eval(1, "can_see", e, lane, 1) == 0.0
    # in the curved corridor: the annulus' inner arc hides both
    corr = state(pkg, (1.0, 5.0))
    assert oracle.ho_eval(1, "can_see", e, corr, 0) == 0.0 and oracle.ho_eval(1, "can_see", e, corr, 1) == 0.0
    # corridor width: between the arcs at r = 4 the distance is 1 to both
    ring = state(pkg, (5.0 - 4.0 / np.sqrt(2), 3.0 + 4.0 / np.sqrt(2)))
    assert oracle.ho_eval(1, "bdist", e, ring, 0) == pytest.approx(1.0, abs=1e-6)  # position held in FP32


def test_costs_closed_forms(oracle, pkg):
    e = course(pkg)
    K = pkg.scenarios.ho_cost()
    r = 0.3
    for y, want in [(0.45, 0.0),                                   # outside the band 1.4 r = 0.42
                    (0.40, 10.0 * (0.40 / r) ** 2),                # in the band, beyond 1.2 r = 0.36
                    (0.35, 10.0 * (0.35 / r) ** 2 + 1e7)]:         # inside the collision distance
        s = state(pkg, (1.0, y))
        assert oracle.ho_eval(1, "state_cost", e, s, 0, K) == pytest.approx(want, rel=1e-6)
    won = state(pkg, (1.0, 1.5), laps=1)
    assert oracle.ho_eval(1, "state_cost", e, won, 1, K) == pytest.approx(-1e6, rel=1e-9)
    s = state(pkg, (1.0, 1.5))
    assert oracle.ho_eval(1, "final_cost", e, s, 0, K) == pytest.approx(-500.0 * oracle.ho_eval(1, "advance", e, s), rel=1e-9)


# ------------------------------------------------------------------ world step

def test_step_clamp_and_speed_cap(oracle, pkg):
    e = course(pkg)
    s = state(pkg, (1.0, 1.5), (0.5, 0.2))
    s2, _, term = oracle.ho_env_step(1, e, 0, [30.0, 40.0], s)
    assert term == 0
    dv = np.array(s2.vel[:]) - np.array([0.5, 0.2])
    assert np.linalg.norm(dv) == pytest.approx(0.1 * 3.0, rel=1e-6)                 # |a| clamped to max_accel
    np.testing.assert_allclose(dv / np.linalg.norm(dv), [0.6, 0.8], rtol=1e-6)       # direction kept
    np.testing.assert_allclose(s2.pos[:], [1.05, 1.52], rtol=1e-6)                    # position with the old speed
    s = state(pkg, (1.0, 1.5), (1.9, 0.5))
    s2, _, _ = oracle.ho_env_step(1, e, 0, [3.0, 0.0], s)
    v = np.array(s2.vel[:])
    assert np.linalg.norm(v) == pytest.approx(2.0, rel=1e-6)                          # speed capped
    np.testing.assert_allclose(v / np.linalg.norm(v), np.array([2.2, 0.5]) / np.hypot(2.2, 0.5), rtol=1e-6)


def test_gates_and_laps(oracle, pkg):
    e = course(pkg)
    s2, _, term = oracle.ho_env_step(1, e, 0, [0.0, 0.0], state(pkg, (1.0, 2.9), (0.0, 2.0)))
    assert (s2.gates, s2.laps, term) == (1, 0, 0)
    s2, _, _ = oracle.ho_env_step(1, e, 0, [0.0, 0.0], state(pkg, (1.0, 2.8), (0.0, 1.0)))
    assert s2.gates == 0                                                              # not across yet
    s2, _, _ = oracle.ho_env_step(1, e, 0, [0.0, 0.0], state(pkg, (2.2, 2.9), (0.0, 2.0)))
    assert s2.gates == 0                                                              # outside the gate radius
    s2, _, term = oracle.ho_env_step(1, e, 0, [0.0, 0.0], state(pkg, (11.9, 7.0), (2.0, 0.0), gates=2))
    assert (s2.gates, s2.laps, term) == (0, 1, 2)                                     # lap wrap: a win
    s2, _, term = oracle.ho_env_step(1, e, 0, [0.0, 0.0], state(pkg, (1.0, 0.35), (0.0, -0.5)))
    assert term == 1                                                                  # within a radius of the floor


def test_noise_draw_order(oracle, pkg):
    e = course(pkg, pos_noise=0.2, speed_noise=0.1, action_noise=0.7)
    seed, k, t, step, stream = 0x1234_5678_9ABC, 17, 5, 3, 4
    key = (seed & 0xFFFFFFFF, seed >> 32)
    z = np.concatenate([oracle.normals4(key, (k, 4 * t + b, step, stream)) for b in range(2)]).astype(np.float64)
    s = state(pkg, (1.0, 1.5), (0.5, 0.2))
    s2, _, _ = oracle.ho_env_step(1, e, 0, [0.0, 0.0], s, seed=seed, k=k, t=t, step=step, stream=stream)
    f = lambda v: float(np.float32(v))
    a = z[:2] * f(0.7)
    vel = np.array([0.5, 0.2]) + f(0.1) * a
    pos = np.array([1.0, 1.5]) + f(0.1) * np.array([f(0.5), f(0.2)])
    pos = pos + z[[2, 4]] * f(0.2)
    vel = vel + z[[3, 5]] * f(0.1)
    np.testing.assert_allclose(s2.pos[:], pos.astype(np.float32), rtol=1e-6)
    np.testing.assert_allclose(s2.vel[:], vel.astype(np.float32), rtol=1e-6)
    s3, _, _ = oracle.ho_env_step(0, e, 0, [0.0, 0.0], s, seed=seed, k=k, t=t, step=step, stream=stream)
    np.testing.assert_allclose(s3.pos[:], s2.pos[:], rtol=1e-6)


def test_belief_collapses_in_sight(oracle, pkg):
    e = course(pkg)
    hidden = state(pkg, (1.0, 5.0), (0.0, 0.1))
    lane = state(pkg, (7.5, 5.5), (0.1, 0.0))
    for th in (0, 1):
        _, b, _ = oracle.ho_env_step(1, e, th, [0.0, 0.0], hidden, belief=[0.5, 0.5])
        np.testing.assert_array_equal(b, [0.5, 0.5])
        # the bottom obstacle in sight: the world's model is revealed whichever it is
        _, b, _ = oracle.ho_env_step(1, e, th, [0.0, 0.0], lane, belief=[0.5, 0.5])
        np.testing.assert_array_equal(b, np.eye(2, dtype=np.float32)[th])
        # a belief already at 0.9 is kept
        _, b, _ = oracle.ho_env_step(1, e, th, [0.0, 0.0], lane, belief=[0.9, 0.1])
        np.testing.assert_array_equal(b, np.float32([0.9, 0.1]))


# ------------------------------------------------------------------ the branching step

def make(pkg, K, T, tau, sigma=0.0, min_conf=0.95, belief=(0.5, 0.5), seed=13):
    return pkg.scenarios.bm_instance(K, T, seed, knots=tau, sigma=sigma, min_confidence=min_conf, belief=belief, adapt=1, lam=10.0)


def replay(oracle, pkg, e, cost, ins, x0, th, actions_of, step, belief, f64=1):
    """Rollout of one sample under model th through the world step: (cost, branch used, branching time)."""
    st, b = x0, np.array(belief, np.float32)
    pred = next((i + 1 for i in range(2) if b[i] > ins.min_confidence), 0)
    time, origin = (-1 if pred == 0 else 0), 0
    acc, decay, n = 0.0, 1.0, 0
    for t in range(ins.n_timesteps):
        n += 1
        u = actions_of(pred, t - origin)
        st, b, term = oracle.ho_env_step(f64, e, th, u, st, belief=b, seed=ins.seed, k=th * ins.n_samples + ins._k, t=t, step=step)
        if pred == 0:
            np_ = next((i + 1 for i in range(2) if b[i] > ins.min_confidence), 0)
            if np_:
                pred, time, origin = np_, t + 1, t + 1
        acc += oracle.ho_eval(f64, "state_cost", e, st, th, cost) * decay
        decay *= float(np.float32(0.97))
        if term:
            break
    return acc + oracle.ho_eval(f64, "final_cost", e, st, th, cost), pred, time, n


@pytest.mark.parametrize("start,belief", [((1.0, 4.6), (0.5, 0.5)), ((6.0, 5.4), (0.5, 0.5)), ((6.0, 5.4), (1.0, 0.0))])
def test_constant_plan_rollouts_replay(oracle, pkg, start, belief):
    e = course(pkg, action_noise=0.2)
    cost = pkg.scenarios.ho_cost()
    K, T, tau = 4, 24, [0, 5, 11, 16, 22]
    ins = make(pkg, K, T, tau, belief=belief)
    x0 = state(pkg, start, (0.3, 0.4))
    c = np.array([1.0, 0.6], np.float32)
    r = oracle.bm_step(1, e, cost, x0, ins, np.float32(belief), 5, 10.0, np.tile(c, (3, len(tau), 1)))
    committed = next((i + 1 for i in range(2) if belief[i] > 0.95), 0)
    for th in range(2):
        if committed and committed != th + 1:
            assert np.all(np.isinf(r["ctrue"][th])) and np.all(r["len"][th] == 0)
            continue
        for k in range(K):
            ins._k = k
            want = replay(oracle, pkg, e, cost, ins, x0, th, lambda p, t: c, 5, belief)
            assert r["len"][th, k] == want[3]
            # the replay holds the world in FP32 between steps
            assert r["ctrue"][th, k] == pytest.approx(want[0], rel=1e-6)
            assert (r["used"][th, k], r["btime"][th, k]) == want[1:3]
    if start == (6.0, 5.4) and not committed:
        assert np.all(r["btime"] >= 1), "the bottom obstacle comes into sight on the way"


def test_random_knot_rollouts_replay(oracle, pkg):
    """Spline actions of the current branch at local time, noise per (plan, knot, sample)."""
    e = course(pkg)
    cost = pkg.scenarios.ho_cost()
    K, T, tau = 12, 30, [0, 6, 12, 19, 28]
    ins = make(pkg, K, T, tau, sigma=0.8)
    x0 = state(pkg, (1.0, 4.0), (0.3, 1.0))
    knots = np.random.default_rng(6).normal(0, 1.5, (3, len(tau), 2)).astype(np.float32)
    r = oracle.bm_step(1, e, cost, x0, ins, np.float32([0.5, 0.5]), 2, 10.0, knots)
    B = oracle.bm_basis(1, tau, T)
    for th in range(2):
        for k in range(K):
            v = knots.astype(np.float64) + r["noise"][:, :, k, :].astype(np.float64)
            ins._k = k
            want = replay(oracle, pkg, e, cost, ins, x0, th, lambda p, t: (B[t] @ v[p]).astype(np.float32), 2, (0.5, 0.5))
            assert (r["used"][th, k], r["btime"][th, k], r["len"][th, k]) == want[1:]
            assert r["ctrue"][th, k] == pytest.approx(want[0], rel=1e-6)
    assert np.any(r["len"] < T), "some rollouts end early"
    assert 0 < np.mean(r["btime"] > 0) < 1, "some rollouts branch, some do not"


def test_verification_counts_replay(oracle, pkg):
    e = course(pkg, action_noise=0.4)
    cost = pkg.scenarios.ho_cost()
    K, T, tau = 16, 24, [0, 5, 11, 16, 22]
    ins = make(pkg, K, T, tau, sigma=0.5, min_conf=0.75, belief=(0.3, 0.7))
    ins.n_verif_samples, ins.verif_horizon, ins.beta, ins.max_verif_eps = 64, 18, 1e-5, 0.001
    x0 = state(pkg, (1.0, 0.6), (0.0, 0.5))
    rng = np.random.default_rng(9)
    knots = rng.normal(0, 1.2, (3, len(tau), 2)).astype(np.float32)
    prev = rng.normal(0, 2.5, (3, T, 2)).astype(np.float32)
    b0 = np.float32([0.3, 0.7])
    r = oracle.bm_step(0, e, cost, x0, ins, b0, 4, 10.0, knots, prev_dense=prev)
    key = (ins.seed & 0xFFFFFFFF, ins.seed >> 32)
    for set_, plans, got in [(0, r["dense_new"].astype(np.float32), r["fail_new"]), (1, prev, r["fail_old"])]:
        fail = np.zeros(T, np.int64)
        for s in range(ins.n_verif_samples):
            word = s | (set_ << 30)
            u = ((int(oracle.philox((word, 0xFFFFFFFF, 4, 5), key)[0]) >> 8) + 1) * 2.0 ** -24
            theta = 0 if u <= float(b0[0]) else 1
            st, b = x0, b0.copy()
            pred = next((i + 1 for i in range(2) if b[i] > ins.min_confidence), 0)
            origin = 0
            for t in range(ins.verif_horizon):
                st, b, term = oracle.ho_env_step(0, e, theta, plans[pred, t - origin], st, belief=b, seed=ins.seed, k=word, t=t,
                                                 step=4, stream=5)
                if pred == 0:
                    np_ = next((i + 1 for i in range(2) if b[i] > ins.min_confidence), 0)
                    if np_:
                        pred, origin = np_, t + 1
                if term == 1:
                    fail[t] += 1
                if term:
                    break
        np.testing.assert_array_equal(got, fail)
    assert r["fail_old"].sum() > 0, "the wild previous plans crash some samples"
