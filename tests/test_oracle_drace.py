"""Pins of the oracle's DroneRace environment (include/environments/env_dronerace.h) and of
the branching spline MPPI step on it.

* projection: the full scan is NumPy's first argmin over the 512 samples; the greedy
  search from a nearby previous progress finds the same sample; sampleCenterline is the
  linear interpolation in closed form (wrap at 1, negative progress)
* PD opponent: at rest with the target ahead the action is kp along the line, clamped to
  max accel; velocity across the line is damped by kd; repulsion pushes away from the
  other agent; the action never exceeds max accel
* belief update: the Gaussian likelihood written out in NumPy, the uniform fallback
* boundary distance, winner tests (incl. the opponent wall test as the reference writes
  it), advance, gate crossing and the lap wrap in closed form
* world step: clamp and speed cap, the draw order of the opponent, action and state noise
  against Philox normals
* the branching step: rollouts with a constant plan are replays through the world step
  (costs, branch used, time, length); verification counts replay
"""
import numpy as np
import pytest


def course(pkg, **kw):
    return pkg.scenarios.drace_env(**kw)


def state(pkg, e, p0, p1, v0=(0.0, 0.0), v1=(0.0, 0.0), s0=None, s1=None, gates=(1, 1), laps=(0, 0)):
    st = pkg.abi.DraceState()
    st.pos[:] = [p0[0], p0[1], p1[0], p1[1]]
    st.vel[:] = [v0[0], v0[1], v1[0], v1[1]]
    for a, (p, sv) in enumerate(((p0, s0), (p1, s1))):
        sv = p[0] / (e.track_points[0][511][0]) if sv is None else sv
        for r in range(4):
            st.s[a][r] = sv if r < e.n_racelines else 0.0
    st.gates[:] = list(gates)
    st.laps[:] = list(laps)
    return st


def lines(e):
    return np.array([[list(p) for p in e.track_points[r]] for r in range(e.n_racelines)], np.float32)


# ------------------------------------------------------------------ projection

def test_projection_full_and_greedy(oracle, pkg):
    e = course(pkg)
    L = lines(e)
    rng = np.random.default_rng(1)
    for r in range(2):
        tp = L[r].astype(np.float64)
        for _ in range(60):
            i = rng.integers(5, 500)
            x, y = tp[i] + rng.normal(0, 0.6, 2)
            d2 = ((tp - [np.float32(x), np.float32(y)]) ** 2).sum(1)
            want = int(np.argmin(d2)) / 512
            assert oracle.drace_project(1, L[r], x, y, -1.0) == pytest.approx(want, abs=1e-12)
            # from a previous progress a few samples away the greedy search lands on the same sample
            prev = (i + rng.integers(-6, 7)) / 512
            assert oracle.drace_project(1, L[r], x, y, prev) == pytest.approx(want, abs=1e-12)
            assert oracle.drace_project(0, L[r], x, y, prev) == pytest.approx(want, abs=1e-7)


def test_pid_target_and_line_interpolation(oracle, pkg):
    """The PD target is the raceline at S + targetDistance, linearly interpolated: an agent
    at rest exactly on the target's direction is pushed along it by kp (clamped)."""
    e = course(pkg)
    L = lines(e).astype(np.float64)
    for s in (0.0, 0.33, 0.4102, 0.97):
        st = state(pkg, e, (float(L[0][int(s * 512)][0]) - 1.0, 0.0), (20.0, 3.0), s0=s, s1=0.9)
        t = s + np.float32(0.05)
        t = t - np.floor(t)
        i0 = int(t * 512)
        fr = t * 512 - i0
        target = (1 - fr) * L[0][i0] + fr * L[0][(i0 + 1) % 512]
        for model in (0, 1):
            tl = target if model == 0 else np.array([target[0], -target[1]])
            a = oracle.drace_pid(1, e, st, 0, model)
            d = tl - np.array(st.pos[:2], np.float64)
            want = 5.0 * d / np.sqrt(d @ d + 1e-5)
            n = np.linalg.norm(want)
            if n > 3.0:
                want *= 3.0 / n
            np.testing.assert_allclose(a, want, rtol=1e-6, atol=1e-9)


def test_pid_damping_repulsion_and_clamp(oracle, pkg):
    e = course(pkg, repulsion=30.0)
    rng = np.random.default_rng(2)
    for _ in range(40):
        p0 = (rng.uniform(0, 24), rng.uniform(-1.5, 1.5))
        p1 = (p0[0] + rng.uniform(-2, 2), p0[1] + rng.uniform(-1.5, 1.5))
        v = rng.normal(0, 1.5, 2)
        st = state(pkg, e, p0, p1, v0=v)
        for model in (0, 1):
            a = oracle.drace_pid(1, e, st, 0, model)
            assert np.linalg.norm(a) <= 3.0 * (1 + 1e-12)
    # damping: a velocity across the line, nothing else, no clamp (small gains)
    e2 = course(pkg)
    for m in range(2):
        e2.opp_pid[m].kp, e2.opp_pid[m].kd = 0.5, 0.4
    st = state(pkg, e2, (5.0, 0.0), (20.0, 2.0), v0=(0.0, 0.7), s0=5.0 / 27.5)
    a = oracle.drace_pid(1, e2, st, 0, 0)
    assert a[1] == pytest.approx(-0.4 * 0.7, rel=1e-6)
    # repulsion: the other agent close by pushes the action away from it
    e3 = course(pkg, repulsion=2.0)
    e4 = course(pkg)
    for m in range(2):
        for ee in (e3, e4):
            ee.opp_pid[m].kp, ee.opp_pid[m].kd = 0.2, 0.1
    st = state(pkg, e3, (5.0, 0.0), (5.5, 0.8), s0=5.0 / 27.5)
    diff = np.array([0.5, 0.8])
    push = oracle.drace_pid(1, e3, st, 0, 0) - oracle.drace_pid(1, e4, st, 0, 0)
    assert push @ diff < 0


def test_belief_update_closed_form(oracle, pkg):
    e = course(pkg)
    # the opponent (agent 1) at rest behind a point where the racelines split
    st = state(pkg, e, (2.0, 0.0), (9.0, 0.0), s0=2.0 / 27.5, s1=9.0 / 27.5)
    noms = np.array([oracle.drace_pid(1, e, st, 1, m) for m in range(2)])
    assert abs(noms[0][1] - noms[1][1]) > 0.1, "the two models disagree here"
    b = np.float32([0.5, 0.5])
    st2, b2, _ = oracle.drace_env_step(1, e, 1, [0.0, 0.0], st, belief=b, seed=0)
    # no noise draw changes the check: recompute the observation the same way
    e_quiet = course(pkg, action_noise=0.0, pid_noise=0.0)
    st3, b3, _ = oracle.drace_env_step(1, e_quiet, 1, [0.0, 0.0], st, belief=b, seed=0)
    obs = noms[1].copy()
    n = np.linalg.norm(obs)
    if n > 3.0:
        obs *= 3.0 / n
    sig2 = np.float32(0.2) ** 2 + np.float32(0.0) ** 2
    lik = [np.exp(-0.5 * np.sum((obs - noms[m]) ** 2) / sig2) / (2 * np.pi * sig2) for m in range(2)]
    want = 0.5 * np.array(lik)
    want /= want.sum()
    np.testing.assert_allclose(b3, want.astype(np.float32), rtol=2e-6, atol=1e-30)
    assert b3[1] > 0.99
    # far-off observations (environment action noise 50 on top of the model's action, which
    # the likelihood's sigma = 0.2 does not cover): both likelihoods underflow, uniform belief
    e_far = course(pkg, action_noise=50.0, pid_noise=0.0)
    for seed in range(3):
        _, b4, _ = oracle.drace_env_step(0, e_far, 0, [0.0, 0.0], st, belief=np.float32([0.9, 0.1]), seed=seed)
        np.testing.assert_array_equal(b4, np.float32([0.5, 0.5]))


def test_boundary_winner_advance_closed_forms(oracle, pkg):
    e = course(pkg)
    K = pkg.scenarios.drace_cost()
    # walls: the nearest wall of the controlled agent at (1, 2.4) is y = 3 (0.6 away)
    st = state(pkg, e, (1.0, 2.4), (10.0, -1.0))
    assert oracle.drace_eval(1, "bdist", e, st) == pytest.approx(0.6, abs=1e-6)
    # the round obstacle at (12.5, 0), radius 0.25
    st = state(pkg, e, (12.5, 1.0), (5.0, -2.0))
    assert oracle.drace_eval(1, "bdist", e, st) == pytest.approx(0.75, abs=1e-6)
    # the other agent, less one radius
    st = state(pkg, e, (6.0, 1.0), (6.0, -0.2))
    assert oracle.drace_eval(1, "bdist", e, st) == pytest.approx(1.2 - 0.75, abs=1e-6)
    # the opponent's win is a loss only when asked
    st = state(pkg, e, (6.0, 1.0), (6.0, -2.0), laps=(0, 1))
    assert oracle.drace_eval(1, "bdist_lose", e, st) == -1.0
    assert oracle.drace_eval(1, "opp_winner", e, st) == 1.0
    assert oracle.drace_eval(1, "state_cost", e, st, K) == pytest.approx(1000.0)
    # winner: own lap; opponent beyond the wall plus a radius; the reference's opponent wall
    # test adds the radius on the high side too, so 0.5 past the top wall is not a win
    assert oracle.drace_eval(1, "winner", e, state(pkg, e, (6.0, 1.0), (10.0, 0.0), laps=(1, 0))) == 1.0
    assert oracle.drace_eval(1, "winner", e, state(pkg, e, (6.0, 1.0), (10.0, -2.9))) == 1.0   # below -3 + 0.75
    assert oracle.drace_eval(1, "winner", e, state(pkg, e, (6.0, 1.0), (10.0, 3.5))) == 0.0    # above 3, below 3 + 0.75
    assert oracle.drace_eval(1, "winner", e, state(pkg, e, (6.0, 1.0), (10.0, 3.9))) == 1.0
    assert oracle.drace_eval(1, "winner", e, state(pkg, e, (6.0, 1.0), (12.5, 0.5))) == 1.0    # in the obstacle's reach
    # advance: gates + 2 laps + 0.8 (1 - |p - next| / |next - gate|); next of gate 1 is gate 0 at (25, 0)
    st = state(pkg, e, (10.0, 0.0), (25.0, 0.0))
    assert oracle.drace_eval(1, "advance", e, st) == pytest.approx(1 + 0.8 * (1 - 15.0 / 24.75), rel=1e-7)  # 0.8f
    assert oracle.drace_eval(1, "opp_advance", e, st) == pytest.approx(1.8, rel=1e-7)
    assert oracle.drace_eval(1, "final_cost", e, st, K) == pytest.approx(-200.0 * (1 + 0.8 * (1 - 15.0 / 24.75)), rel=1e-7)


def test_step_gates_clamp_and_noise_order(oracle, pkg):
    e = course(pkg, action_noise=0.0, pid_noise=0.0)
    # crossing gate 0 at x = 25 within its radius: lap, win
    st = state(pkg, e, (24.9, 0.3), (10.0, -1.5), v0=(2.0, 0.0))
    s2, _, term = oracle.drace_env_step(1, e, 0, [0.0, 0.0], st)
    assert (s2.gates[0], s2.laps[0], term) == (0, 1, 2)
    st = state(pkg, e, (24.9, 1.2), (10.0, -1.5), v0=(2.0, 0.0))
    s2, _, _ = oracle.drace_env_step(1, e, 0, [0.0, 0.0], st)
    assert (s2.gates[0], s2.laps[0]) == (1, 0)  # outside the gate radius
    # the ego action is clamped to max accel, the speed capped at max speed
    st = state(pkg, e, (5.0, 0.0), (10.0, -1.5), v0=(0.5, 0.2))
    s2, _, _ = oracle.drace_env_step(1, e, 0, [30.0, 40.0], st)
    dv = np.array(s2.vel[:2]) - [0.5, 0.2]
    assert np.linalg.norm(dv) == pytest.approx(0.3, rel=1e-6)
    st = state(pkg, e, (5.0, 0.0), (10.0, -1.5), v0=(2.0, 0.3))
    s2, _, _ = oracle.drace_env_step(1, e, 0, [3.0, 0.0], st)
    assert np.linalg.norm(s2.vel[:2]) == pytest.approx(2.05, rel=1e-6)
    # draw order: the opponent's PD noise (2), the action noise of agent 0 then 1 (2 + 2), then
    # position / speed noise per coordinate
    en = course(pkg, action_noise=0.3, pid_noise=0.7)
    en.pos_noise, en.speed_noise = 0.05, 0.02
    seed, k, t, step = 0x51ED, 9, 4, 2
    key = (seed & 0xFFFFFFFF, seed >> 32)
    z = np.concatenate([oracle.normals4(key, (k, 4 * t + b, step, 4)) for b in range(4)]).astype(np.float64)
    st = state(pkg, en, (5.0, 0.0), (10.0, -1.5), s0=5.0 / 27.5, s1=10.0 / 27.5)
    nom = oracle.drace_pid(1, en, st, 1, 1)
    s2, _, _ = oracle.drace_env_step(1, en, 1, [0.0, 0.0], st, seed=seed, k=k, t=t, step=step)
    f = lambda v: float(np.float32(v))
    opp = nom + f(0.7) * z[0:2]
    n = np.linalg.norm(opp)
    if n > 3.0:
        opp = opp * (3.0 / n)
    acts = np.concatenate([z[2:4] * f(0.3), opp + z[4:6] * f(0.3)])
    vel = f(0.1) * acts
    pos = np.array(st.pos[:], np.float64) + z[[6, 8, 10, 12]] * f(0.05)
    vel = vel + z[[7, 9, 11, 13]] * f(0.02)
    np.testing.assert_allclose(s2.pos[:], pos.astype(np.float32), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(s2.vel[:], vel.astype(np.float32), rtol=1e-5, atol=1e-7)


# ------------------------------------------------------------------ the branching step

def make(pkg, K, T, tau, sigma=0.0, min_conf=0.95, belief=(0.5, 0.5), seed=17):
    return pkg.scenarios.bm_instance(K, T, seed, knots=tau, sigma=sigma, min_confidence=min_conf, belief=belief, adapt=1, lam=10.0)


def replay(oracle, e, cost, ins, x0, th, k, actions_of, step, belief):
    st, b = x0, np.array(belief, np.float32)
    pred = next((i + 1 for i in range(e.n_models) if b[i] > ins.min_confidence), 0)
    time, origin = (-1 if pred == 0 else 0), 0
    acc, decay, n = 0.0, 1.0, 0
    for t in range(ins.n_timesteps):
        n += 1
        st, b, term = oracle.drace_env_step(1, e, th, actions_of(pred, t - origin), st, belief=b, seed=ins.seed,
                                            k=th * ins.n_samples + k, t=t, step=step)
        if pred == 0:
            np_ = next((i + 1 for i in range(e.n_models) if b[i] > ins.min_confidence), 0)
            if np_:
                pred, time, origin = np_, t + 1, t + 1
        acc += oracle.drace_eval(1, "state_cost", e, st, cost) * decay
        decay *= float(np.float32(0.97))
        if term:
            break
    return acc + oracle.drace_eval(1, "final_cost", e, st, cost), pred, time, n


@pytest.mark.parametrize("belief", [(0.5, 0.5), (0.02, 0.98)])
def test_constant_plan_rollouts_replay(oracle, pkg, belief):
    e = course(pkg)
    cost = pkg.scenarios.drace_cost()
    K, T, tau = 4, 30, [0, 7, 14, 21, 28]
    ins = make(pkg, K, T, tau, belief=belief, min_conf=0.9)
    x0 = state(pkg, e, (4.0, 0.0), (7.0, 0.0), v0=(1.0, 0.0), v1=(1.8, 0.0), s0=4.0 / 27.5, s1=7.0 / 27.5)
    c = np.array([2.0, 0.4], np.float32)
    r = oracle.bm_step(1, e, cost, x0, ins, np.float32(belief), 3, 10.0, np.tile(c, (3, len(tau), 1)))
    committed = next((i + 1 for i in range(2) if belief[i] > 0.9), 0)
    for th in range(2):
        if committed and committed != th + 1:
            assert np.all(np.isinf(r["ctrue"][th])) and np.all(r["len"][th] == 0)
            continue
        for k in range(K):
            want = replay(oracle, e, cost, ins, x0, th, k, lambda p, t: c, 3, belief)
            assert r["ctrue"][th, k] == pytest.approx(want[0], rel=1e-6)  # the replay holds the world in FP32
            assert (r["used"][th, k], r["btime"][th, k], r["len"][th, k]) == want[1:]
    if not committed:
        assert np.all(r["btime"] >= 1), "the racelines split and the opponent's model shows"


def test_verification_counts_replay(oracle, pkg):
    e = course(pkg)
    cost = pkg.scenarios.drace_cost()
    K, T, tau = 16, 24, [0, 5, 11, 16, 22]
    ins = make(pkg, K, T, tau, sigma=0.5, min_conf=0.8, belief=(0.4, 0.6))
    ins.n_verif_samples, ins.verif_horizon, ins.beta, ins.max_verif_eps = 48, 20, 1e-5, 0.001
    x0 = state(pkg, e, (6.0, 0.0), (7.5, 0.0), v0=(1.5, 0.0), v1=(1.8, 0.0), s0=6.0 / 27.5, s1=7.5 / 27.5)
    rng = np.random.default_rng(9)
    knots = rng.normal(0, 1.5, (3, len(tau), 2)).astype(np.float32)
    prev = np.tile(np.float32([2.5, 0.0]), (3, T, 1)) + rng.normal(0, 0.3, (3, T, 2)).astype(np.float32)
    b0 = np.float32([0.4, 0.6])
    r = oracle.bm_step(0, e, cost, x0, ins, b0, 4, 10.0, knots, prev_dense=prev)
    key = (ins.seed & 0xFFFFFFFF, ins.seed >> 32)
    for set_, plans, got in [(0, r["dense_new"].astype(np.float32), r["fail_new"]), (1, prev, r["fail_old"])]:
        fail = np.zeros(T, np.int64)
        for s in range(ins.n_verif_samples):
            word = s | (set_ << 30)
            u = np.float32(((int(oracle.philox((word, 0xFFFFFFFF, 4, 5), key)[0]) >> 8) + 1) * 2.0 ** -24)
            theta = 0 if u <= b0[0] else 1
            st, b = x0, b0.copy()
            pred = next((i + 1 for i in range(2) if b[i] > ins.min_confidence), 0)
            origin = 0
            for t in range(ins.verif_horizon):
                st, b, term = oracle.drace_env_step(0, e, theta, plans[pred, t - origin], st, belief=b, seed=ins.seed, k=word, t=t, step=4,
                                                    stream=5, lose_on_opp_win=False)
                if pred == 0:
                    np_ = next((i + 1 for i in range(2) if b[i] > ins.min_confidence), 0)
                    if np_:
                        pred, origin = np_, t + 1
                if term == 1:
                    fail[t] += 1
                if term:
                    break
        np.testing.assert_array_equal(got, fail)
    assert r["fail_old"].sum() > 0, "the previous plans drive into the slower opponent"
