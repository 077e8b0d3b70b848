"""Pins of the oracle's branching spline MPPI on StratRace2D (kernels.cu:34-486,
mppi_controller.cu:67-141 and 250-466 without verification).

* spline basis: SciPy's natural cubic spline through e_i; B at the knots is
  the identity, rows sum to one, linear functions are reproduced
* belief update: the Gaussian likelihood written out in NumPy; the world
  step's belief from the models' nominal actions and the applied action
* rollouts with a constant plan and sigma = 0 equal the plain StratRace2D
  rollouts of each model (same environment noise); branching bookkeeping
  equals a replay through the belief-updating world step
* reductions: aggregated costs, masked minima, nu and the temperature rule
  in closed form; a committed controller's update is the textbook MPPI
  average; linear knots interpolate, shift and act in closed form; the L2
  clamp
"""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

KNOT_SETS = [(80, list(range(0, 77, 4))), (30, [0, 3, 4, 10, 17, 25, 28]), (12, [2, 9]), (15, [0, 5, 13])]


def quiet_env(pkg, n_models=2):
    """The race environment with the opponents' action noise off (deterministic worlds)."""
    return pkg.scenarios.sr_env(n_models=n_models, action_noise=0.0)


def make(pkg, K, T, tau, sigma=0.5, min_conf=0.95, belief=(0.5, 0.5), seed=11, adapt=1, lam=10.0):
    return pkg.scenarios.bm_instance(K, T, seed, knots=tau, sigma=sigma, min_confidence=min_conf, belief=belief, adapt=adapt,
                                     lam=lam)


@pytest.mark.parametrize("T,tau", KNOT_SETS)
def test_spline_basis_is_scipy_natural_spline(oracle, T, tau):
    B64 = oracle.bm_basis(1, tau, T)
    B32 = oracle.bm_basis(0, tau, T)
    M = len(tau)
    t = np.arange(T, dtype=np.float64)
    for i in range(M):
        y = np.zeros(M)
        y[i] = 1.0
        if M == 2:
            want = np.interp(t, tau, y)
        else:
            want = CubicSpline(np.array(tau, float), y, bc_type="natural")(np.clip(t, tau[0], tau[-1]))
        np.testing.assert_allclose(B64[:, i], want, atol=1e-12)
    np.testing.assert_allclose(B32, B64, atol=2e-6)


@pytest.mark.parametrize("T,tau", KNOT_SETS)
def test_spline_basis_invariants(oracle, T, tau):
    B = oracle.bm_basis(1, tau, T)
    M = len(tau)
    np.testing.assert_allclose(B[tau, :], np.eye(M), atol=1e-13)
    np.testing.assert_allclose(B.sum(axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(B @ np.array(tau, float), np.clip(np.arange(T), tau[0], tau[-1]), atol=1e-9)


def belief_numpy(env, belief, obs, noms):
    b = np.array(belief, np.float64)
    for th in range(env.n_models):
        nom = np.array(noms[th], np.float64)
        amax = env.max_accel[1]
        sq = nom @ nom
        scale = amax / np.sqrt(sq) if sq > amax * amax else 1.0
        d = np.asarray(obs, np.float64) - nom * scale
        s2 = float(np.float32(0.2)) ** 2 + env.opp[th].action_noise ** 2  # the reference's constants are floats
        b[th] *= np.exp(-0.5 * (d @ d) / s2 - np.log(2 * float(np.float32(np.pi))) - np.log(s2))
    tot = b.sum()
    return np.full_like(b, 1.0 / len(b)) if tot < 1e-20 else b / tot


def test_update_belief_closed_form(oracle, pkg):
    env = pkg.scenarios.sr_env(n_models=3, action_noise=0.1)
    env.opp[2].action_noise = 0.4
    rng = np.random.default_rng(5)
    for case in range(40):
        b0 = rng.dirichlet(np.ones(3))
        noms = rng.normal(0, 2.0, (3, 2))  # several above the opponent's max acceleration
        obs = noms[case % 3] * 0.8 + rng.normal(0, 0.3, 2)
        want = belief_numpy(env, b0, obs, noms)
        np.testing.assert_allclose(oracle.sr_update_belief(1, env, b0, obs, noms), want, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(oracle.sr_update_belief(0, env, b0, obs, noms), want, rtol=2e-5, atol=1e-30)
    # every likelihood underflows: uniform
    far = oracle.sr_update_belief(1, env, [0.2, 0.3, 0.5], [40.0, 40.0], np.zeros((3, 2)))
    np.testing.assert_array_equal(far, np.full(3, 1.0 / 3))


def test_world_step_updates_belief_from_applied_action(oracle, pkg):
    env = quiet_env(pkg)
    st = pkg.scenarios.sr_start_state(env, start=(0.15, 0.2))
    st.vel[2], st.vel[3] = -0.6, 1.1
    b = np.array([0.5, 0.5], np.float32)
    for t in range(6):
        noms = np.array([oracle.sr_opp_action(1, env, th, st) for th in range(2)])
        obs = noms[1].copy()  # the world follows model 1; no action noise: applied = clamped nominal
        n = np.linalg.norm(obs)
        if n > env.max_accel[1]:
            obs *= env.max_accel[1] / n
        s_plain, term_plain = oracle.sr_env_step(1, env, 1, [0.4, -0.2], st, 3, 0xFFFFFFFF, t, 0, lose_on_opp_win=False)
        s_b, b_new, term = oracle.sr_env_step_belief(1, env, 1, [0.4, -0.2], st, b, 3, 0xFFFFFFFF, t, 0, lose_on_opp_win=False)
        assert term == term_plain
        assert bytes(s_b) == bytes(s_plain)
        np.testing.assert_allclose(b_new, belief_numpy(env, b, obs, noms).astype(np.float32), rtol=1e-6, atol=1e-30)
        st, b = s_b, b_new


def replay(oracle, env, ins, x0, th, k, c, step, belief):
    """Rollout of sample k under model th with the constant action c, through the world step."""
    st, b = x0, np.array(belief, np.float32)
    pred = next((i + 1 for i in range(env.n_models) if b[i] > ins.min_confidence), 0)
    time = -1 if pred == 0 else 0
    n = 0
    for t in range(ins.n_timesteps):
        st, b, term = oracle.sr_env_step_belief(1, env, th, c, st, b, ins.seed, th * ins.n_samples + k, t, step)
        n += 1
        if pred == 0:
            np_ = next((i + 1 for i in range(env.n_models) if b[i] > ins.min_confidence), 0)
            if np_:
                pred, time = np_, t + 1
        if term:
            break
    return pred, time, n


@pytest.mark.parametrize("min_conf,belief", [(2.0, (0.5, 0.5)), (0.95, (0.99, 0.01)), (0.52, (0.5, 0.5)), (0.7, (0.4, 0.6))])
def test_constant_plan_rollouts_are_plain_rollouts(oracle, pkg, min_conf, belief):
    env = pkg.scenarios.sr_env(n_models=2, action_noise=0.1)
    cost = pkg.scenarios.bm_cost()
    K, T, tau = 6, 24, [0, 5, 11, 16, 22]
    ins = make(pkg, K, T, tau, sigma=0.0, min_conf=min_conf, belief=belief)
    x0 = pkg.scenarios.sr_start_state(env, start=(0.15, 0.2))
    c = np.array([0.5, -0.25], np.float32)
    knots = np.tile(c, (3, len(tau), 1))
    r = oracle.bm_step(1, env, cost, x0, ins, np.array(belief, np.float32), 7, 10.0, knots)
    committed = next((i + 1 for i in range(2) if belief[i] > min_conf), 0)
    for th in range(2):
        if committed and committed != th + 1:
            assert np.all(np.isinf(r["ctrue"][th]))
            assert np.all(r["used"][th] == committed) and np.all(r["btime"][th] == 0) and np.all(r["len"][th] == 0)
            continue
        plain, _ = oracle.sr_rollout(1, env, cost, x0, th, np.tile(c, (T, 1)), 2 * K, T, ins.seed, 7, 0.0, want_noise=False)
        np.testing.assert_allclose(r["ctrue"][th], plain[th * K:(th + 1) * K], rtol=1e-9)
        for k in range(K):
            assert (r["used"][th, k], r["btime"][th, k], r["len"][th, k]) == replay(oracle, env, ins, x0, th, k, c, 7, belief)
    if min_conf == 0.52:
        assert np.any(r["btime"] >= 1), "some rollouts commit on the way"


def test_reductions_closed_forms(oracle, pkg):
    env = quiet_env(pkg)
    cost = pkg.scenarios.bm_cost()
    K, T, tau = 16, 20, [0, 4, 9, 14, 18]
    x0 = pkg.scenarios.sr_start_state(env, start=(0.15, 0.2))
    ins = make(pkg, K, T, tau, sigma=0.0, min_conf=2.0)
    knots = np.tile(np.float32([0.3, 0.1]), (3, len(tau), 1))
    r = oracle.bm_step(1, env, cost, x0, ins, np.float32([0.5, 0.5]), 0, 10.0, knots)
    ct = r["ctrue"]
    assert np.all(r["used"] == 0) and np.all(r["btime"] == -1)
    np.testing.assert_array_equal(r["costs"][1], ct[0])
    np.testing.assert_array_equal(r["costs"][2], ct[1])
    np.testing.assert_allclose(r["costs"][0], 0.5 * ct[0] + 0.5 * ct[1], rtol=1e-15)
    np.testing.assert_array_equal(r["mins"][0], np.full(len(tau), r["costs"][0].min()))
    assert np.all(np.isinf(r["mins"][1:])) and np.all(r["nu"][1:] == 0)
    # deterministic worlds, sigma = 0: every sample identical, every weight 1
    assert np.all(r["costs"][0] == r["costs"][0][0])
    assert r["nu"][0] == K and r["branch"] == 0
    assert r["lambda_"] == pytest.approx(10.0 * 0.9, rel=1e-7)
    # nu above the band -> 0.9; nu below -> 1.2; inside -> unchanged (K = 2^14: band [8.2, 16.4])
    ins2 = make(pkg, K, T, tau, sigma=0.0, min_conf=2.0, lam=3.0)
    ins2.adapt_lambda = 0
    assert oracle.bm_step(1, env, cost, x0, ins2, np.float32([0.5, 0.5]), 0, 3.0, knots)["lambda_"] == pytest.approx(3.0)


def test_committed_update_is_textbook_mppi(oracle, pkg):
    env = pkg.scenarios.sr_env(n_models=2, action_noise=0.1)
    cost = pkg.scenarios.bm_cost()
    K, T, tau = 96, 20, [0, 4, 9, 14, 18]
    M = len(tau)
    x0 = pkg.scenarios.sr_start_state(env, start=(0.15, 0.2))
    ins = make(pkg, K, T, tau, sigma=0.6, belief=(0.99, 0.01))
    rng = np.random.default_rng(2)
    knots = rng.normal(0, 0.4, (3, M, 2)).astype(np.float32)
    lam = 4.0
    r = oracle.bm_step(1, env, cost, x0, ins, np.float32([0.99, 0.01]), 3, lam, knots)
    assert r["branch"] == 1 and np.all(np.isinf(r["ctrue"][1]))
    c = r["ctrue"][0]
    w = np.exp(-(c - c.min()) / lam)
    eps = r["noise"][1].astype(np.float64)  # [M][K][2]
    upd = knots[1].astype(np.float64) + np.einsum("k,mkd->md", w, eps) / w.sum()
    B = oracle.bm_basis(1, tau, T)
    dense = B @ upd
    np.testing.assert_allclose(r["action"], dense[0], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(r["dense"][1, :-1], dense[1:], rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(r["dense"][1, -1], 0.0)
    np.testing.assert_allclose(r["nu"][1], float(np.float32(0.99)) * w.sum(), rtol=1e-12)
    assert np.all(np.isinf(r["mins"][[0, 2]])) and r["mins"][1].min() == c.min()
    # plans incompatible with the commitment are untouched
    np.testing.assert_array_equal(r["knots"][[0, 2]], knots[[0, 2]].astype(np.float64))


def test_linear_knots_interpolate_shift_and_act(oracle, pkg):
    env = quiet_env(pkg)
    cost = pkg.scenarios.bm_cost()
    K, T, tau = 4, 30, [0, 3, 4, 10, 17, 25, 28]
    x0 = pkg.scenarios.sr_start_state(env, start=(0.15, 0.2))
    ins = make(pkg, K, T, tau, sigma=0.0, min_conf=2.0)
    a, b = np.array([0.25, -0.5]), np.array([0.015625, 0.03125])  # exact in binary
    lin = lambda t: a + np.outer(np.clip(t, tau[0], tau[-1]), b)  # noqa: E731
    knots = np.stack([lin(np.array(tau))] * 3).astype(np.float32)
    r = oracle.bm_step(1, env, cost, x0, ins, np.float32([0.5, 0.5]), 0, 10.0, knots)
    t = np.arange(T)
    np.testing.assert_allclose(r["action"], a, atol=1e-12)
    np.testing.assert_allclose(r["dense"][0, :-1], lin(t[1:]), atol=1e-12)
    np.testing.assert_array_equal(r["dense"][0, -1], 0.0)
    for p in (1, 2):
        np.testing.assert_allclose(r["dense"][p], lin(t), atol=1e-12)
        np.testing.assert_array_equal(r["knots"][p], knots[p].astype(np.float64))
    np.testing.assert_allclose(r["knots"][0], lin(np.minimum(np.array(tau) + 1, T - 1)), atol=1e-12)


def test_knot_clamp(oracle, pkg):
    env = quiet_env(pkg)
    cost = pkg.scenarios.bm_cost()
    K, T, tau = 4, 12, [0, 5, 11]
    x0 = pkg.scenarios.sr_start_state(env, start=(0.15, 0.2))
    ins = make(pkg, K, T, tau, sigma=0.0, min_conf=2.0)
    knots = np.zeros((3, 3, 2), np.float32)
    knots[2] = [1.5, 2.0]  # norm 2.5 > max accel 2
    knots[1] = [0.0, 2.0]  # norm exactly 2: kept
    r = oracle.bm_step(1, env, cost, x0, ins, np.float32([0.5, 0.5]), 0, 10.0, knots)
    np.testing.assert_allclose(r["knots"][2], np.tile([1.2, 1.6], (3, 1)), rtol=1e-12)
    np.testing.assert_array_equal(r["knots"][1], knots[1].astype(np.float64))
    np.testing.assert_allclose(r["dense"][2], np.tile([1.5, 2.0], (T, 1)), rtol=1e-12)  # interpolated before the clamp


def test_masked_minima_and_averages_brute_force(oracle, pkg):
    env = pkg.scenarios.sr_env(n_models=2, action_noise=0.1)
    cost = pkg.scenarios.bm_cost()
    K, T, tau = 64, 24, [0, 5, 11, 16, 22]
    M = len(tau)
    x0 = pkg.scenarios.sr_start_state(env, start=(0.15, 0.2))
    belief = np.float32([0.45, 0.55])
    ins = make(pkg, K, T, tau, sigma=0.7, min_conf=0.6, belief=belief)
    knots = np.random.default_rng(4).normal(0, 0.5, (3, M, 2)).astype(np.float32)
    lam = 5.0
    r = oracle.bm_step(1, env, cost, x0, ins, belief, 2, lam, knots)
    used, bt, costs = r["used"], r["btime"], r["costs"]
    assert len(set(bt.ravel())) > 2, "rollouts commit at several times"

    def informs(b, t_local, th, k):
        """Plan b at local time t_local was followed by rollout (th, k) (state.h:203-219)."""
        origin = bt[th, k] if (b != 0 and bt[th, k] > 0) else 0
        tg = t_local + origin
        if tg >= T:
            return False
        if b != 0:
            return used[th, k] == b
        return used[th, k] == 0 or tg < bt[th, k]

    B = oracle.bm_basis(1, tau, T)
    for b in range(3):
        for m in range(M):
            coeff = np.zeros(K)
            for k in range(K):
                for th in range(2):
                    if b != 0 and b != th + 1:
                        continue
                    nb = [tau[j] for j in (m - 1, m, m + 1) if 0 <= j < M]
                    if any(informs(b, x, th, k) for x in nb):
                        coeff[k] += belief[th]
            ok = coeff > 0.01
            want_min = costs[b][ok].min() if ok.any() else np.inf
            assert r["mins"][b, m] == want_min
            if not ok.any():
                continue
            w = np.exp(-(costs[b][ok] - want_min) / lam) * coeff[ok]
            inc = (w[:, None] * r["noise"][b, m][ok]).sum(axis=0) / w.sum()
            if m == 0:
                np.testing.assert_allclose(r["nu"][b], w.sum(), rtol=1e-12)
            if b != r["branch"]:
                np.testing.assert_allclose(r["knots"][b, m], knots[b, m] + inc, rtol=1e-9, atol=1e-12)
    assert B.shape == (T, M)


def verify_replay(oracle, env, ins, x0, belief, step, plans, set_):
    """Failure counts of verifyNominalFailureKernel rebuilt sample by sample through the
    belief-updating world step (noise stream 5) and a NumPy inverse-CDF model draw."""
    T, H, NT = ins.n_timesteps, ins.verif_horizon, env.n_models
    key = [ins.seed & 0xFFFFFFFF, ins.seed >> 32]
    fail = np.zeros(T, np.int32)
    for s in range(ins.n_verif_samples):
        word = s | (set_ << 30)
        r = oracle.philox([word, 0xFFFFFFFF, step, 5], key)
        u = np.float32(((r[0] >> 8) + 1) * 2.0 ** -24)
        acc, theta = np.float32(0), NT - 1
        for q in range(NT):
            acc = np.float32(acc + np.float32(belief[q]))
            if u <= acc:
                theta = q
                break
        st, b = x0, np.array(belief, np.float32)
        pred = next((i + 1 for i in range(NT) if b[i] > ins.min_confidence), 0)
        origin = 0
        for t in range(H):
            st, b, term = oracle.sr_env_step_belief(0, env, theta, plans[pred, t - origin], st, b, ins.seed, word, t, step,
                                                    lose_on_opp_win=False, stream=5)
            if pred == 0:
                np_ = next((i + 1 for i in range(NT) if b[i] > ins.min_confidence), 0)
                if np_:
                    pred, origin = np_, t + 1
            if term == 1:
                fail[t] += 1
                break
            if term:
                break
    return fail


@pytest.mark.parametrize("start,belief", [((0.15, 0.2), (0.5, 0.5)), ((0.18, 0.2), (0.3, 0.7))])
def test_verification_counts_replay(oracle, pkg, start, belief):
    env = pkg.scenarios.sr_env(n_models=2, action_noise=0.3)
    cost = pkg.scenarios.bm_cost()
    K, T, tau = 32, 24, [0, 5, 11, 16, 22]
    ins = make(pkg, K, T, tau, sigma=0.5, min_conf=0.6, belief=belief)
    ins.n_verif_samples, ins.verif_horizon, ins.beta, ins.max_verif_eps = 48, 18, 1e-5, 0.001
    x0 = pkg.scenarios.sr_start_state(env, start=start)
    rng = np.random.default_rng(8)
    knots = rng.normal(0, 1.2, (3, len(tau), 2)).astype(np.float32)
    prev = rng.normal(0, 1.5, (3, T, 2)).astype(np.float32)
    b = np.float32(belief)
    r = oracle.bm_step(0, env, cost, x0, ins, b, 6, 10.0, knots, prev_dense=prev)
    assert np.array_equal(r["fail_new"], verify_replay(oracle, env, ins, x0, b, 6, r["dense_new"].astype(np.float32), 0))
    assert np.array_equal(r["fail_old"], verify_replay(oracle, env, ins, x0, b, 6, prev, 1))
    assert np.all(r["fail_new"][ins.verif_horizon:] == 0)
    # the executed plans with the new plan chosen: dense_new with the chosen branch shifted left, zero fill
    want = r["dense_new"].copy()
    want[r["branch"], :-1] = r["dense_new"][r["branch"], 1:]
    want[r["branch"], -1] = 0.0
    np.testing.assert_array_equal(r["dense"], want)


def test_verification_closed_forms(oracle, pkg):
    """Deterministic worlds: a plan that leaves the track fails every sample at the same step;
    a standing plan fails none; the counts do not depend on the plan set's noise draw."""
    env = quiet_env(pkg)
    cost = pkg.scenarios.bm_cost()
    K, T, tau = 8, 30, [0, 7, 14, 21, 28]
    ins = make(pkg, K, T, tau, sigma=0.0, min_conf=2.0)
    ins.n_verif_samples, ins.verif_horizon, ins.beta, ins.max_verif_eps = 64, 30, 1e-5, 0.001
    x0 = pkg.scenarios.sr_start_state(env, start=(0.15, 0.2))
    out = np.zeros((3, T, 2), np.float32)
    tang = np.array(env.t_grid[77][:])
    normal = np.array([-tang[1], tang[0]], np.float32)
    out[:] = 2.0 * normal  # full acceleration across the track
    r = oracle.bm_step(0, env, cost, x0, ins, np.float32([0.5, 0.5]), 0, 10.0, np.zeros((3, len(tau), 2), np.float32), prev_dense=out)
    assert r["fail_old"].sum() == 64 and np.count_nonzero(r["fail_old"]) == 1
    assert r["fail_new"].sum() == 0  # zero knots, sigma 0: the ego stands still, nothing fails within 30 steps
