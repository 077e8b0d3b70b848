"""Pins of the oracle's PR-MPPI step (prmppi_kernels.cu / prmppi_controller.cu
restated for the quadrotor race).

* rollouts: without a collision the PR running cost equals the (pinned) plain
  MPPI running cost, so pair costs equal or_rollout's costs; the safety of a
  one-step rollout equals r*coll - bd computed in NumPy from the stepped state
  and the shifted-velocity opponent prediction
* aggregation, minima, three weighted averages: an independent FP64 NumPy
  restatement (belief-weighted or uniform model mean, safety penalty,
  exp(-(c - min)/lambda), sum w(u + eps)/sum w) on small K; the PR formula
  equals the plain MPPI formula u + sum w eps / sum w (the pinned or_mppi_update)
* choice: model-averaged candidate costs, pick, use_nom, reset_nom from the
  returned per-model costs; with sigma = 0 every sample is the plan itself,
  so candidate 1 is the shifted plan and its candidate cost is the pair cost
* model draws: inverse CDF of the belief at u from NumPy's Philox; frequencies
* temperature adaptation: thresholds 0.001 K / 0.0005 K with the 0.01 / 100 clamps
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def world(pkg):
    sc = pkg.scenarios
    return sc, sc.c3_pr_four_drones(seed=2, n_samples=96, n_timesteps=20)


def _belief(ins):
    return np.array(ins.belief[:ins.n_branches], dtype=np.float32)


def _plans(sc, ins, rng=None, scale=0.0):
    T = ins.n_timesteps
    hov = np.tile(np.array(sc.hover_control(), dtype=np.float32), (T, 1))
    if rng is None:
        return hov, hov.copy()
    return ((hov + rng.normal(0, scale, hov.shape)).astype(np.float32),
            (hov + rng.normal(0, scale, hov.shape)).astype(np.float32))


def brute_reduce(ins, belief, pairs, noise, nom, rob, lam):
    """FP64 restatement of prmppi_kernels.cu:134-282 from the pair costs."""
    M = pairs.shape[1]
    c, sf = pairs[..., 0], pairs[..., 1]                      # [plan][m][k]
    if ins.sampled:
        avg = c.mean(axis=1)
    else:
        avg = np.einsum("m,pmk->pk", belief.astype(np.float64), c)
    mx = sf.max(axis=1)
    full = avg + np.where(mx > 0, ins.safety_weight, 0.0)
    sets = [(full[0], nom), (full[1], rob), (mx[1], rob)]
    cands, mins, nus = [], [], []
    for j, (cost, base) in enumerate(sets):
        w = np.exp(-(cost - cost.min()) / float(np.float32(lam[j])))
        u = base.astype(np.float64)[None] + noise.astype(np.float64)
        cands.append(np.einsum("k,ktd->td", w, u) / w.sum())
        mins.append(cost.min())
        nus.append(w.sum())
    assert M == pairs.shape[1]
    return np.array(cands), np.array(mins), np.array(nus), full, mx


@pytest.mark.parametrize("sampled", [0, 1])
def test_reduce_matches_brute_force(oracle, world, pkg, sampled):
    sc, w = world
    rng = np.random.default_rng(11 + sampled)
    ins = sc.pr_instance(64, 12, 0, 1, 777, _belief(w.instances[1]), np.asarray(w.instances[1].vel_shift)[:, :3],
                         sampled=sampled, n_models=5)
    belief = _belief(ins)
    nom, rob = _plans(sc, ins, rng, 0.5)
    lam = [1.0, 0.5, 0.2]
    opps = oracle.world_opponents(w, w.drones, 1)
    r = oracle.pr_step(1, w.params, w.cost, w.tracks[0], w.drones[1], opps, ins, belief, 3, lam, nom, rob)
    cands, mins, nus, full, mx = brute_reduce(ins, belief, r["pairs"], r["noise"], nom, rob, lam)
    np.testing.assert_allclose(r["agg"][:, 0], full[0], rtol=1e-12)
    np.testing.assert_allclose(r["agg"][:, 2], full[1], rtol=1e-12)
    np.testing.assert_allclose(r["agg"][:, 3], mx[1], rtol=1e-12)
    np.testing.assert_allclose(r["min"], mins, rtol=1e-12)
    np.testing.assert_allclose(r["nu"], nus, rtol=1e-12)
    np.testing.assert_allclose(r["cand"], cands, rtol=1e-11, atol=1e-11)
    # FP32 oracle stays within float rounding of the FP64 restatement of its own costs
    r32 = oracle.pr_step(0, w.params, w.cost, w.tracks[0], w.drones[1], opps, ins, belief, 3, lam, nom, rob)
    c32, m32, n32, _, _ = brute_reduce(ins, belief, r32["pairs"], r32["noise"], nom, rob, lam)
    np.testing.assert_allclose(r32["cand"], c32, rtol=3e-6, atol=3e-6)
    np.testing.assert_array_equal(r32["min"], m32.astype(np.float32))


def test_pr_formula_equals_plain_update(oracle, world):
    """sum w (u + eps) / sum w = u + sum w eps / sum w (the plain MPPI formula)."""
    sc, w = world
    ins = w.instances[0]
    rng = np.random.default_rng(4)
    nom, rob = _plans(sc, ins, rng, 0.3)
    lam = [1.0, 1.0, 1.0]
    opps = oracle.world_opponents(w, w.drones, 0)
    r = oracle.pr_step(1, w.params, w.cost, w.tracks[0], w.drones[0], opps, ins, _belief(ins), 0, lam, nom, rob)
    plain = oracle.update(1, w.params, r["agg"][:, 0], r["noise"], nom, 1.0, nom[0])
    # the plain update also clamps; compare where no clamp acts
    P = w.params
    u = r["cand"][0]
    inside = (u[:, 0] > P.thrust_min) & (u[:, 0] < P.thrust_max) & np.all(np.abs(u[:, 1:]) < P.rate_max, axis=1)
    assert inside.sum() > 10
    np.testing.assert_allclose(plain["updated"][inside], u[inside], rtol=1e-12, atol=1e-12)
    assert plain["nu"] == pytest.approx(r["nu"][0], rel=1e-12)


def test_zero_noise_reduces_to_plan_rollout(oracle, world, pkg):
    """sigma = 0: all samples equal the shifted plan, so candidate 1 is the plan and
    the candidate costs are the pair costs (+ the safety penalty when unsafe)."""
    sc, w = world
    src = w.instances[2]
    ins = sc.pr_instance(8, 15, 0, 2, 99, _belief(src), np.asarray(src.vel_shift)[:, :3], sigma=(0, 0, 0, 0))
    rng = np.random.default_rng(8)
    nom, rob = _plans(sc, ins, rng, 1.0)
    opps = oracle.world_opponents(w, w.drones, 2)
    for f64 in (0, 1):
        r = oracle.pr_step(f64, w.params, w.cost, w.tracks[0], w.drones[2], opps, ins, _belief(ins), 0, [1, 1, 1], nom, rob)
        assert np.all(r["noise"] == 0)
        np.testing.assert_array_equal(r["cand"][0], nom.astype(np.float64))
        np.testing.assert_array_equal(r["cand"][1], rob.astype(np.float64))
        assert np.all(r["pairs"] == r["pairs"][:, :, :1, :])
        c = oracle.pr_choose(f64, w.params, w.cost, w.tracks[0], w.drones[2], opps, ins, _belief(ins), r["models"], r["cand"],
                             [1, 1, 1], r["nu"])
        pc = r["pairs"][0, :, 0, 0]
        ps = r["pairs"][0, :, 0, 1]
        np.testing.assert_array_equal(c["cc"][0], np.where(ps > 0, pc + ins.safety_weight, pc))
        np.testing.assert_array_equal(c["cs"][0], ps)


def test_pair_costs_equal_plain_rollout_without_collision(oracle, world, pkg):
    sc, w = world
    ego_i = 0
    src = w.instances[ego_i]
    B = src.n_branches
    zero = np.zeros((B, 3, 3), dtype=np.float32)
    ins = sc.pr_instance(128, 20, 0, ego_i, 4242, _belief(src), zero)
    nom, rob = _plans(sc, ins)
    opps = oracle.world_opponents(w, w.drones, ego_i)
    r = oracle.pr_step(0, w.params, w.cost, w.tracks[0], w.drones[ego_i], opps, ins, _belief(ins), 5, [1, 1, 1], nom, rob)
    costs, noise, _ = oracle.rollout(0, w.params, w.cost, w.tracks[0], w.drones[ego_i], opps, nom, 128, 20, 4242, 5,
                                     ins.sigma[:])
    np.testing.assert_array_equal(r["noise"], noise)
    safe = r["pairs"][0, 0, :, 1]
    ok = safe <= 0
    assert ok.sum() > 100
    for m in range(B):
        np.testing.assert_array_equal(r["pairs"][0, m, ok, 0], costs[ok])


def test_one_step_safety_closed_form(oracle, world, pkg):
    sc, w = world
    ego_i = 1
    src = w.instances[ego_i]
    ins = sc.pr_instance(4, 1, 0, ego_i, 5, _belief(src), np.asarray(src.vel_shift)[:, :3])
    nom, rob = _plans(sc, ins)
    opps = oracle.world_opponents(w, w.drones, ego_i)
    r = oracle.pr_step(1, w.params, w.cost, w.tracks[0], w.drones[ego_i], opps, ins, _belief(ins), 0, [1, 1, 1], nom, rob)
    P, Kc = w.params, w.cost
    s = w.drones[ego_i]
    x0 = np.array(list(s.p) + list(s.q) + list(s.v) + list(s.w))
    shifts = np.asarray(ins.vel_shift)
    for k in range(4):
        u = nom[0].astype(np.float64) + r["noise"][k, 0].astype(np.float64)
        x, _, _ = oracle.quad_step(1, P, x0, u, w.tracks[0], Kc.gate_traversal_margin, s.gate, s.laps)
        for m in range(ins.n_branches):
            d = []
            for j, o in enumerate(opps):
                v = np.array(o.v[:], dtype=np.float32) + shifts[m, j].astype(np.float32)
                pred = np.array(o.p[:], dtype=np.float64) + v.astype(np.float64) * float(np.float32(P.dt))
                d.append(np.linalg.norm(x[:3] - pred))
            bd = min(min(d) - P.drone_radius, x[2] - w.tracks[0].ground_z)
            safety = float(np.float32(P.drone_radius) * np.float32(Kc.coll_dist_factor)) - bd
            assert r["pairs"][0, m, k, 1] == pytest.approx(safety, abs=1e-9)


def test_model_draws_inverse_cdf(oracle, world, pkg):
    sc, w = world
    belief = np.array([0.1, 0.6, 0.3], dtype=np.float32)
    ins = sc.pr_instance(2, 2, 0, 0, 0x1234_5678_9ABC_DEF0, belief, np.zeros((3, 3, 3), np.float32), sampled=1, n_models=16)
    opps = oracle.world_opponents(w, w.drones, 0)
    nom, rob = _plans(sc, ins)
    key = (ins.seed & 0xFFFFFFFF, ins.seed >> 32)
    counts = np.zeros(3)
    for step in range(40):
        r = oracle.pr_step(0, w.params, w.cost, w.tracks[0], w.drones[0], opps, ins, belief, step, [1, 1, 1], nom, rob,
                           want_noise=False, want_pairs=False)
        ctr = np.zeros((16, 4), dtype=np.uint32)
        ctr[:, 0] = np.arange(16)
        ctr[:, 2] = step
        ctr[:, 3] = 3
        words = pkg.philox.philox4x32_10(ctr, key)[:, 0]
        u = ((words >> np.uint32(8)).astype(np.float64) + 1.0) * 2.0 ** -24
        cdf = np.cumsum(belief.astype(np.float32), dtype=np.float32)
        want = np.array([next((i for i in range(3) if ui <= cdf[i]), 2) for ui in u])
        np.testing.assert_array_equal(r["models"], want)
        counts += np.bincount(r["models"], minlength=3)
    freq = counts / counts.sum()
    np.testing.assert_allclose(freq, belief, atol=0.06)


def test_choice_and_temperature_rules(oracle, world, pkg):
    sc, w = world
    for sampled in (0, 1):
        src = w.instances[3]
        ins = sc.pr_instance(64, 16, 0, 3, 31 + sampled, _belief(src), np.asarray(src.vel_shift)[:, :3], sampled=sampled,
                             n_models=4, adapt=1)
        opps = oracle.world_opponents(w, w.drones, 3)
        rng = np.random.default_rng(sampled)
        nom, rob = _plans(sc, ins, rng, 2.0)
        r = oracle.pr_step(0, w.params, w.cost, w.tracks[0], w.drones[3], opps, ins, _belief(ins), 0, [1, 1, 1], nom, rob)
        for nu in ([1.0, 0.01, 0.05], [0.02, 0.064, 0.032]):
            lam = np.array([0.5, 0.011, 99.0], dtype=np.float32)
            c = oracle.pr_choose(0, w.params, w.cost, w.tracks[0], w.drones[3], opps, ins, _belief(ins), r["models"], r["cand"],
                                 lam, nu)
            cc, cs = c["cc"], c["cs"]
            if sampled:
                avg = cc.astype(np.float32).sum(axis=1, dtype=np.float32) / np.float32(cc.shape[1])
            else:
                avg = np.array([np.float32(0)] * 2)
                for which in range(2):
                    a = np.float32(0)
                    for m in range(cc.shape[1]):
                        a = np.float32(np.float64(_belief(ins)[m]) * np.float64(np.float32(cc[which, m])) + np.float64(a))
                    avg[which] = a
            np.testing.assert_allclose(c["cand_cost"], avg, rtol=1e-6)
            pick = 0 if c["cand_cost"][0] <= c["cand_cost"][1] else 1
            assert c["reset_nom"] == int(c["cand_cost"][1] < c["cand_cost"][0])
            assert c["use_nom"] == int(not np.any(cs[pick] > 0))
            plan = r["cand"][pick] if c["use_nom"] else r["cand"][2]
            np.testing.assert_array_equal(c["action"], plan[0])
            K = ins.n_samples
            want = []
            for j in range(3):
                v, l = np.float32(nu[j]), lam[j]
                if v > np.float32(0.001) * np.float32(K) and l > np.float32(0.01):
                    l = np.float32(l * np.float32(0.9))
                elif v < np.float32(0.0005) * np.float32(K) and l < np.float32(100.0):
                    l = np.float32(l * np.float32(1.2))
                want.append(l)
            np.testing.assert_array_equal(np.float32(c["lam"]), np.array(want, dtype=np.float32))


def test_fp32_close_to_fp64(oracle, world):
    sc, w = world
    ins = w.instances[0]
    nom, rob = _plans(sc, ins)
    opps = oracle.world_opponents(w, w.drones, 0)
    b = _belief(ins)
    r32 = oracle.pr_step(0, w.params, w.cost, w.tracks[0], w.drones[0], opps, ins, b, 0, [1, 1, 1], nom, rob)
    r64 = oracle.pr_step(1, w.params, w.cost, w.tracks[0], w.drones[0], opps, ins, b, 0, [1, 1, 1], nom, rob)
    ok = (r64["pairs"][..., 1] < -0.05).all(axis=(0, 1))
    np.testing.assert_allclose(r32["pairs"][:, :, ok, 0], r64["pairs"][:, :, ok, 0], rtol=1e-4)
    np.testing.assert_allclose(r32["cand"], r64["cand"], rtol=1e-4, atol=1e-4)
