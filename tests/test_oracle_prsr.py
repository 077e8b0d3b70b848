"""Pins of the oracle's PR-MPPI on StratRace2D (prmppi_kernels.cu on the
reference's environment, costs.cuh:44-86).

* one-step rollouts: pair cost = PRMPPIstateCost + finalCost and safety =
  r*coll - trackBoundaryDist<false>, computed in NumPy from the state that
  the (pinned) environment step produces with the same noise
* aggregation, minima and the three weighted averages: FP64 NumPy restatement
* sigma = 0: candidate 1 is the shifted plan
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def race(pkg):
    return pkg.scenarios.pr_sr_race(n_samples=48, n_timesteps=20)


def _bd(e, st, lose):
    if lose and st.laps[1] >= e.n_win_laps:
        return -1.0
    d = np.hypot(np.float32(st.pos[2]) - np.float32(st.pos[0]), np.float32(st.pos[3]) - np.float32(st.pos[1]))
    return min(d - e.drone_radius, e.track_width - abs(st.lat_dist[0]))


def test_one_step_costs_closed_form(oracle, race, pkg):
    e, K = race.env, race.cost
    sc = pkg.scenarios
    for sampled in (0, 1):
        ins = sc.prsr_instance(6, 1, 77, [0.5, 0.5], sampled=sampled, n_draws=3)
        st = race.states[0]
        r = oracle.prsr_step(1, e, K, st, ins, np.array([0.5, 0.5], np.float32), 4, [10, 10, 10],
                             np.array([[0.7, -0.3]], np.float32), np.array([[-0.2, 0.4]], np.float32))
        M = r["models"].size
        for plan, base in ((0, [0.7, -0.3]), (1, [-0.2, 0.4])):
            for m in range(M):
                for k in range(6):
                    u = np.float32(base) + r["noise"][k, 0]
                    s1, term = oracle.sr_env_step(1, e, int(r["models"][m]), u, st, ins.seed, m * 6 + k, 0, 4)
                    bd = _bd(e, s1, True)
                    c = 0.0
                    if bd < K.boundary_threshold_factor * e.drone_radius:
                        c += K.boundary_cost * (bd / e.drone_radius) ** 2
                    if s1.laps[0] >= e.n_win_laps or abs(s1.lat_dist[1]) > e.track_width - e.drone_radius:
                        c -= K.win_cost
                    if s1.laps[1] >= e.n_win_laps:
                        c += K.opp_win_cost
                    c += -K.final_adv_weight * (s1.s[0] / e.track_length + s1.laps[0])
                    safety = e.drone_radius * K.coll_dist_factor - _bd(e, s1, False)
                    assert r["pairs"][plan, m, k, 0] == pytest.approx(c, rel=1e-6, abs=1e-6)
                    assert r["pairs"][plan, m, k, 1] == pytest.approx(safety, rel=1e-6, abs=1e-6)


@pytest.mark.parametrize("sampled", [0, 1])
def test_reduce_matches_brute_force(oracle, race, pkg, sampled):
    e, K = race.env, race.cost
    ins = pkg.scenarios.prsr_instance(48, 20, 31 + sampled, [0.3, 0.7], sampled=sampled, n_draws=5, safety_weight=50.0)
    belief = np.array([0.3, 0.7], np.float32)
    rng = np.random.default_rng(sampled)
    nom = rng.normal(0, 0.8, (20, 2)).astype(np.float32)
    rob = rng.normal(0, 0.8, (20, 2)).astype(np.float32)
    lam = [10.0, 5.0, 1.0]
    r = oracle.prsr_step(1, e, K, race.states[0], ins, belief, 2, lam, nom, rob)
    c, sf = r["pairs"][..., 0], r["pairs"][..., 1]
    avg = c.mean(axis=1) if sampled else np.einsum("m,pmk->pk", belief.astype(np.float64), c)
    mx = sf.max(axis=1)
    full = avg + np.where(mx > 0, ins.safety_weight, 0.0)
    np.testing.assert_allclose(r["agg"][:, 0], full[0], rtol=1e-12)
    np.testing.assert_allclose(r["agg"][:, 2], full[1], rtol=1e-12)
    for j, (cost, base) in enumerate([(full[0], nom), (full[1], rob), (mx[1], rob)]):
        w = np.exp(-(cost - cost.min()) / float(np.float32(lam[j])))
        u = base.astype(np.float64)[None] + r["noise"].astype(np.float64)
        np.testing.assert_allclose(r["cand"][j], np.einsum("k,ktd->td", w, u) / w.sum(), rtol=1e-11, atol=1e-11)
        assert r["nu"][j] == pytest.approx(w.sum(), rel=1e-12)


def test_zero_noise_candidate_is_plan(oracle, race, pkg):
    e, K = race.env, race.cost
    ins = pkg.scenarios.prsr_instance(8, 12, 5, [0.5, 0.5], sigma=0.0)
    nom = np.full((12, 2), 0.3, np.float32)
    rob = np.full((12, 2), -0.2, np.float32)
    r = oracle.prsr_step(0, e, K, race.states[0], ins, np.array([0.5, 0.5], np.float32), 0, [10, 10, 10], nom, rob)
    np.testing.assert_array_equal(r["cand"][0], nom.astype(np.float64))
    np.testing.assert_array_equal(r["cand"][1], rob.astype(np.float64))
