"""Pins of the oracle's MPPI step: brute-force FP64 softmax update on small K,
weight invariants, clamp/shift semantics, and the rollout loop structure
(discount 0.97, early stop on a terminal state, terminal progress cost)."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def setup(pkg):
    sc = pkg.scenarios
    P, K = sc.quad_params(), sc.cost_cfg()
    tr = sc.track_circle(0, 8, 20.0, 1.0)
    ego = sc.start_states(tr, 1)[0]
    return sc, P, K, tr, ego


def brute_update(P, costs, noise, nominal, lam, fill):
    """Independent FP64 restatement of kernels.cu:376-403 + clamp + shift."""
    c = np.asarray(costs, dtype=np.float64)
    lam = float(np.float32(lam))  # the temperature is an FP32 field of the ABI
    w = np.exp(-(c - c.min()) / lam)
    delta = np.einsum("k,ktd->td", w, noise.astype(np.float64)) / w.sum()
    u = nominal.astype(np.float64) + delta
    u[:, 0] = np.clip(u[:, 0], P.thrust_min, P.thrust_max)
    u[:, 1:] = np.clip(u[:, 1:], -P.rate_max, P.rate_max)
    sh = np.vstack([u[1:], np.asarray(fill, dtype=np.float64)[None]])
    return u, sh, w


def test_update_matches_brute_force(oracle, setup):
    sc, P, K, tr, ego = setup
    rng = np.random.default_rng(3)
    for Kn, T, lam in [(1, 1, 1.0), (7, 3, 1.0), (64, 10, 0.3), (200, 25, 5.0)]:
        costs = rng.normal(0, 20, Kn)
        noise = rng.normal(0, 1, (Kn, T, 4)).astype(np.float32)
        nominal = (np.array(sc.hover_control(P)) + rng.normal(0, 0.5, (T, 4))).astype(np.float32)
        fill = np.array(sc.hover_control(P), dtype=np.float32)
        u, sh, w = brute_update(P, costs, noise, nominal, lam, fill)
        r64 = oracle.update(1, P, costs, noise, nominal, lam, fill)
        np.testing.assert_allclose(r64["updated"], u, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(r64["shifted"], sh, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(r64["action"], u[0], rtol=1e-12, atol=1e-12)
        assert r64["nu"] == pytest.approx(w.sum(), rel=1e-12)
        # FP32 oracle: FP32 weights and increment, within float rounding of the FP64 result
        r32 = oracle.update(0, P, costs.astype(np.float32).astype(np.float64), noise, nominal, lam, fill)
        u32, _, _ = brute_update(P, costs.astype(np.float32), noise, nominal, lam, fill)
        np.testing.assert_allclose(r32["updated"], u32, rtol=2e-6, atol=2e-6)


def test_weight_invariants(oracle, setup):
    sc, P, K, tr, ego = setup
    rng = np.random.default_rng(5)
    Kn, T = 500, 8
    costs = rng.normal(100, 30, Kn)
    noise = rng.normal(0, 1, (Kn, T, 4)).astype(np.float32)
    nominal = np.tile(np.array(sc.hover_control(P), dtype=np.float32), (T, 1))
    r = oracle.update(1, P, costs, noise, nominal, 2.0, nominal[0])
    w = r["weights"]
    assert w[np.argmin(costs)] == 1.0          # the best sample has weight exactly 1
    assert np.all(w <= 1.0) and np.all(w >= 0.0)
    assert r["nu"] == pytest.approx(w.sum())   # nu is the sum of weights (kernels.cu:405-406)
    assert r["min_cost"] == costs.min()


def test_low_temperature_selects_best_sample(oracle, setup):
    sc, P, K, tr, ego = setup
    rng = np.random.default_rng(9)
    Kn, T = 50, 6
    costs = rng.normal(0, 10, Kn)
    noise = rng.normal(0, 0.1, (Kn, T, 4)).astype(np.float32)
    nominal = np.tile(np.array(sc.hover_control(P), dtype=np.float32), (T, 1))
    r = oracle.update(1, P, costs, noise, nominal, 1e-3, nominal[0])
    best = np.argmin(costs)
    np.testing.assert_allclose(r["updated"], nominal + noise[best], atol=1e-9)


def test_zero_fill_shift_as_reference(oracle, setup):
    sc, P, K, tr, ego = setup
    T, Kn = 5, 3
    nominal = np.arange(T * 4, dtype=np.float32).reshape(T, 4) * 0.1
    noise = np.zeros((Kn, T, 4), dtype=np.float32)
    r = oracle.update(1, P, np.zeros(Kn), noise, nominal, 1.0, np.zeros(4, dtype=np.float32))
    np.testing.assert_allclose(r["shifted"][:-1], r["updated"][1:])
    assert np.all(r["shifted"][-1] == 0.0)


def test_rollout_loop_structure(oracle, setup):
    """or_rollout for one sample equals the loop of kernels.cu:100-155 composed
    from the separately pinned step / running-cost / progress primitives."""
    sc, P, K, tr, ego = setup
    T = 40
    nominal = np.tile(np.array(sc.hover_control(P), dtype=np.float32), (T, 1))
    nominal[:, 0] += 1.5
    nominal[:, 2] = 0.8
    sigma = np.array([2.0, 0.5, 0.5, 0.5], dtype=np.float32)
    costs, noise, traj = oracle.rollout(1, P, K, tr, ego, [], nominal, 3, T, 77, 4, sigma)
    x = np.array(list(ego.p) + list(ego.q) + list(ego.v) + list(ego.w), dtype=np.float64)
    for k in range(3):
        xs, g, l = x.copy(), ego.gate, ego.laps
        acc, decay = 0.0, 1.0
        for t in range(T):
            u = nominal[t].astype(np.float64) + noise[k, t].astype(np.float64)
            xs, g, l = oracle.quad_step(1, P, xs, u, track=tr, margin=K.gate_traversal_margin, gate=g, laps=l)
            c, term = oracle.running_cost(1, P, K, tr, xs, g, l, [], float(np.float32(P.dt)) * (t + 1))
            acc = acc + c * decay
            decay = decay * float(np.float32(0.97))
            if term:
                break
        acc = acc - K.final_adv_weight * oracle.progress(1, tr, xs, g, l)
        assert costs[k] == pytest.approx(acc, rel=1e-12, abs=1e-9)


def test_noise_free_samples_equal_nominal_cost(oracle, setup):
    sc, P, K, tr, ego = setup
    T = 30
    nominal = np.tile(np.array(sc.hover_control(P), dtype=np.float32), (T, 1))
    costs, noise, traj = oracle.rollout(0, P, K, tr, ego, [], nominal, 16, T, 1, 0, np.zeros(4, dtype=np.float32))
    assert np.all(noise == 0.0)
    assert np.all(costs == traj)


def test_fp32_oracle_tracks_fp64(oracle, setup):
    sc, P, K, tr, ego = setup
    T = 50
    nominal = np.tile(np.array(sc.hover_control(P), dtype=np.float32), (T, 1))
    sigma = np.array([2.0, 0.5, 0.5, 0.5], dtype=np.float32)
    c32, n32, _ = oracle.rollout(0, P, K, tr, ego, [], nominal, 512, T, 11, 0, sigma)
    c64, n64, _ = oracle.rollout(1, P, K, tr, ego, [], nominal, 512, T, 11, 0, sigma)
    assert np.array_equal(n32, n64)                 # same noise in both
    rel = np.abs(c32 - c64) / np.maximum(1.0, np.abs(c64))
    assert np.median(rel) < 1e-5
