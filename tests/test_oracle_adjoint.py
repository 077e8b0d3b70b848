"""Pins of the oracle's reverse-mode rollout gradient: central finite
differences (h = 1e-3, FP64) on coordinates whose discrete decisions (clamp
activity, gate crossings, stop step, active cost branches) do not change
within +-h; invariance to the checkpoint interval; J equal to the
noise-free rollout cost; FP32 adjoint close to FP64."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def world(pkg):
    sc = pkg.scenarios
    scn = sc.c1_four_drones()
    ego = scn.drones[0]
    opp = []
    for d in range(1, 4):
        o = pkg.abi.Opponent()
        for c in range(3):
            o.p[c] = scn.drones[d].p[c]
            o.v[c] = 1.5 * (c == 0)
        opp.append(o)
    return scn, ego, opp


def sequences(pkg, n, T, seed):
    rng = np.random.default_rng(seed)
    hover = np.array(pkg.scenarios.hover_control(), np.float32)
    sigma = np.array([2.0, 0.5, 0.5, 0.5], np.float32)
    return (hover + sigma * rng.standard_normal((n, T, 4))).astype(np.float32)


def test_gradient_matches_central_differences(oracle, pkg, world):
    scn, ego, opp = world
    T, h = 40, 1e-3
    U = sequences(pkg, 6, T, 1)
    J, G, sig = oracle.cost_grad(1, scn.params, scn.cost, scn.tracks[0], ego, opp, U, ck=10)
    n, checked, kinks = U.shape[0], 0, 0
    for i in range(n):
        Up = np.repeat(U[i][None], T * 4, axis=0).astype(np.float64)
        Um = Up.copy()
        idx = np.arange(T * 4)
        Up.reshape(T * 4, T * 4)[idx, idx] += h
        Um.reshape(T * 4, T * 4)[idx, idx] -= h
        # perturbations are applied in float32 (the ABI type): use the exact step taken
        Up32, Um32 = Up.astype(np.float32), Um.astype(np.float32)
        hp = (Up32.reshape(T * 4, T * 4)[idx, idx].astype(np.float64) - U[i].reshape(-1))
        hm = (U[i].reshape(-1) - Um32.reshape(T * 4, T * 4)[idx, idx].astype(np.float64))
        Jp, _, sp = oracle.cost_grad(1, scn.params, scn.cost, scn.tracks[0], ego, opp, Up32, grad=False)
        Jm, _, sm = oracle.cost_grad(1, scn.params, scn.cost, scn.tracks[0], ego, opp, Um32, grad=False)
        ok = (sp == sig[i]) & (sm == sig[i])
        fd = (Jp - Jm) / (hp + hm)
        g = G[i].reshape(-1)
        kinks += int((~ok).sum())
        scale = max(1.0, np.abs(g).max())
        err = np.abs(fd[ok] - g[ok]) / scale
        assert err.max() < 1e-5, (i, err.max())
        checked += int(ok.sum())
    assert checked > 0.8 * n * T * 4, (checked, kinks)


@pytest.mark.parametrize("ck", [1, 3, 7, 40, 100])
def test_checkpoint_interval_does_not_change_the_gradient(oracle, pkg, world, ck):
    scn, ego, opp = world
    U = sequences(pkg, 4, 40, 2)
    J0, G0, _ = oracle.cost_grad(1, scn.params, scn.cost, scn.tracks[0], ego, opp, U, ck=10)
    J1, G1, _ = oracle.cost_grad(1, scn.params, scn.cost, scn.tracks[0], ego, opp, U, ck=ck)
    assert np.array_equal(J0, J1)
    assert np.array_equal(G0, G1)


def test_cost_equals_noise_free_rollout(oracle, pkg, world):
    scn, ego, opp = world
    U = sequences(pkg, 3, 50, 3)
    for f64 in (0, 1):
        J, _, _ = oracle.cost_grad(f64, scn.params, scn.cost, scn.tracks[0], ego, opp, U, grad=False)
        for i in range(3):
            _, _, traj = oracle.rollout(f64, scn.params, scn.cost, scn.tracks[0], ego, opp, U[i], 1, 50, 1, 0,
                                        np.zeros(4, np.float32), want_noise=False)
            assert J[i] == traj


def test_fp32_adjoint_tracks_fp64(oracle, pkg, world):
    scn, ego, opp = world
    U = sequences(pkg, 16, 100, 4)
    J32, G32, s32 = oracle.cost_grad(0, scn.params, scn.cost, scn.tracks[0], ego, opp, U)
    J64, G64, s64 = oracle.cost_grad(1, scn.params, scn.cost, scn.tracks[0], ego, opp, U)
    same = s32 == s64
    assert same.mean() > 0.7
    for i in np.where(same)[0]:
        scale = max(1.0, np.abs(G64[i]).max())
        assert np.abs(G32[i] - G64[i]).max() / scale < 1e-3
