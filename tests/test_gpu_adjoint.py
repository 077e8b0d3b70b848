"""GPU adjoint rollout (dr_engine_cost_grad): J bit-exact with the FP32 oracle,
dJ/dU against the FP32 and FP64 oracle adjoints and against central finite
differences (h = 1e-3) of the FP64 oracle on 256 random control sequences
(hover + N(0, sigma)); coordinates whose discrete decisions change within
+-h are excluded and counted.  Also: checkpoint invariance, the autograd
wrapper, and the adjoint-refined control step."""
import sys
from types import SimpleNamespace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def setup(pkg, T=100, n_drones=4):
    sc = pkg.scenarios
    tr = sc.track_lemniscate(1, 12, 30.0, 1.5, 0.3)
    drones = sc.start_states(tr, n_drones)
    for j in range(1, n_drones):
        drones[j].v[0] = 1.5
    inst = [sc.instance(256, T, 0, 0, 11)]
    scn = sc.Scenario("adj", [tr], inst, [0] * n_drones, drones, 1, "adjoint")
    opp = []
    for d in range(1, n_drones):
        o = pkg.abi.Opponent()
        for c in range(3):
            o.p[c], o.v[c] = drones[d].p[c], drones[d].v[c]
        opp.append(o)
    return scn, opp


def sequences(pkg, n, T, seed):
    rng = np.random.default_rng(seed)
    hover = np.array(pkg.scenarios.hover_control(), np.float32)
    sigma = np.array([2.0, 0.5, 0.5, 0.5], np.float32)
    return (hover + sigma * rng.standard_normal((n, T, 4))).astype(np.float32)


def test_gradient_against_oracles_and_fd(pkg, oracle):
    say("case: adjoint, 256 sequences, T=100, checkpoints every 10")
    T = 100
    scn, opp = setup(pkg, T)
    eng = pkg.Engine(scn)
    U = sequences(pkg, 256, T, 7)
    J, G = eng.cost_grad(0, U, ckpt_every=10)
    J32, G32, s32 = oracle.cost_grad(0, scn.params, scn.cost, scn.tracks[0], scn.drones[0], opp, U, ck=10, threads=16)
    J64, G64, s64 = oracle.cost_grad(1, scn.params, scn.cost, scn.tracks[0], scn.drones[0], opp, U, ck=10, threads=16)
    assert np.array_equal(J.astype(np.float64), J32), "J must be bit-exact with the FP32 oracle"
    scale = np.maximum(1.0, np.abs(G64).reshape(256, -1).max(axis=1))
    e32 = (np.abs(G - G32).reshape(256, -1).max(axis=1) / scale)
    same = s32 == s64
    e64 = (np.abs(G - G64).reshape(256, -1).max(axis=1) / scale)[same]
    say(f"  dJ/dU vs FP32 oracle: max rel {e32.max():.2e}; vs FP64 oracle ({same.sum()} sequences with equal decisions): "
        f"max rel {e64.max():.2e}, median {np.median(e64):.2e}")
    assert e32.max() < 1e-5
    assert np.median(e64) < 1e-4
    # central differences of the FP64 oracle
    h = 1e-3
    checked = kinks = 0
    worst = 0.0
    idx = np.arange(T * 4)
    for i in range(256):
        base = np.repeat(U[i].reshape(1, -1), T * 4, axis=0).astype(np.float64)
        up, um = base.copy(), base.copy()
        up[idx, idx] += h
        um[idx, idx] -= h
        up32, um32 = up.astype(np.float32), um.astype(np.float32)
        step = up32[idx, idx].astype(np.float64) - um32[idx, idx].astype(np.float64)
        Jp, _, sp = oracle.cost_grad(1, scn.params, scn.cost, scn.tracks[0], scn.drones[0], opp, up32.reshape(-1, T, 4),
                                     grad=False, threads=16)
        Jm, _, sm = oracle.cost_grad(1, scn.params, scn.cost, scn.tracks[0], scn.drones[0], opp, um32.reshape(-1, T, 4),
                                     grad=False, threads=16)
        ok = (sp == s64[i]) & (sm == s64[i])
        fd = (Jp - Jm) / step
        err = np.abs(fd[ok] - G[i].reshape(-1)[ok].astype(np.float64)) / scale[i]
        if err.size:
            worst = max(worst, float(err.max()))
        checked += int(ok.sum())
        kinks += int((~ok).sum())
    say(f"  central differences h=1e-3: {checked} coordinates checked, {kinks} excluded (decision change within +-h), "
        f"max rel error {worst:.2e}")
    assert checked > 0.5 * 256 * T * 4
    assert worst < 1e-3
    eng.close()


def test_checkpoint_interval_invariance(pkg):
    T = 60
    scn, opp = setup(pkg, T)
    eng = pkg.Engine(scn)
    U = sequences(pkg, 40, T, 3)
    ref = eng.cost_grad(0, U, ckpt_every=10)
    for ck in (1, 4, 7, 60):
        J, G = eng.cost_grad(0, U, ckpt_every=ck)
        assert np.array_equal(J, ref[0]) and np.array_equal(G, ref[1]), ck
    eng.close()


def test_autograd_wrapper(pkg):
    import torch
    sys.path.insert(0, "")
    from drone_race_amd.autograd import rollout_cost
    T = 50
    scn, opp = setup(pkg, T)
    eng = pkg.Engine(scn)
    U = torch.tensor(sequences(pkg, 8, T, 5), device="cuda:0", requires_grad=True)
    J = rollout_cost(U, eng, 0, 10)
    (J * torch.arange(1, 9, device="cuda:0", dtype=torch.float32)).sum().backward()
    Jh, Gh = eng.cost_grad(0, U.detach().cpu().numpy(), ckpt_every=10)
    assert np.array_equal(J.detach().cpu().numpy(), Jh)
    expect = Gh * np.arange(1, 9, dtype=np.float32)[:, None, None]
    assert np.allclose(U.grad.cpu().numpy(), expect, rtol=0, atol=0)
    eng.close()


def refine_oracle(oracle, scn, opp, U, n_steps, step, ck=10):
    """Iterates U_0 = U, U_k+1 = clamp(U_k - step dJ/dU(U_k)) and their J (FP32
    oracle adjoint).  The engine keeps the iterate of lowest J, U_0 (the plain
    MPPI update) on ties."""
    P = scn.params
    lo = np.array([P.thrust_min, -P.rate_max, -P.rate_max, -P.rate_max], np.float32)
    hi = np.array([P.thrust_max, P.rate_max, P.rate_max, P.rate_max], np.float32)
    Us, Js = [U], []
    for k in range(n_steps + 1):
        J, G, _ = oracle.cost_grad(0, P, scn.cost, scn.tracks[0], scn.drones[0], opp, Us[-1][None], ck=ck, grad=k < n_steps)
        Js.append(float(J[0]))
        if k < n_steps:
            Us.append(np.clip((Us[-1].astype(np.float64) - np.float64(np.float32(step)) * G[0]).astype(np.float32), lo, hi))
    return Us, np.array(Js)


def check_refined(eng, act, Us, Js, tol=1e-5):
    """The engine's shifted nominal and action against the iterate of lowest J
    (within tol relative: the forward recurrences run as prefix scans, the
    same arithmetic in another order, and each gradient step multiplies the
    gradient's rounding by the step size); iterates whose J is within rounding
    (1e-5 relative) of the lowest are all acceptable.  Returns (lowest,
    matched)."""
    best = int(np.argmin(Js))
    tie = 1e-5 * max(1.0, abs(Js[best]))
    ok = [k for k in range(len(Js)) if Js[k] <= Js[best] + tie]
    got = eng.get_nominal(0)[:-1]
    errs = {k: np.max(np.abs(got - Us[k][1:]), initial=0.0) / max(1.0, np.abs(Us[k]).max()) for k in ok}  # T = 1: action only
    k = min(errs, key=errs.get)
    say(f"  J of the iterates {np.array2string(Js, precision=4)}; lowest {best}; engine kept {k} (max rel err {errs[k]:.2e})")
    assert errs[k] < tol
    assert np.abs(act[0] - Us[k][0]).max() / max(1.0, np.abs(Us[k][0]).max()) < tol
    return best, k


@pytest.mark.parametrize("step,kept", [(0.05, 3), (5.0, 0), (50.0, 2)])
def test_refined_step_matches_oracle(pkg, oracle, step, kept):
    """MPPI update, then 3 projected gradient steps (checkpoints every 10), the
    lowest-J iterate kept, then action and shift.  At step 0.05 every gradient
    step lowers J; at step 5 all three overshoot and the plain update is kept;
    at step 50 the second iterate is the lowest (J of the iterates, FP32
    oracle: [-88.7, -103.0, -107.4, -109.0], [-88.7, 159.7, 135.4, 109.1],
    [-88.7, 77.8, -92.4, 72.0]).  Plans within 1e-5 relative of the oracle's
    iterate up to step 5; at step 50 the two gradient steps multiply the
    gradient's rounding 50-fold, and the bar is 1e-4."""
    say(f"case: adjoint-refined control step, step {step}")
    sc = pkg.scenarios
    T, K = 60, 2048
    scn, opp = setup(pkg, T)
    scn.instances[0] = sc.instance(K, T, 0, 0, 21)
    eng = pkg.Engine(scn)
    eng.set_refine(3, step, 10)
    nom0 = eng.get_nominal(0).copy()
    act = eng.step(1)
    ins = scn.instances[0]
    costs, noise, _ = oracle.rollout(0, scn.params, scn.cost, scn.tracks[0], scn.drones[0], opp, nom0, K, T, ins.seed, 0,
                                     np.array(ins.sigma[:], np.float32))
    upd = oracle.update(0, scn.params, costs, noise, nom0, ins.lambda_, np.array(ins.shift_fill[:], np.float32))
    Us, Js = refine_oracle(oracle, scn, opp, upd["updated"].astype(np.float32), 3, step)
    best, _ = check_refined(eng, act, Us, Js, tol=1e-5 if step <= 5.0 else 1e-4)
    assert best == kept
    eng.close()


@pytest.mark.parametrize("T,ck,collide", [(37, 7, "mid"), (50, 1, "none"), (40, 64, "mid"), (256, 16, "none"), (30, 10, "start"),
                                             (3, 2, "none"), (1, 1, "none")])
def test_refined_step_edge_cases(pkg, oracle, T, ck, collide):
    """The refinement at horizons that are not a multiple of the checkpoint
    interval, at an interval of 1 and one longer than T, at the largest
    horizon, and on trajectories that end in a collision mid-horizon (an
    opponent closing head-on) or at the first step (an opponent 0.3 m away):
    the reverse segments then cover fewer steps than the horizon."""
    say(f"case: refined step T={T} ck={ck} collision={collide}")
    sc = pkg.scenarios
    K = 1024
    scn, opp = setup(pkg, T)
    if collide != "none":
        d0, d1 = scn.drones[0], scn.drones[1]
        gap = 0.3 if collide == "start" else 1.2
        d1.p[0], d1.p[1], d1.p[2] = d0.p[0] + gap, d0.p[1], d0.p[2]
        d1.v[0], d1.v[1], d1.v[2] = (-3.0 if collide == "mid" else 0.0), 0.0, 0.0
        for c in range(3):
            opp[0].p[c], opp[0].v[c] = d1.p[c], d1.v[c]
    scn.instances[0] = sc.instance(K, T, 0, 0, 23)
    eng = pkg.Engine(scn)
    eng.set_refine(3, 0.05, ck)
    nom0 = eng.get_nominal(0).copy()
    act = eng.step(1)
    ins = scn.instances[0]
    costs, noise, _ = oracle.rollout(0, scn.params, scn.cost, scn.tracks[0], scn.drones[0], opp, nom0, K, T, ins.seed, 0,
                                     np.array(ins.sigma[:], np.float32))
    upd = oracle.update(0, scn.params, costs, noise, nom0, ins.lambda_, np.array(ins.shift_fill[:], np.float32))
    U0 = upd["updated"].astype(np.float32)
    _, G, _ = oracle.cost_grad(0, scn.params, scn.cost, scn.tracks[0], scn.drones[0], opp, U0[None], ck=ck)
    live = np.nonzero(np.abs(G[0]).sum(axis=1))[0]
    last = int(live.max()) if live.size else -1
    say(f"  last step with a nonzero gradient: {last} of {T - 1}")
    if collide == "mid":
        assert 0 <= last < T - 8, "the trajectory ends in the collision"
    if collide == "start":
        assert last < 1
    Us, Js = refine_oracle(oracle, scn, opp, U0, 3, 0.05, ck=ck)
    check_refined(eng, act, Us, Js)
    eng.close()


def test_refined_batch_of_uneven_instances(pkg, oracle):
    """Three controllers in one engine (K = 512, 1000, 2048; T = 30, 57, 100;
    each one's opponents are the other drones): every instance's refined
    nominal and action against the oracle's MPPI update and refinement."""
    from test_gpu_parity import opponents_of, small_scenario
    say("case: refinement of a batch of uneven instances")
    sc = pkg.scenarios
    scn = small_scenario(sc, n_samples=512, n_timesteps=30, n_drones=4, controlled=3, seed=9, track="lemniscate")
    scn.instances[1] = sc.instance(1000, 57, 0, 1, sc.seed_mix(9, 1))
    scn.instances[2] = sc.instance(2048, 100, 0, 2, sc.seed_mix(9, 2))
    eng = pkg.Engine(scn)
    eng.set_refine(3, 0.05, 10)
    noms = [eng.get_nominal(i).copy() for i in range(3)]
    acts = eng.step(1)
    for i, ins in enumerate(scn.instances):
        K, T = ins.n_samples, ins.n_timesteps
        opp = opponents_of(pkg, scn, ins.drone)
        costs, noise, _ = oracle.rollout(0, scn.params, scn.cost, scn.tracks[0], scn.drones[ins.drone], opp, noms[i], K, T,
                                         ins.seed, 0, np.array(ins.sigma[:], np.float32))
        upd = oracle.update(0, scn.params, costs, noise, noms[i], ins.lambda_, np.array(ins.shift_fill[:], np.float32))
        one = SimpleNamespace(params=scn.params, cost=scn.cost, tracks=scn.tracks, drones=[scn.drones[ins.drone]])
        Us, Js = refine_oracle(oracle, one, opp, upd["updated"].astype(np.float32), 3, 0.05)
        view = SimpleNamespace(get_nominal=lambda _, i=i: eng.get_nominal(i))  # instance i seen as instance 0
        check_refined(view, acts[i:i + 1], Us, Js)
    eng.close()


def test_cost_grad_rejects_mismatched_shapes(pkg):
    """U and dJ/dU carry n_seq * T * 4 floats; any other length is rejected by
    the C ABI (DR_EINVAL, nothing read or written) and any other shape by the
    host and autograd wrappers."""
    import ctypes as C
    import torch
    from drone_race_amd.autograd import rollout_cost
    T = 20
    scn, _ = setup(pkg, T)
    eng = pkg.Engine(scn)
    U = sequences(pkg, 4, T + 1, 1)  # room for the over-long cases
    J = np.full(4, 7.0, np.float32)
    G = np.full_like(U, 7.0)
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    for n_floats in (4 * T * 4 - 4, 4 * T * 4 + 4, 4 * (T + 1) * 4, 0):
        rc = eng.lib.dr_engine_cost_grad(eng.h, 0, f(U), 4, n_floats, f(J), f(G), 10)
        assert rc == pkg.abi.DR_EINVAL, n_floats
        assert "n_floats" in eng.lib.dr_last_error().decode()
    assert np.all(J == 7.0) and np.all(G == 7.0)
    assert eng.lib.dr_engine_cost_grad(eng.h, 0, f(U), 4, 4 * T * 4, f(J), f(G), 10) == pkg.abi.DR_OK
    for bad in (U[:, :T - 1], U, U[:, :T, :3], U[0, :T]):
        with pytest.raises(ValueError):
            eng.cost_grad(0, bad)
    with pytest.raises(ValueError):
        rollout_cost(torch.zeros((2, T + 1, 4), device="cuda:0"), eng, 0, 10)
    eng.close()


def test_refined_step_through_many_gates(pkg, oracle):
    """Eight gates 0.5 m apart on a straight line, the drone flying through them at 5 m/s: within
    the 60-step horizon (1.2 s) every iterate passes most of the gates (J below -10 x 5: the
    terminal cost is -10 per gate of progress and the running costs are not negative), so the
    refinement's gate counter runs through many crossings per pass.  The refined nominal and
    action against the oracle's iterate of lowest J."""
    say("case: refined step through eight gates 0.5 m apart")
    sc = pkg.scenarios
    T, K, n = 60, 1024, 8
    centers = [[0.5 * i, 0.0, 4.0] for i in range(n)]
    tr = sc._make_track(centers, [[1.0, 0.0, 0.0]] * n, 1.0)
    drones = (pkg.abi.DroneState * 2)()
    for j, y in enumerate((0.0, 6.0)):
        d = drones[j]
        d.p[0], d.p[1], d.p[2] = -0.6, y, 4.0
        d.q[0] = 1.0
        d.v[0] = 5.0
        d.gate, d.laps = n - 1, 0
    scn = sc.Scenario("gates", [tr], [sc.instance(K, T, 0, 0, 31)], [0, 0], drones, 1, "gates")
    o = pkg.abi.Opponent()
    for c in range(3):
        o.p[c], o.v[c] = drones[1].p[c], drones[1].v[c]
    opp = [o]
    eng = pkg.Engine(scn)
    eng.set_refine(3, 0.05, 10)
    nom0 = eng.get_nominal(0).copy()
    act = eng.step(1)
    ins = scn.instances[0]
    costs, noise, _ = oracle.rollout(0, scn.params, scn.cost, scn.tracks[0], scn.drones[0], opp, nom0, K, T, ins.seed, 0,
                                     np.array(ins.sigma[:], np.float32))
    upd = oracle.update(0, scn.params, costs, noise, nom0, ins.lambda_, np.array(ins.shift_fill[:], np.float32))
    Us, Js = refine_oracle(oracle, scn, opp, upd["updated"].astype(np.float32), 3, 0.05)
    assert Js.max() < -50.0
    check_refined(eng, act, Us, Js)
    eng.close()
