"""GPU parity of branching spline MPPI on DroneRace (HIP engine vs the FP32 oracle; FP64
distance reported), to the bar of the StratRace2D tests: spline basis, knot noise, per-model
costs, branch used / branching time / rollout length, aggregated costs and masked minima
bit-exact; nu within 1e-6; knots, dense plans and action within 1e-5 relative; closed-loop
worlds and beliefs bit-exact; verification failure counts bit-exact.
"""
import copy

import numpy as np
import pytest

from test_gpu_bm import REL, check_step, check_verification, rel_err, say

pytestmark = pytest.mark.gpu


def drace_scenario(pkg, specs, min_conf=0.9, belief=(0.5, 0.5), sigma=1.0, p0=(4.0, 0.0), p1=(6.0, 0.0), v0=(1.5, 0.0), v1=(1.8, 0.0),
                   **env_kw):
    """One DroneRace world per spec (K, T, M), world i's opponent on model i % 2."""
    sc = pkg.scenarios
    env = sc.drace_env(**env_kw)
    inst = [sc.bm_instance(K, T, sc.seed_mix(41, i), n_knots=M, sigma=sigma, min_confidence=min_conf, belief=belief, true_theta=i % 2)
            for i, (K, T, M) in enumerate(specs)]
    st = []
    for _ in specs:
        s = pkg.abi.DraceState()
        s.pos[:] = [p0[0], p0[1], p1[0], p1[1]]
        s.vel[:] = [v0[0], v0[1], v1[0], v1[1]]
        for a, p in enumerate((p0, p1)):
            for r in range(2):
                s.s[a][r] = p[0] / 27.5
        s.gates[:] = [1, 1]
        st.append(s)
    return sc.SrScenario("drace_test", env, sc.drace_cost(), inst, st, 1, "test")


@pytest.mark.parametrize("case", ["branching", "committed", "wide", "noisy"])
def test_drace_step_parity(pkg, oracle, case):
    say(f"[drace parity] {case}")
    if case == "wide":  # M > 20: knot noise read back from HBM
        sc = drace_scenario(pkg, [(1000, 40, 10), (2048, 64, 30)])
    elif case == "committed":
        sc = drace_scenario(pkg, [(1000, 40, 10), (2048, 48, 20)], belief=(0.02, 0.98))
    elif case == "noisy":
        sc = drace_scenario(pkg, [(1536, 40, 12)], repulsion=3.0)
        sc.env.pos_noise, sc.env.speed_noise = 0.02, 0.05
    else:
        sc = drace_scenario(pkg, [(1000, 40, 10), (2048, 48, 20)])
    eng = pkg.DraceBMEngine(sc)
    if case in ("branching", "noisy"):
        eng.set_record_lengths(True)
    rng = np.random.default_rng(23)
    knots = []
    for i, ins in enumerate(sc.instances):
        kn = (np.float32([2.0, 0.0]) + rng.normal(0, 0.8, (3, ins.n_knots, 2))).astype(np.float32)
        eng.set_knots(i, kn)
        knots.append(kn)
        want = oracle.bm_basis(0, list(ins.knots[:ins.n_knots]), ins.n_timesteps).astype(np.float32)
        assert np.array_equal(eng.basis(i).view(np.uint32), want.view(np.uint32)), "spline basis not bit-exact"
    states = [copy.deepcopy(s) for s in eng.get_states()]
    beliefs = [eng.belief(i) for i in range(len(sc.instances))]
    lambdas = [ins.lambda_ for ins in sc.instances]
    acts = eng.step(1, advance=False)
    check_step(oracle, eng, sc, states, beliefs, knots, lambdas, acts, 0, want64=True)
    if case == "branching":
        used, _ = eng.branches(1)
        assert 0 < np.mean(used != 0) <= 1, "the opponent's model shows where the racelines split"


def test_drace_closed_loop(pkg, oracle):
    say("[drace closed loop] 2 worlds (one per opponent model), 8 steps")
    sc = drace_scenario(pkg, [(768, 30, 8), (1024, 36, 12)])
    eng = pkg.DraceBMEngine(sc)
    n = len(sc.instances)
    states = [copy.deepcopy(s) for s in eng.get_states()]
    beliefs = [eng.belief(i) for i in range(n)]
    knots = [eng.knots(i) for i in range(n)]
    lambdas = [ins.lambda_ for ins in sc.instances]
    for step in range(8):
        eng.set_record_lengths(step % 3 == 2)
        acts = eng.step(1, advance=True)
        check_step(oracle, eng, sc, states, beliefs, knots, lambdas, acts, step)
        g = eng.get_states()
        ev = eng.events()
        diag = eng.diag()
        for i, ins in enumerate(sc.instances):
            s2, b2, term = oracle.drace_env_step(0, sc.env, ins.true_theta, acts[i], states[i], belief=beliefs[i], seed=ins.seed,
                                                 k=0xFFFFFFFF, t=step, step=0, lose_on_opp_win=False)
            assert bytes(g[i]) == bytes(s2), f"world {i} differs at step {step}"
            assert np.array_equal(eng.belief(i).view(np.uint32), b2.view(np.uint32)), "belief not bit-exact"
            assert ev[i] == term
            states[i], beliefs[i] = s2, b2
            knots[i] = eng.knots(i)
            lambdas[i] = diag[i].lambda_
        say(f"  step {step}: worlds and beliefs identical to the oracle; agents "
            + " ".join(f"({s.pos[0]:.2f}, {s.pos[1]:.2f})/({s.pos[2]:.2f}, {s.pos[3]:.2f})" for s in states) + "; beliefs "
            + " ".join(np.array2string(b, precision=2) for b in beliefs))


def test_drace_reference_size_step(pkg, oracle):
    """One step at the reference's DroneRace spline setting (tinyGateEnv): K = 2^16, T = 60, 10 knots."""
    say("[drace reference size] K=65536 T=60 M=10")
    sc = pkg.scenarios.drace_race(n_inst=1)
    eng = pkg.DraceBMEngine(sc)
    eng.set_record_lengths(True)
    states = [copy.deepcopy(s) for s in eng.get_states()]
    acts = eng.step(1, advance=False)
    ins = sc.instances[0]
    check_step(oracle, eng, sc, states, [eng.belief(0)], [np.zeros((3, ins.n_knots, 2), np.float32)], [ins.lambda_], acts, 0)


def test_drace_verification_parity(pkg, oracle):
    say("[drace verification] 4096 samples")
    sc = drace_scenario(pkg, [(1024, 30, 8)], min_conf=0.8, belief=(0.4, 0.6), sigma=0.5, p0=(6.0, 0.0), p1=(7.5, 0.0))
    ins = sc.instances[0]
    ins.n_verif_samples, ins.verif_horizon, ins.beta, ins.max_verif_eps = 4096, 24, 1e-5, 0.01
    eng = pkg.DraceBMEngine(sc)
    kn = np.tile(np.float32([2.5, 0.0]), (3, 8, 1))
    eng.set_knots(0, kn)
    prev = eng.nominal(0)
    states = [copy.deepcopy(s) for s in eng.get_states()]
    belief = eng.belief(0)
    acts = eng.step(1, advance=False)
    r = oracle.bm_step(0, sc.env, sc.cost, states[0], ins, belief, 0, ins.lambda_, kn, prev_dense=prev)
    assert np.array_equal(eng.noise(0).view(np.uint32), r["noise"].view(np.uint32))
    np.testing.assert_array_equal(eng.costs_true(0).astype(np.float64), r["ctrue"])
    use_new, a0, cert = check_verification(eng, 0, ins, r, prev, 30)
    assert rel_err(acts[0], a0) <= REL
    say(f"  failures new {int(r['fail_new'].sum())} old {int(r['fail_old'].sum())}; certified loss {cert:.3e} -> "
        f"{'new' if use_new else 'previous'} plan; bit-exact counts")
