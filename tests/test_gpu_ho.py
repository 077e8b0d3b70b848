"""GPU parity of branching spline MPPI on HiddenObs (HIP engine vs the FP32 oracle;
FP64 distance reported).

The bar is the StratRace2D one: spline basis, knot noise, per-model costs, branch
used / branching time, aggregated costs and masked minima bit-exact; nu within 1e-6;
knots, dense plans and action within 1e-5 relative; temperature and chosen branch
exact; closed-loop worlds and beliefs bit-exact; verification failure counts
bit-exact with the Clopper-Pearson outcome against SciPy.
"""
import copy

import numpy as np
import pytest

from test_gpu_bm import REL, check_step, check_verification, rel_err, say

pytestmark = pytest.mark.gpu


def ho_scenario(pkg, specs, min_conf=0.95, belief=(0.5, 0.5), sigma=0.8, pos=(1.0, 4.0), vel=(0.3, 1.0), **env_kw):
    """One HiddenObs world per spec (K, T, M), world i with hidden obstacle i % 2."""
    sc = pkg.scenarios
    env = sc.ho_env(**env_kw)
    inst = [sc.bm_instance(K, T, sc.seed_mix(31, i), n_knots=M, sigma=sigma, min_confidence=min_conf, belief=belief, true_theta=i % 2)
            for i, (K, T, M) in enumerate(specs)]
    st = []
    for _ in specs:
        s = pkg.abi.HoState()
        s.pos[:], s.vel[:] = pos, vel
        st.append(s)
    return sc.SrScenario("ho_test", env, sc.ho_cost(), inst, st, 1, "test")


@pytest.mark.parametrize("case", ["branching", "committed", "wide", "noisy"])
def test_ho_step_parity(pkg, oracle, case):
    say(f"[ho parity] {case}")
    if case == "wide":  # M > 20: knot noise read back from HBM
        sc = ho_scenario(pkg, [(1000, 40, 10), (2048, 64, 30)])
    elif case == "committed":
        sc = ho_scenario(pkg, [(1000, 40, 10), (2048, 48, 20)], belief=(1.0, 0.0))
    elif case == "noisy":
        sc = ho_scenario(pkg, [(1536, 40, 12)], pos_noise=0.02, speed_noise=0.05, action_noise=0.3)
    else:
        sc = ho_scenario(pkg, [(1000, 40, 10), (2048, 48, 20)])
    eng = pkg.HoBMEngine(sc)
    if case in ("branching", "noisy"):  # the rollout instantiation that records its lengths
        eng.set_record_lengths(True)
    rng = np.random.default_rng(17)
    knots = []
    for i, ins in enumerate(sc.instances):
        kn = rng.normal(0, 1.0, (3, ins.n_knots, 2)).astype(np.float32)
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
        assert 0 < np.mean(used != 0) < 1, "some rollouts see a hidden obstacle, some do not"


def test_ho_closed_loop(pkg, oracle):
    say("[ho closed loop] 2 worlds (one per hidden obstacle), 8 steps")
    sc = ho_scenario(pkg, [(768, 30, 8), (1024, 36, 12)], pos=(1.0, 3.6), vel=(0.2, 1.2))
    eng = pkg.HoBMEngine(sc)
    n = len(sc.instances)
    states = [copy.deepcopy(s) for s in eng.get_states()]
    beliefs = [eng.belief(i) for i in range(n)]
    knots = [eng.knots(i) for i in range(n)]
    lambdas = [ins.lambda_ for ins in sc.instances]
    for step in range(8):
        eng.set_record_lengths(step % 3 == 1)
        acts = eng.step(1, advance=True)
        check_step(oracle, eng, sc, states, beliefs, knots, lambdas, acts, step)
        g = eng.get_states()
        ev = eng.events()
        diag = eng.diag()
        for i, ins in enumerate(sc.instances):
            s2, b2, term = oracle.ho_env_step(0, sc.env, ins.true_theta, acts[i], states[i], belief=beliefs[i], seed=ins.seed,
                                              k=0xFFFFFFFF, t=step, step=0)
            assert bytes(g[i]) == bytes(s2), f"world {i} differs at step {step}"
            assert np.array_equal(eng.belief(i).view(np.uint32), b2.view(np.uint32)), "belief not bit-exact"
            assert ev[i] == term
            states[i], beliefs[i] = s2, b2
            knots[i] = eng.knots(i)
            lambdas[i] = diag[i].lambda_
        say(f"  step {step}: worlds and beliefs identical to the oracle; positions "
            + " ".join(f"({s.pos[0]:.2f}, {s.pos[1]:.2f})" for s in states) + "; beliefs "
            + " ".join(np.array2string(b, precision=2) for b in beliefs))


def test_ho_reference_size_step(pkg, oracle):
    """One step at the reference's HiddenObs setting: K = 2^16, T = 80, 20 knots, from the start."""
    say("[ho reference size] K=65536 T=80 M=20")
    sc = pkg.scenarios.ho_race(n_inst=1)
    eng = pkg.HoBMEngine(sc)
    states = [copy.deepcopy(s) for s in eng.get_states()]
    acts = eng.step(1, advance=False)
    ins = sc.instances[0]
    check_step(oracle, eng, sc, states, [eng.belief(0)], [np.zeros((3, ins.n_knots, 2), np.float32)], [ins.lambda_], acts, 0)


def test_ho_verification_parity(pkg, oracle):
    say("[ho verification] 4096 samples, wild previous plans")
    sc = ho_scenario(pkg, [(1024, 30, 8)], min_conf=0.75, belief=(0.3, 0.7), sigma=0.3, pos=(1.0, 1.5), vel=(0.0, 0.8),
                     action_noise=0.3)
    ins = sc.instances[0]
    ins.n_verif_samples, ins.verif_horizon, ins.beta, ins.max_verif_eps = 4096, 24, 1e-5, 0.01
    eng = pkg.HoBMEngine(sc)
    kn = np.zeros((3, 8, 2), np.float32)
    kn[:, :, 1] = 1.0
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
