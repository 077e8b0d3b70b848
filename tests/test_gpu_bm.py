"""GPU parity of branching spline MPPI on StratRace2D (HIP engine vs the FP32
oracle; FP64 distance reported).

Bar: spline basis, knot noise, per-model costs, branch used / branching
time, aggregated costs and masked minima bit-exact; nu within 1e-6 and
knots, dense plans and action within 1e-5 relative (the FP64 weighted sums
are added in another order); temperature and chosen branch exact; closed-loop
worlds and beliefs bit-exact.
"""
import copy
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)))) if a.size else 0.0


def scenario(pkg, specs, min_conf=0.6, belief=(0.5, 0.5), sigma=0.6, start=(0.15, 0.2)):
    """One world per spec (K, T, M); knots at the reference's default spacing."""
    sc = pkg.scenarios
    env = sc.sr_env(n_models=2)
    inst = [sc.bm_instance(K, T, sc.seed_mix(21, i), n_knots=M, sigma=sigma, min_confidence=min_conf, belief=belief, true_theta=i % 2)
            for i, (K, T, M) in enumerate(specs)]
    st = [sc.sr_start_state(env, start=start) for _ in specs]
    return sc.SrScenario("bm_test", env, sc.bm_cost(), inst, st, 1, "test")


def check_step(oracle, eng, sc, states, beliefs, knots, lambdas, acts, step, want64=False):
    diag = eng.diag()
    for i, ins in enumerate(sc.instances):
        r = oracle.bm_step(0, sc.env, sc.cost, states[i], ins, beliefs[i], step, lambdas[i], knots[i])
        assert np.array_equal(eng.noise(i).view(np.uint32), r["noise"].view(np.uint32)), "knot noise not bit-exact"
        used, bt = eng.branches(i)
        np.testing.assert_array_equal(used, r["used"])
        np.testing.assert_array_equal(bt, r["btime"])
        if eng.recording:
            np.testing.assert_array_equal(eng.rollout_lengths(i), r["len"])
        g_ct = eng.costs_true(i)
        np.testing.assert_array_equal(g_ct.astype(np.float64), r["ctrue"])
        np.testing.assert_array_equal(eng.costs(i).astype(np.float64), r["costs"])
        np.testing.assert_array_equal(eng.min_costs(i).astype(np.float64), r["mins"])
        assert rel_err(eng.nu(i), r["nu"]) <= 1e-6
        err = max(rel_err(eng.knots(i), r["knots"]), rel_err(eng.nominal(i), r["dense"]), rel_err(acts[i], r["action"]))
        assert err <= REL, err
        assert diag[i].branch == r["branch"]
        assert diag[i].lambda_ == np.float32(r["lambda_"])
        d64 = ""
        if want64:
            r64 = oracle.bm_step(1, sc.env, sc.cost, states[i], ins, beliefs[i], step, lambdas[i], knots[i], want_noise=False)
            ok = (r64["used"] == used) & (r64["btime"] == bt) & np.isfinite(r64["ctrue"]) & (np.abs(r64["ctrue"]) < 1e4)
            d64 = f"; costs vs FP64 {rel_err(g_ct[ok], r64['ctrue'][ok]):.2e} on {ok.mean():.0%} of rollouts"
        say(f"  inst {i}: K={ins.n_samples} T={ins.n_timesteps} M={ins.n_knots} branch {diag[i].branch} "
            f"commits {np.mean(used != 0):.0%}; exact, plans max rel {err:.2e}{d64}")


@pytest.mark.parametrize("case", ["branching", "committed", "never", "wide"])
def test_bm_step_parity(pkg, oracle, case):
    say(f"[bm parity] {case}")
    if case == "wide":  # M > 20: knot noise read back from HBM
        sc = scenario(pkg, [(1000, 40, 10), (2048, 64, 30)])
    elif case == "committed":
        sc = scenario(pkg, [(1000, 40, 10), (2048, 48, 20)], belief=(0.99, 0.01), min_conf=0.95)
    elif case == "never":
        sc = scenario(pkg, [(1536, 30, 6)], min_conf=2.0)
    else:
        sc = scenario(pkg, [(1000, 40, 10), (2048, 48, 20)])
    eng = pkg.BMEngine(sc)
    if case in ("branching", "wide"):  # the rollout instantiation that records its lengths
        eng.set_record_lengths(True)
    rng = np.random.default_rng(7)
    knots = []
    for i, ins in enumerate(sc.instances):
        kn = rng.normal(0, 0.5, (3, ins.n_knots, 2)).astype(np.float32)
        eng.set_knots(i, kn)
        knots.append(kn)
        B = eng.basis(i)
        want = oracle.bm_basis(0, list(ins.knots[:ins.n_knots]), ins.n_timesteps).astype(np.float32)
        assert np.array_equal(B.view(np.uint32), want.view(np.uint32)), "spline basis not bit-exact"
    states = [copy.deepcopy(s) for s in eng.get_states()]
    beliefs = [eng.belief(i) for i in range(len(sc.instances))]
    lambdas = [ins.lambda_ for ins in sc.instances]
    acts = eng.step(1, advance=False)
    check_step(oracle, eng, sc, states, beliefs, knots, lambdas, acts, 0, want64=True)


def test_bm_closed_loop(pkg, oracle):
    say("[bm closed loop] 2 worlds, 5 steps")
    sc = scenario(pkg, [(768, 30, 8), (1024, 36, 12)], min_conf=0.8)
    eng = pkg.BMEngine(sc)
    n = len(sc.instances)
    states = [copy.deepcopy(s) for s in eng.get_states()]
    beliefs = [eng.belief(i) for i in range(n)]
    knots = [eng.knots(i) for i in range(n)]
    lambdas = [ins.lambda_ for ins in sc.instances]
    for step in range(5):
        eng.set_record_lengths(step in (2, 3))  # the captured graphs follow the switch
        acts = eng.step(1, advance=True)
        check_step(oracle, eng, sc, states, beliefs, knots, lambdas, acts, step)
        g = eng.get_states()
        ev = eng.events()
        diag = eng.diag()
        for i, ins in enumerate(sc.instances):
            s2, b2, term = oracle.sr_env_step_belief(0, sc.env, ins.true_theta, acts[i], states[i], beliefs[i], ins.seed,
                                                     0xFFFFFFFF, step, 0, lose_on_opp_win=False)
            assert bytes(g[i]) == bytes(s2), f"world {i} differs at step {step}"
            assert np.array_equal(eng.belief(i).view(np.uint32), b2.view(np.uint32)), "belief not bit-exact"
            assert ev[i] == term
            states[i], beliefs[i] = s2, b2
            knots[i] = eng.knots(i)
            lambdas[i] = diag[i].lambda_
        say(f"  step {step}: worlds and beliefs identical to the oracle; beliefs "
            + " ".join(np.array2string(b, precision=3) for b in beliefs))


def test_bm_reference_size_step(pkg, oracle):
    """One step at the reference's setting: K = 2^16, T = 80, 20 knots, two opponent models."""
    say("[bm reference size] K=65536 T=80 M=20")
    sc = pkg.scenarios.bm_race(n_inst=1)
    eng = pkg.BMEngine(sc)
    states = [copy.deepcopy(s) for s in eng.get_states()]
    acts = eng.step(1, advance=False)
    ins = sc.instances[0]
    check_step(oracle, eng, sc, states, [eng.belief(0)], [np.zeros((3, ins.n_knots, 2), np.float32)], [ins.lambda_], acts, 0)


def cp_bounds(n):
    from scipy.stats import beta as bd

    def upper(k, a):
        return 1.0 if k == n else float(bd.ppf(1 - a, k + 1, n - k))

    def lower(k, a):
        return 0.0 if k == 0 else float(bd.ppf(a, k, n - k + 1))
    return upper, lower


def check_verification(eng, i, ins, r, prev, T):
    """Failure counts bit-exact; certified loss, decision, epsilon against SciPy's beta quantiles
    (computeCertifiedLoss / computeEpsilon, mppi_controller.cu:468-514); the executed plans."""
    fnew, fold = eng.fail_counts(i)
    np.testing.assert_array_equal(fnew, r["fail_new"])
    np.testing.assert_array_equal(fold, r["fail_old"])
    n = ins.n_verif_samples
    upper, lower = cp_bounds(n)
    half = float(np.float32(ins.beta)) / 2
    cn, co = np.cumsum(fnew), np.cumsum(fold)
    cert = max([-1.0] + [upper(int(cn[t]), half) - lower(int(co[t]), half) for t in range(T)])
    d = eng.diag()[i]
    assert d.certified_loss == pytest.approx(cert, rel=1e-9, abs=1e-12)
    use_new = int(cert < float(np.float32(ins.max_verif_eps)))
    assert d.use_new_plan == use_new
    total = int(cn[-1] if use_new else co[-1])
    assert d.fail_count == total
    eps_p = upper(total, float(np.float32(ins.beta)))
    assert d.epsilon_partial == pytest.approx(eps_p, rel=1e-9, abs=1e-12)
    assert d.epsilon == pytest.approx(eps_p + ins.verif_horizon * float(np.float32(ins.max_verif_eps)), rel=1e-9)
    chosen = (r["dense_new"] if use_new else prev.astype(np.float64)).copy()
    br = r["branch"]
    want = chosen.copy()
    want[br, :-1] = chosen[br, 1:]
    want[br, -1] = 0.0
    assert rel_err(eng.nominal(i), want) <= REL
    return use_new, chosen[br, 0], cert


@pytest.mark.parametrize("plan", ["safe", "crash"])
def test_bm_verification_parity(pkg, oracle, plan):
    say(f"[bm verification] {plan} new plan")
    sc = scenario(pkg, [(1024, 30, 8)], min_conf=0.6, sigma=0.2)
    ins = sc.instances[0]
    # 4096 samples certify no loss below upper(0) = 1 - (beta / 2)^(1/4096) ~ 3e-3: the bound is
    # maxVerifEps = 0.01 (MPPIConfig's default, utils.py:488)
    ins.n_verif_samples, ins.verif_horizon, ins.beta, ins.max_verif_eps = 4096, 24, 1e-5, 0.01
    eng = pkg.BMEngine(sc)
    env = sc.env
    tang = np.array(env.t_grid[77][:])
    normal = np.float32([-tang[1], tang[0]])
    kn = np.zeros((3, 8, 2), np.float32)
    if plan == "crash":
        kn[:] = 1.9 * normal
    eng.set_knots(0, kn)
    prev = eng.nominal(0)
    states = [copy.deepcopy(s) for s in eng.get_states()]
    belief = eng.belief(0)
    acts = eng.step(1, advance=False)
    r = oracle.bm_step(0, env, sc.cost, states[0], ins, belief, 0, ins.lambda_, kn, prev_dense=prev)
    assert np.array_equal(eng.noise(0).view(np.uint32), r["noise"].view(np.uint32))
    np.testing.assert_array_equal(eng.costs_true(0).astype(np.float64), r["ctrue"])
    use_new, a0, cert = check_verification(eng, 0, ins, r, prev, 30)
    assert rel_err(acts[0], a0) <= REL
    say(f"  failures new {int(r['fail_new'].sum())} old {int(r['fail_old'].sum())}; certified loss {cert:.3e} -> "
        f"{'new' if use_new else 'previous'} plan; bit-exact counts")
    assert use_new == (1 if plan == "safe" else 0)


def test_bm_verification_closed_loop(pkg, oracle):
    say("[bm verification closed loop] 3 steps")
    sc = scenario(pkg, [(768, 30, 8)], min_conf=0.8, sigma=0.5)
    ins = sc.instances[0]
    ins.n_verif_samples, ins.verif_horizon, ins.beta, ins.max_verif_eps = 2048, 20, 1e-5, 0.05
    eng = pkg.BMEngine(sc)
    state = copy.deepcopy(eng.get_states()[0])
    belief = eng.belief(0)
    knots, prev, lam = eng.knots(0), eng.nominal(0), ins.lambda_
    for step in range(3):
        acts = eng.step(1, advance=True)
        r = oracle.bm_step(0, sc.env, sc.cost, state, ins, belief, step, lam, knots, prev_dense=prev)
        use_new, a0, cert = check_verification(eng, 0, ins, r, prev, 30)
        assert rel_err(acts[0], a0) <= REL
        s2, b2, term = oracle.sr_env_step_belief(0, sc.env, ins.true_theta, acts[0], state, belief, ins.seed, 0xFFFFFFFF, step, 0,
                                                 lose_on_opp_win=False)
        assert bytes(eng.get_states()[0]) == bytes(s2)
        assert np.array_equal(eng.belief(0).view(np.uint32), b2.view(np.uint32))
        state, belief = s2, b2
        knots, prev, lam = eng.knots(0), eng.nominal(0), eng.diag()[0].lambda_
        say(f"  step {step}: certified loss {cert:.3e}, {'new' if use_new else 'previous'} plan, world identical")
