"""GPU parity of PR-MPPI on the reference's StratRace2D environment (HIP
engine vs the FP32 oracle; FP64 distance reported).

Bar: control noise, per-rollout costs and safeties, aggregates, minima and
model draws bit-exact; candidates within 1e-5 relative; given the engine's
candidates, candidate costs, flags, action and plans bit-exact; closed-loop
worlds bit-exact.
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


def shifted(plan, fill=(0.0, 0.0)):
    return np.vstack([plan[1:], np.asarray(fill, np.float32)[None]]).astype(np.float32)


def check_step(oracle, eng, sc, states, plans, acts, step=0, beliefs=None):
    diag = eng.diag()
    for i, ins in enumerate(sc.instances):
        belief = np.array(ins.belief[:sc.env.n_models], np.float32) if beliefs is None else beliefs[i]
        lam = [ins.lambda_] * 3
        nom_w, rob_w = shifted(plans[i][0]), shifted(plans[i][1])
        r = oracle.prsr_step(0, sc.env, sc.cost, states[i], ins, belief, step, lam, nom_w, rob_w)
        assert np.array_equal(eng.noise(i).view(np.uint32), r["noise"].view(np.uint32)), "noise not bit-exact"
        np.testing.assert_array_equal(eng.models(i), r["models"])
        g_pairs = eng.pair_costs(i)
        np.testing.assert_array_equal(g_pairs.astype(np.float64), r["pairs"])
        np.testing.assert_array_equal(eng.aggregates(i).astype(np.float64), r["agg"])
        np.testing.assert_array_equal(np.array(diag[i].min_costs[:], np.float64), r["min"])
        assert rel_err(diag[i].nu[:], r["nu"]) <= 1e-6
        c1, c2 = eng.candidates(i)
        nom_after, rob_after = eng.get_nominals(i)
        err = max(rel_err(c1, r["cand"][0]), rel_err(c2, r["cand"][1]), rel_err(rob_after, r["cand"][2]))
        r64 = oracle.prsr_step(1, sc.env, sc.cost, states[i], ins, belief, step, lam, nom_w, rob_w, want_noise=False)
        ok = (np.abs(r64["pairs"][..., 0]) < 1e4).all(axis=(0, 1))
        d64 = rel_err(g_pairs[:, :, ok, 0], r64["pairs"][:, :, ok, 0]) if ok.any() else 0.0
        say(f"  inst {i}: K={ins.n_samples} T={ins.n_timesteps} M={r['models'].size} pairs exact, candidates max rel {err:.2e}; "
            f"costs vs FP64 {d64:.2e}")
        assert err <= REL
        cand = np.stack([c1, c2, rob_after]).astype(np.float64)
        c = oracle.prsr_choose(0, sc.env, sc.cost, states[i], ins, belief, r["models"], cand, step, lam,
                               np.array(diag[i].nu[:], np.float64))
        np.testing.assert_array_equal(np.array(diag[i].cand_costs[:], np.float64), c["cand_cost"])
        assert diag[i].use_nom_plan == c["use_nom"] and diag[i].reset_nom == c["reset_nom"]
        np.testing.assert_array_equal(acts[i].astype(np.float64), c["action"])
        pick = 0 if c["cand_cost"][0] <= c["cand_cost"][1] else 1
        np.testing.assert_array_equal(nom_after, cand[pick].astype(np.float32))


@pytest.mark.parametrize("sampled", [0, 1])
def test_prsr_step_parity(pkg, oracle, sampled):
    say(f"[pr-sr parity] sampled={sampled}")
    sc = pkg.scenarios.pr_sr_race(n_inst=2, n_samples=1024, n_timesteps=30, sampled=sampled)
    eng = pkg.PRSREngine(sc)
    rng = np.random.default_rng(9)
    plans = []
    for i in range(2):
        nom = rng.normal(0, 0.6, (30, 2)).astype(np.float32)
        rob = rng.normal(0, 0.6, (30, 2)).astype(np.float32)
        eng.set_nominals(i, nom, rob)
        plans.append((nom, rob))
    states = [copy.deepcopy(s) for s in eng.get_states()]
    acts = eng.step(1, advance=False)
    check_step(oracle, eng, sc, states, plans, acts)


def test_prsr_closed_loop(pkg, oracle):
    say("[pr-sr closed loop] 2 worlds, 4 steps, belief updated by the world step")
    sc = pkg.scenarios.pr_sr_race(n_inst=2, n_samples=512, n_timesteps=25)
    eng = pkg.PRSREngine(sc)
    NT = sc.env.n_models
    plans = [(np.zeros((25, 2), np.float32), np.zeros((25, 2), np.float32)) for _ in range(2)]
    states = [copy.deepcopy(s) for s in eng.get_states()]
    beliefs = [np.array(ins.belief[:NT], np.float32) for ins in sc.instances]
    for n in range(4):
        acts = eng.step(1, advance=True)
        check_step(oracle, eng, sc, states, plans, acts, step=n, beliefs=beliefs)
        g = eng.get_states()
        ev = eng.events()
        for i, ins in enumerate(sc.instances):
            s2, b2, term = oracle.sr_env_step_belief(0, sc.env, ins.true_theta, acts[i], states[i], beliefs[i], ins.seed, 0xFFFFFFFF, n,
                                                     0, lose_on_opp_win=False)
            got = np.array(list(g[i].pos) + list(g[i].vel) + list(g[i].s) + list(g[i].lat_dist))
            want = np.array(list(s2.pos) + list(s2.vel) + list(s2.s) + list(s2.lat_dist))
            np.testing.assert_array_equal(got, want)
            assert np.array_equal(eng.belief(i, NT).view(np.uint32), b2.view(np.uint32)), "belief not bit-exact"
            assert ev[i] == term
            states[i], beliefs[i] = s2, b2
        plans = [eng.get_nominals(i) for i in range(2)]
        say(f"  step {n}: worlds and beliefs identical to the oracle; beliefs "
            + " ".join(np.array2string(b, precision=3) for b in beliefs))


def test_prsr_reference_size_step(pkg, oracle):
    """One step at the reference's PR-MPPI setting: K = 2^16, T = 60, P = 9 sampled models."""
    say("[pr-sr reference size] K=65536 T=60 P=9")
    sc = pkg.scenarios.pr_sr_race(n_inst=1, n_samples=65536, n_timesteps=60)
    eng = pkg.PRSREngine(sc)
    plans = [(np.zeros((60, 2), np.float32), np.zeros((60, 2), np.float32))]
    states = [copy.deepcopy(s) for s in eng.get_states()]
    acts = eng.step(1, advance=False)
    check_step(oracle, eng, sc, states, plans, acts)
