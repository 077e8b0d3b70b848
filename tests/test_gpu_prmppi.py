"""GPU parity of the PR-MPPI control step (HIP engine vs the FP32 oracle, the
FP64 oracle's distance reported).

Bar: noise, per-rollout costs and safeties, per-sample aggregates, minima and
model draws bit-exact (same FP32 operations in the same order); the three
weighted averages within 1e-5 relative (FP64 sums in another order, so the
FP32 result can differ by one rounding); given the engine's candidates, the
candidate costs, the choice flags, the action and the next plans bit-exact.
"""
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


def scenario(sc, K=512, T=30, sampled=0, n_models=0, seed=2, moving=True):
    s = sc.c3_pr_four_drones(seed=seed, n_samples=K, n_timesteps=T, sampled=sampled, n_models=n_models)
    if moving:
        for j, d in enumerate(s.drones):
            d.v[0], d.v[1], d.v[2] = 0.8 * (j - 1.5), 0.6, 0.1 * j
    return s


def shifted(plan, fill):
    return np.vstack([plan[1:], np.asarray(fill, np.float32)[None]]).astype(np.float32)


def check_step(pkg, oracle, eng, s, plans, step=0, lams=None):
    """Compare the engine's last step of every instance with the oracle.
    plans[i] = (nom, rob) unshifted plans before the step."""
    acts = eng.actions_last
    diag = eng.diag()
    for i, ins in enumerate(s.instances):
        fill = np.array(ins.shift_fill[:], np.float32)
        nom_w, rob_w = shifted(plans[i][0], fill), shifted(plans[i][1], fill)
        belief = np.array(ins.belief[:ins.n_branches], np.float32)
        lam = [ins.lambda_] * 3 if lams is None else lams[i]
        opps = oracle.world_opponents(s, s.drones_before, ins.drone)
        ego = s.drones_before[ins.drone]
        tr = s.tracks[ins.track]
        r = oracle.pr_step(0, s.params, s.cost, tr, ego, opps, ins, belief, step, lam, nom_w, rob_w)
        g_noise = eng.noise(i)
        assert np.array_equal(g_noise.view(np.uint32), r["noise"].view(np.uint32)), "noise not bit-exact"
        np.testing.assert_array_equal(eng.models(i), r["models"])
        g_pairs = eng.pair_costs(i)
        exact = np.mean(g_pairs.astype(np.float64) == r["pairs"])
        np.testing.assert_array_equal(g_pairs.astype(np.float64), r["pairs"])
        np.testing.assert_array_equal(eng.aggregates(i).astype(np.float64), r["agg"])
        np.testing.assert_array_equal(np.array(diag[i].min_costs[:], np.float64), r["min"])
        assert rel_err(diag[i].nu[:], r["nu"]) <= 1e-6
        c1, c2 = eng.candidates(i)
        nom_after, rob_after = eng.get_nominals(i)
        e1, e2, e3 = rel_err(c1, r["cand"][0]), rel_err(c2, r["cand"][1]), rel_err(rob_after, r["cand"][2])
        n_exact = np.mean(np.concatenate([c1.ravel(), c2.ravel(), rob_after.ravel()]).astype(np.float64) ==
                          r["cand"].ravel())
        r64 = oracle.pr_step(1, s.params, s.cost, tr, ego, opps, ins, belief, step, lam, nom_w, rob_w, want_noise=False)
        ok = (r64["pairs"][..., 1] < -0.05).all(axis=(0, 1))
        d64 = rel_err(g_pairs[:, :, ok, 0], r64["pairs"][:, :, ok, 0]) if ok.any() else 0.0
        say(f"  inst {i}: K={ins.n_samples} T={ins.n_timesteps} M={r['models'].size} pairs exact={exact:.4f} "
            f"cands exact={n_exact:.4f} max rel {max(e1, e2, e3):.2e}; costs vs FP64 {d64:.2e}, "
            f"cands vs FP64 {rel_err(np.stack([c1, c2, rob_after]), r64['cand']):.2e}")
        assert max(e1, e2, e3) <= REL
        # choice from the engine's candidates
        cand = np.stack([c1, c2, rob_after]).astype(np.float64)
        c = oracle.pr_choose(0, s.params, s.cost, tr, ego, opps, ins, belief, r["models"], cand, lam,
                             np.array(diag[i].nu[:], np.float64))
        np.testing.assert_array_equal(np.array(diag[i].cand_costs[:], np.float64), c["cand_cost"])
        assert diag[i].use_nom_plan == c["use_nom"] and diag[i].reset_nom == c["reset_nom"]
        np.testing.assert_array_equal(acts[i].astype(np.float64), c["action"])
        pick = 0 if c["cand_cost"][0] <= c["cand_cost"][1] else 1
        np.testing.assert_array_equal(nom_after, cand[pick].astype(np.float32))
        np.testing.assert_array_equal(np.array(diag[i].lambda_[:], np.float32), np.float32(c["lam"]))


class Eng:
    """PREngine plus the inputs of its last step."""

    def __init__(self, pkg, s):
        import copy
        self.e = pkg.PREngine(s)
        self.s = s
        self.s.drones_before = copy.deepcopy(s.drones)
        self.actions_last = None

    def __getattr__(self, k):
        return getattr(self.e, k)

    def step(self, advance=False):
        import copy
        self.s.drones_before = copy.deepcopy(self.e.get_drones())
        self.actions_last = self.e.step(1, advance)
        return self.actions_last


@pytest.mark.parametrize("sampled,n_models", [(0, 0), (1, 9)])
def test_pr_step_parity(pkg, oracle, sampled, n_models):
    say(f"[pr parity] sampled={sampled} n_models={n_models}")
    s = scenario(pkg.scenarios, K=512, T=30, sampled=sampled, n_models=n_models)
    eng = Eng(pkg, s)
    rng = np.random.default_rng(5)
    plans = []
    for i, ins in enumerate(s.instances):
        hov = np.tile(np.array(pkg.scenarios.hover_control(), np.float32), (ins.n_timesteps, 1))
        nom = (hov + rng.normal(0, 0.3, hov.shape)).astype(np.float32)
        rob = (hov + rng.normal(0, 0.3, hov.shape)).astype(np.float32)
        eng.set_nominals(i, nom, rob)
        plans.append((nom, rob))
    eng.step(advance=False)
    check_step(pkg, oracle, eng, s, plans)


def test_pr_closed_loop(pkg, oracle):
    say("[pr closed loop] 4 drones, 5 steps")
    s = scenario(pkg.scenarios, K=256, T=25)
    eng = Eng(pkg, s)
    import copy
    drones = copy.deepcopy(s.drones)
    works = []
    for ins in s.instances:
        hov = np.tile(np.array(ins.shift_fill[:], np.float32), (ins.n_timesteps, 1))
        works.append((shifted(hov, ins.shift_fill[:]), shifted(hov, ins.shift_fill[:])))
    steps = [0] * len(s.instances)
    lams = [np.array([ins.lambda_] * 3, np.float32) for ins in s.instances]
    for k in range(5):
        g_act = eng.step(advance=True)
        o_act, _ = oracle.pr_world_step(s, drones, works, steps, lams)
        err = rel_err(g_act, o_act)
        g_dr = eng.get_drones()
        pos = max(rel_err(g_dr[d].p[:], drones[d].p[:]) for d in range(4))
        say(f"  step {k}: actions max rel {err:.2e}, positions max rel {pos:.2e}")
        assert err <= REL and pos <= REL
        for d in range(4):
            assert g_dr[d].gate == drones[d].gate and g_dr[d].laps == drones[d].laps


def test_pr_launch_shapes_agree(pkg):
    say("[pr launch shapes]")
    s = scenario(pkg.scenarios, K=700, T=33)
    ref = None
    for waves, chunk, tg in [(0, 1024, 16), (1, 64, 7), (3, 2048, 32), (8, 256, 1), (16, 512, 64)]:
        eng = pkg.PREngine(s)
        eng.set_launch(rollout_block=64 * waves, reduce_chunk=chunk, reduce_tgroup=tg)
        eng.step(1, advance=True)
        out = [(eng.pair_costs(i), eng.aggregates(i), eng.candidates(i), eng.get_nominals(i)) for i in range(4)]
        dr = eng.get_drones()
        eng.close()
        if ref is None:
            ref = (out, dr)
            continue
        for i in range(4):
            np.testing.assert_array_equal(out[i][0], ref[0][i][0])
            np.testing.assert_array_equal(out[i][1], ref[0][i][1])
            assert rel_err(out[i][2][0], ref[0][i][2][0]) <= 1e-6
            assert rel_err(out[i][2][1], ref[0][i][2][1]) <= 1e-6
            assert rel_err(out[i][3][1], ref[0][i][3][1]) <= 1e-6
        assert max(rel_err(dr[d].p[:], ref[1][d].p[:]) for d in range(4)) <= 1e-6
        say(f"  waves={waves} chunk={chunk} tgroup={tg}: rollouts identical, reductions within 1e-6")


def test_pr_graph_replay_matches_eager(pkg):
    say("[pr graph replay]")
    s = scenario(pkg.scenarios, K=300, T=20)
    res = []
    for graph, spg in [(0, 1), (1, 1), (1, 3)]:
        eng = pkg.PREngine(s)
        eng.set_launch(use_graph=graph, steps_per_graph=spg)
        acts = [eng.step(3, advance=True) for _ in range(2)]
        res.append((acts, [eng.get_nominals(i) for i in range(4)], eng.get_drones()))
        eng.close()
    for r in res[1:]:
        for a, b in zip(r[0], res[0][0]):
            np.testing.assert_array_equal(a, b)
        for i in range(4):
            np.testing.assert_array_equal(r[1][i][0], res[0][1][i][0])
            np.testing.assert_array_equal(r[1][i][1], res[0][1][i][1])
        for d in range(4):
            np.testing.assert_array_equal(np.array(r[2][d].p[:]), np.array(res[0][2][d].p[:]))
    say("  eager, graph (1 step), graph (3 steps): identical over 6 closed-loop steps")


def test_pr_set_belief_changes_draws(pkg, oracle):
    say("[pr belief]")
    s = scenario(pkg.scenarios, K=128, T=10, sampled=1, n_models=16)
    eng = Eng(pkg, s)
    ins = s.instances[0]
    B = ins.n_branches
    b = np.zeros(B, np.float32)
    b[-1] = 1.0
    eng.set_belief(0, b)
    eng.step()
    assert np.all(eng.models(0) == B - 1)
    b = np.zeros(B, np.float32)
    b[0] = 1.0
    eng.set_belief(0, b)
    eng.step()
    assert np.all(eng.models(0) == 0)


def test_pr_full_size_step(pkg, oracle):
    """One step of the 4-drone PR-MPPI workload at K=8192, T=80."""
    say("[pr full size] K=8192 T=80")
    s = scenario(pkg.scenarios, K=8192, T=80, moving=False)
    eng = Eng(pkg, s)
    plans = []
    for ins in s.instances:
        hov = np.tile(np.array(ins.shift_fill[:], np.float32), (ins.n_timesteps, 1))
        plans.append((hov, hov))
    eng.step(advance=False)
    check_step(pkg, oracle, eng, s, plans)
