"""GPU parity of plain MPPI on the reference's StratRace2D environment (HIP
engine vs the FP32 oracle restating env_stratrace.h / costs.cuh /
kernels.cu; FP64 distance reported).

Bar: control noise bit-exact; per-sample costs bit-exact (same FP32
operations in the same order, environment noise included); updated nominal,
action, min cost and nu within 1e-5 relative (FP64 sums in another order);
closed-loop world states within 1e-5 relative, events equal.
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


def scenario(pkg, n_inst=2, K=2048, T=40, **env_kw):
    return pkg.scenarios.sr_race(n_inst=n_inst, n_samples=K, n_timesteps=T, **env_kw)


def check_step(oracle, eng, sc, states, noms, step=0):
    acts = eng.last_actions
    diag = eng.diag()
    for i, ins in enumerate(sc.instances):
        costs, noise = oracle.sr_rollout(0, sc.env, sc.cost, states[i], ins.true_theta, noms[i], ins.n_samples, ins.n_timesteps,
                                         ins.seed, step, ins.sigma)
        g_noise = eng.noise(i)
        assert np.array_equal(g_noise.view(np.uint32), noise.view(np.uint32)), "noise not bit-exact"
        g_costs = eng.costs(i)
        exact = np.mean(g_costs.astype(np.float64) == costs)
        c64, _ = oracle.sr_rollout(1, sc.env, sc.cost, states[i], ins.true_theta, noms[i], ins.n_samples, ins.n_timesteps,
                                   ins.seed, step, ins.sigma, want_noise=False)
        ok = (np.abs(c64) < 1e4) & (np.abs(costs) < 1e4)
        d64 = rel_err(g_costs[ok], c64[ok]) if ok.any() else 0.0
        upd = oracle.sr_update(0, costs, noise, noms[i], ins.lambda_, sc.env.max_accel[0])
        g_nom = eng.get_nominal(i)
        e_nom = rel_err(g_nom, upd["shifted"])
        e_act = rel_err(acts[i], upd["action"])
        say(f"  inst {i}: K={ins.n_samples} T={ins.n_timesteps} theta={ins.true_theta} costs exact={exact:.4f} "
            f"vs FP64 (no discrete jump) {d64:.2e}; nominal {e_nom:.2e} action {e_act:.2e}")
        np.testing.assert_array_equal(g_costs.astype(np.float64), costs)
        assert e_nom <= REL and e_act <= REL
        assert diag[i]["min_cost"] == np.float32(upd["min_cost"])
        assert rel_err(diag[i]["nu"], upd["nu"]) <= REL


class Eng:
    def __init__(self, pkg, sc):
        self.e = pkg.SREngine(sc)
        self.last_actions = None

    def __getattr__(self, k):
        return getattr(self.e, k)

    def step(self, advance=False):
        self.last_actions = self.e.step(1, advance)
        return self.last_actions


@pytest.mark.parametrize("noisy", [False, True])
def test_sr_step_parity(pkg, oracle, noisy):
    say(f"[sr parity] environment noise={noisy}")
    kw = dict(pos_noise=0.01, speed_noise=0.02, env_action_noise=0.1) if noisy else {}
    sc = scenario(pkg, **kw)
    eng = Eng(pkg, sc)
    rng = np.random.default_rng(3)
    noms = []
    for i, ins in enumerate(sc.instances):
        nom = rng.normal(0, 0.8, (ins.n_timesteps, 2)).astype(np.float32)
        eng.set_nominal(i, nom)
        noms.append(nom)
    states = [copy.deepcopy(s) for s in eng.get_states()]
    eng.step(advance=False)
    check_step(oracle, eng, sc, states, noms)


def test_sr_closed_loop(pkg, oracle):
    say("[sr closed loop] 2 worlds, 6 steps")
    sc = scenario(pkg, K=1024, T=30)
    eng = Eng(pkg, sc)
    noms = [np.zeros((ins.n_timesteps, 2), np.float32) for ins in sc.instances]
    states = [copy.deepcopy(s) for s in eng.get_states()]
    for n in range(6):
        before = [copy.deepcopy(s) for s in states]
        acts = eng.step(advance=True)
        check_step(oracle, eng, sc, before, noms, step=n)
        g_states = eng.get_states()
        g_ev = eng.events()
        for i, ins in enumerate(sc.instances):
            s2, term = oracle.sr_env_step(0, sc.env, ins.true_theta, acts[i], before[i], ins.seed, 0xFFFFFFFF, n, 0,
                                          lose_on_opp_win=False)
            got = np.array(list(g_states[i].pos) + list(g_states[i].vel) + list(g_states[i].s) + list(g_states[i].lat_dist))
            want = np.array(list(s2.pos) + list(s2.vel) + list(s2.s) + list(s2.lat_dist))
            np.testing.assert_array_equal(got, want)
            assert list(g_states[i].laps) == list(s2.laps) and g_ev[i] == term
            states[i] = s2
        noms = [eng.get_nominal(i) for i in range(len(sc.instances))]
        say(f"  step {n}: worlds identical to the oracle, events {list(g_ev)}")


def test_sr_launch_shapes_agree(pkg):
    say("[sr launch shapes]")
    sc = scenario(pkg, K=1500, T=25)
    ref = None
    for block, chunk, tg, graph in [(128, 1024, 16, 1), (64, 64, 5, 0), (256, 4096, 32, 1)]:
        eng = pkg.SREngine(sc)
        eng.set_launch(rollout_block=block, reduce_chunk=chunk, reduce_tgroup=tg, use_graph=graph)
        eng.step(1, advance=True)
        out = [(eng.costs(i), eng.get_nominal(i)) for i in range(2)]
        eng.close()
        if ref is None:
            ref = out
            continue
        for i in range(2):
            np.testing.assert_array_equal(out[i][0], ref[i][0])
            assert rel_err(out[i][1], ref[i][1]) <= 1e-6
        say(f"  block={block} chunk={chunk} tgroup={tg} graph={graph}: identical costs, nominal within 1e-6")


def test_sr_reference_size_step(pkg, oracle):
    """One step at the reference's dense setting: K = 2^16, T = 70."""
    say("[sr reference size] K=65536 T=70")
    sc = scenario(pkg, n_inst=1, K=65536, T=70)
    eng = Eng(pkg, sc)
    noms = [np.zeros((70, 2), np.float32)]
    states = [copy.deepcopy(s) for s in eng.get_states()]
    eng.step(advance=False)
    check_step(oracle, eng, sc, states, noms)
