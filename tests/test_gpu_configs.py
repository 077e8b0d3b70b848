"""GPU parity at the benchmark configurations' own workloads (HIP engine vs the
FP32 oracle, FP64 distance printed):

  single quadrotor   1 drone, K = 1024, T = 50, 8-gate circle r = 20 m, seed 0:
                     one step sample by sample, then the closed loop
  64 uneven          64 single-drone controllers, K ~ logU[1024, 32768],
                     T ~ U{30..150}, each on its own circle: one step of all 64
                     sample by sample, three closed-loop steps against the
                     oracle's closed loop, then a fourth step sample by sample
                     from the engine's own states
  8 drones refined   8 drones, K = 4096, T = 100, MPPI + 3 projected gradient
                     steps (0.05, checkpoints every 10): every controller's kept
                     iterate against the oracle's MPPI update and refinement

Bar: noise and costs bit-exact, nominals / actions / min cost / nu within
1e-5 relative; closed-loop actions and drone states within 1e-4 relative (the
FP32 update may differ from the oracle's by one rounding per step); refined
plans within 1e-4 of the oracle iterate of lowest J.
"""
import time
from types import SimpleNamespace

import numpy as np
import pytest

from test_gpu_adjoint import check_refined, refine_oracle
from test_gpu_parity import check_instance, make_engine, opponents_of, rel_err, say

pytestmark = pytest.mark.gpu


def _state_rows(drones):
    return np.array([list(d.p) + list(d.q) + list(d.v) + list(d.w) for d in drones], dtype=np.float64)


def _flat_nominals(scn):
    return np.concatenate([np.tile(np.array(i.shift_fill[:], np.float32), (i.n_timesteps, 1)).reshape(-1)
                           for i in scn.instances]).astype(np.float32)


def test_single_quadrotor_workload(pkg, oracle):
    sc = pkg.scenarios
    scn = sc.c0_single()
    ins = scn.instances[0]
    say(f"case: {scn.name}: K={ins.n_samples} T={ins.n_timesteps}, one step sample by sample")
    assert (ins.n_samples, ins.n_timesteps, scn.tracks[0].n_gates) == (1024, 50, 8)
    eng = make_engine(pkg, scn)
    nom0 = eng.get_nominal(0).copy()
    acts = eng.step(1)
    upd = check_instance(pkg, oracle, eng, scn, 0, nom0)
    assert rel_err(acts[0], upd["action"]) <= 1e-5
    eng.close()

    n_steps = 100
    say(f"case: {scn.name}: closed loop, {n_steps} steps against the oracle's closed loop")
    eng = pkg.Engine(scn)
    g_acts = eng.step(n_steps, advance=True)
    g_drones = eng.get_drones()
    drones = (pkg.abi.DroneState * 1)(*scn.drones)
    o_acts, _ = oracle.world_run(scn, drones, _flat_nominals(scn), [0], [ins.lambda_], n_steps, threads=16)
    e_act, e_state = rel_err(g_acts, o_acts), rel_err(_state_rows(g_drones), _state_rows(drones))
    say(f"  after {n_steps} steps: last action max rel {e_act:.2e}, drone state max rel {e_state:.2e}, gate "
        f"{g_drones[0].gate} (oracle {drones[0].gate}), laps {g_drones[0].laps}")
    assert e_act <= 1e-4 and e_state <= 1e-4
    assert (g_drones[0].gate, g_drones[0].laps) == (drones[0].gate, drones[0].laps)
    eng.close()


def test_batch64_uneven_workload(pkg, oracle):
    sc = pkg.scenarios
    scn = sc.c2_batch64()
    n = len(scn.instances)
    Ks = [i.n_samples for i in scn.instances]
    Ts = [i.n_timesteps for i in scn.instances]
    say(f"case: {scn.name}: {n} instances, K in [{min(Ks)}, {max(Ks)}], T in [{min(Ts)}, {max(Ts)}], "
        f"{scn.sample_steps_per_step} sample-steps per step; one step, every instance sample by sample")
    assert n == 64 and 1024 <= min(Ks) and max(Ks) <= 32768 and 30 <= min(Ts) and max(Ts) <= 150
    eng = make_engine(pkg, scn)
    noms = [eng.get_nominal(i).copy() for i in range(n)]
    eng.step(1)
    for i in range(n):
        check_instance(pkg, oracle, eng, scn, i, noms[i])
    eng.close()

    n_loop = 3
    say(f"case: {scn.name}: {n_loop} closed-loop steps against the oracle's closed loop")
    eng = pkg.Engine(scn)
    t0 = time.perf_counter()
    g_acts = eng.step(n_loop, advance=True)
    g_drones = eng.get_drones()
    drones = (pkg.abi.DroneState * n)(*scn.drones)
    o_acts, _ = oracle.world_run(scn, drones, _flat_nominals(scn), [0] * n, [i.lambda_ for i in scn.instances], n_loop,
                                 threads=16)
    e_act, e_state = rel_err(g_acts, o_acts), rel_err(_state_rows(g_drones), _state_rows(drones))
    say(f"  after {n_loop} steps: actions max rel {e_act:.2e}, drone states max rel {e_state:.2e} "
        f"({time.perf_counter() - t0:.1f} s with the oracle)")
    assert e_act <= 1e-4 and e_state <= 1e-4

    say(f"case: {scn.name}: step {n_loop + 1} sample by sample from the engine's own states and plans")
    here = SimpleNamespace(params=scn.params, cost=scn.cost, tracks=scn.tracks, instances=scn.instances,
                           drone_world=scn.drone_world, drones=(pkg.abi.DroneState * n)(*g_drones))
    noms = [eng.get_nominal(i).copy() for i in range(n)]
    eng.set_eval(True)
    eng.step(1)
    for i in range(n):
        check_instance(pkg, oracle, eng, here, i, noms[i], step=n_loop)
    eng.close()


def test_eight_drone_refined_workload(pkg, oracle):
    sc = pkg.scenarios
    scn = sc.c4_eight_drones(4096)
    n = len(scn.instances)
    say(f"case: {scn.name}: MPPI K=4096 T=100 + 3 gradient steps (0.05, checkpoints every 10), every controller")
    assert n == 8 and all((i.n_samples, i.n_timesteps) == (4096, 100) for i in scn.instances)
    eng = pkg.Engine(scn)
    eng.set_refine(3, 0.05, 10)
    noms = [eng.get_nominal(i).copy() for i in range(n)]
    acts = eng.step(1)
    kept = []
    for i, ins in enumerate(scn.instances):
        opp = opponents_of(pkg, scn, ins.drone)
        costs, noise, _ = oracle.rollout(0, scn.params, scn.cost, scn.tracks[0], scn.drones[ins.drone], opp, noms[i],
                                         ins.n_samples, ins.n_timesteps, ins.seed, 0, np.array(ins.sigma[:], np.float32),
                                         threads=16)
        assert np.array_equal(eng.costs(i).astype(np.float64), costs), f"instance {i}: costs are expected to be bit-exact"
        assert np.array_equal(eng.noise(i).view(np.uint32), noise.view(np.uint32)), f"instance {i}: noise not bit-exact"
        upd = oracle.update(0, scn.params, costs, noise, noms[i], ins.lambda_, np.array(ins.shift_fill[:], np.float32))
        one = SimpleNamespace(params=scn.params, cost=scn.cost, tracks=scn.tracks, drones=[scn.drones[ins.drone]])
        Us, Js = refine_oracle(oracle, one, opp, upd["updated"].astype(np.float32), 3, 0.05)
        say(f"  instance {i}:")
        view = SimpleNamespace(get_nominal=lambda _, i=i: eng.get_nominal(i))
        kept.append(check_refined(view, acts[i:i + 1], Us, Js)[1])
    say(f"  kept iterates {kept}")
    eng.close()


def test_executed_trajectory_cost(pkg, oracle):
    """dr_engine_set_exec_cost: after every closed-loop step each controller adds the running cost of
    its drone's new state (the other drones where they now are) and counts collisions; the progress is
    that of the latest state.  Checked step by step against the FP32 oracle's running cost and progress
    on the engine's own drone states (bit-exact), on the 8-drone workload and on two drones placed in
    collision."""
    sc = pkg.scenarios
    for name, scn in (("8 drones", sc.c4_eight_drones(512, n_timesteps=30)),
                      ("collision", sc.c4_eight_drones(256, n_timesteps=20))):
        if name == "collision":
            d0, d1 = scn.drones[0], scn.drones[1]
            d1.p[0], d1.p[1], d1.p[2] = d0.p[0] + 0.2, d0.p[1], d0.p[2]
        say(f"case: executed trajectory cost, {name}")
        eng = pkg.Engine(scn)
        eng.set_exec_cost(True)
        n = len(scn.instances)
        acc = np.zeros(n, np.float32)
        colls = np.zeros(n, np.int64)
        for step in range(6):
            eng.step(1, advance=True)
            drones = eng.get_drones()
            d = eng.diag()
            for i, ins in enumerate(scn.instances):
                me = drones[ins.drone]
                x = np.array(list(me.p) + list(me.q) + list(me.v) + list(me.w), np.float64)
                opp = []
                for j in range(len(drones)):
                    if j != ins.drone:
                        o = pkg.abi.Opponent()
                        for c in range(3):
                            o.p[c], o.v[c] = drones[j].p[c], 0.0
                        opp.append(o)
                c, term = oracle.running_cost(0, scn.params, scn.cost, scn.tracks[0], x, me.gate, me.laps, opp, 0.0)
                acc[i] = np.float32(acc[i] + np.float32(c))
                colls[i] += int(term == 1)
                prog = oracle.progress(0, scn.tracks[0], x, me.gate, me.laps)
                assert d[i]["exec_cost"] == acc[i], (name, step, i, d[i]["exec_cost"], acc[i])
                assert d[i]["exec_collisions"] == colls[i]
                assert d[i]["exec_progress"] == np.float32(prog)
        say(f"  after 6 steps: exec cost {acc.round(3).tolist()}, collisions {colls.tolist()}")
        if name == "collision":
            assert colls[0] >= 1
        eng.set_exec_cost(True)
        assert all(x["exec_cost"] == 0.0 and x["exec_collisions"] == 0 for x in eng.diag())
        eng.close()


def test_eight_drone_refined_closed_loop(pkg, oracle):
    """Three closed-loop steps of the 8-drone refined workload (K = 4096, T = 100, 3 gradient steps)
    against the oracle's closed loop with the same refinement (the CPU baseline's code): actions
    and drone states within 1e-3 relative (the refined plans agree to rounding, which the closed
    loop carries forward)."""
    sc = pkg.scenarios
    scn = sc.c4_eight_drones(4096)
    n = len(scn.instances)
    say(f"case: {scn.name}: 3 refined closed-loop steps against the oracle's")
    eng = pkg.Engine(scn)
    eng.set_refine(3, 0.05, 10)
    g_acts = eng.step(3, advance=True)
    g_drones = eng.get_drones()
    eng.close()
    drones = (pkg.abi.DroneState * n)(*scn.drones)
    o_acts, _ = oracle.world_run(scn, drones, _flat_nominals(scn), [0] * n, [i.lambda_ for i in scn.instances], 3, threads=16,
                                 refine=(3, 0.05, 10))
    e_act, e_state = rel_err(g_acts, o_acts), rel_err(_state_rows(g_drones), _state_rows(drones))
    say(f"  after 3 steps: actions max rel {e_act:.2e}, drone states max rel {e_state:.2e}")
    assert e_act <= 1e-3 and e_state <= 1e-3


def test_refined_closed_loop_launch_paths(pkg):
    """The refined closed loop moves the drones into the other copy of the drone states each step.
    Actions, nominals and drone states are the same bits whether every step is one graph replay
    (the graphs of both copies in turn), several steps share one graph (odd and even counts), or
    the kernels are launched one at a time; steps without advance move nothing; drones without a
    controller keep their velocity."""
    sc = pkg.scenarios
    scn = sc.c4_eight_drones(512, 40)
    scn.instances = scn.instances[:6]  # drones 6 and 7 fly without a controller
    dt = scn.params.dt
    v0 = _state_rows(scn.drones)[6:, 7:10]
    p0 = _state_rows(scn.drones)[6:, 0:3]
    runs = []
    for launch in (dict(use_graph=0), dict(use_graph=1, steps_per_graph=1), dict(use_graph=1, steps_per_graph=3),
                   dict(use_graph=1, steps_per_graph=2)):
        say(f"case: {scn.name} with 6 controllers, refined closed loop, launch {launch}")
        eng = pkg.Engine(scn)
        eng.set_refine(3, 0.05, 10)
        eng.set_launch(**launch)
        acts = [eng.step(1, advance=True), eng.step(4, advance=True)]
        before = _state_rows(eng.get_drones())
        acts.append(eng.step(2, advance=False))
        assert np.array_equal(_state_rows(eng.get_drones()), before)
        acts.append(eng.step(3, advance=True))
        rows = _state_rows(eng.get_drones())
        noms = [eng.get_nominal(i) for i in range(len(scn.instances))]
        eng.close()
        runs.append((acts, rows, noms))
        # 8 closed-loop steps at constant velocity (one FMA per step and axis)
        p = p0.astype(np.float32)
        for _ in range(8):
            p = (p.astype(np.float64) + np.float64(np.float32(dt)) * v0).astype(np.float32)
        assert np.allclose(rows[6:, 0:3], p, rtol=1e-6, atol=1e-6)
        assert np.array_equal(rows[6:, 7:10], v0)
    for acts, rows, noms in runs[1:]:
        for a, b in zip(acts, runs[0][0]):
            assert np.array_equal(a, b)
        assert np.array_equal(rows, runs[0][1])
        for a, b in zip(noms, runs[0][2]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("nb", [1, 2])
def test_noise_made_ahead_matches_eight_stage_rollout(pkg, nb):
    """rollout_split 4 (the refinement launch's spare workgroups make the next step's noise rows,
    the rollout reads them; chunks of 4 * noise_batch steps) against rollout_split 3 (every rollout
    makes its own): the same costs, actions, nominals and drone states bit for bit over closed-loop
    steps, steps without advance, one-step and three-step graphs, plain launches and a step
    counter set by hand; the rows in the buffer after a split-4 step are the next step's."""
    sc = pkg.scenarios
    scn = sc.c4_eight_drones(512, 42)  # T = 42: the last chunk is partial for both chunk lengths
    scn.instances = scn.instances[:6]
    n = len(scn.instances)
    say(f"case: {scn.name} with 6 controllers, refined, rollout_split 4 (noise_batch {nb}) against 3")
    ref = pkg.Engine(scn)
    nz = pkg.Engine(scn)
    for e in (ref, nz):
        e.set_refine(3, 0.05, 10)
    ref.set_launch(rollout_split=3, noise_batch=2, use_graph=1, steps_per_graph=1)
    nz.set_launch(rollout_split=4, noise_batch=nb, use_graph=1, steps_per_graph=3)

    def same(tag):
        for i in range(n):
            assert np.array_equal(ref.costs(i), nz.costs(i)), (tag, i)
            assert np.array_equal(ref.get_nominal(i), nz.get_nominal(i)), (tag, i)
        assert np.array_equal(_state_rows(ref.get_drones()), _state_rows(nz.get_drones())), tag

    for n_steps, adv in ((1, True), (4, True), (2, False), (3, True)):
        assert np.array_equal(ref.step(n_steps, advance=adv), nz.step(n_steps, advance=adv))
        same(f"{n_steps} steps advance={adv}")
    ahead = [nz.noise(i) for i in range(n)]
    ref.step(1, advance=True)
    for i in range(n):
        assert np.array_equal(ref.noise(i), ahead[i])
    nz.step(1, advance=True)
    same("after the rows made ahead")
    for i in range(n):
        ref.set_step(i, 7 + i)
        nz.set_step(i, 7 + i)
    nz.set_launch(rollout_split=4, noise_batch=nb, use_graph=0)
    assert np.array_equal(ref.step(2, advance=True), nz.step(2, advance=True))
    same("step counters set by hand, plain launches")
    nz.set_refine(0, 0.05, 10)  # without the refinement split 4 runs as split 3
    ref.set_refine(0, 0.05, 10)
    assert np.array_equal(ref.step(2, advance=True), nz.step(2, advance=True))
    same("refinement off")
    ref.close()
    nz.close()
