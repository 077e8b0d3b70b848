"""GPU parity of the plain-MPPI control step (HIP engine vs the FP32 oracle,
with the FP64 oracle's distance reported).

Bar: noise bit-exact; costs bit-exact (same FP32 operations in the same
order); updated nominal, action, min cost and nu within 1e-5 relative (the
engine sums weights in FP64 in a different order, so the FP32 increment can
differ by one rounding).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_scenario(sc, n_samples=1000, n_timesteps=37, n_drones=2, controlled=1, seed=5, track="circle"):
    tr = sc.track_circle(seed, 8, 20.0, 1.0) if track == "circle" else sc.track_lemniscate(seed, 12, 30.0, 1.5, 0.3)
    drones = sc.start_states(tr, n_drones, spacing=2.5)
    # give the uncontrolled drones a velocity so the opponent prediction matters
    for j in range(controlled, n_drones):
        drones[j].v[0], drones[j].v[1] = 1.5, -0.5
    inst = [sc.instance(n_samples, n_timesteps, 0, j, sc.seed_mix(seed, j)) for j in range(controlled)]
    return sc.Scenario("test", [tr], inst, [0] * n_drones, drones, 1, "test")


def make_engine(pkg, scn):
    eng = pkg.Engine(scn)
    eng.set_eval(True)
    return eng


def opponents_of(pkg, scn, drone):
    out = []
    for d in range(len(scn.drone_world)):
        if d != drone and scn.drone_world[d] == scn.drone_world[drone]:
            o = pkg.abi.Opponent()
            for c in range(3):
                o.p[c] = scn.drones[d].p[c]
                o.v[c] = scn.drones[d].v[c]
            out.append(o)
    return out


def oracle_step(pkg, oracle, scn, i, nominal, step=0, f64=0):
    ins = scn.instances[i]
    opp = opponents_of(pkg, scn, ins.drone)
    costs, noise, traj = oracle.rollout(f64, scn.params, scn.cost, scn.tracks[ins.track], scn.drones[ins.drone], opp,
                                        nominal, ins.n_samples, ins.n_timesteps, ins.seed, step,
                                        np.array(ins.sigma[:], dtype=np.float32))
    upd = oracle.update(f64, scn.params, costs, noise, nominal, ins.lambda_, np.array(ins.shift_fill[:], np.float32))
    return costs, noise, traj, upd


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)))) if a.size else 0.0


def check_instance(pkg, oracle, eng, scn, i, nominal_before, step=0):
    ins = scn.instances[i]
    costs, noise, traj, upd = oracle_step(pkg, oracle, scn, i, nominal_before, step)
    g_noise = eng.noise(i)
    assert np.array_equal(g_noise.view(np.uint32), noise.view(np.uint32)), "noise not bit-exact"
    g_costs = eng.costs(i)
    exact = np.mean(g_costs.astype(np.float64) == costs)
    err = rel_err(g_costs, costs)
    c64, _, _, u64 = oracle_step(pkg, oracle, scn, i, nominal_before, step, f64=1)
    err64 = rel_err(g_costs, c64)
    say(f"  inst {i}: K={ins.n_samples} T={ins.n_timesteps} costs exact={exact:.4f} max rel vs FP32={err:.2e} "
        f"vs FP64={err64:.2e} median vs FP64={np.median(np.abs(g_costs - c64) / np.maximum(1, np.abs(c64))):.2e}")
    assert err <= REL
    assert exact == 1.0, "costs are expected to be bit-exact"
    g_nom = eng.get_nominal(i)
    assert rel_err(g_nom, upd["shifted"]) <= REL
    d = eng.diag()[i]
    assert d["min_cost"] == pytest.approx(upd["min_cost"], rel=REL)
    assert d["nu"] == pytest.approx(upd["nu"], rel=REL)
    assert d["traj_cost"] == pytest.approx(traj, rel=REL, abs=1e-6)
    say(f"  inst {i}: nominal max rel vs FP32={rel_err(g_nom, upd['shifted']):.2e} vs FP64={rel_err(g_nom, u64['shifted']):.2e}")
    return upd


def test_single_step_parity(pkg, oracle):
    sc = pkg.scenarios
    say("case: single instance K=1000 T=37 with one constant-velocity opponent")
    scn = small_scenario(sc)
    eng = make_engine(pkg, scn)
    nom0 = eng.get_nominal(0).copy()
    nom0[:, 0] += 0.7
    nom0[:, 1] = 0.2
    eng.set_nominal(0, nom0)
    acts = eng.step(1)
    upd = check_instance(pkg, oracle, eng, scn, 0, nom0)
    assert rel_err(acts[0], upd["action"]) <= REL
    eng.close()


@pytest.mark.parametrize("K,T", [(1, 1), (63, 5), (64, 2), (65, 9), (129, 1), (300, 256)])
def test_ragged_and_extreme_sizes(pkg, oracle, K, T):
    sc = pkg.scenarios
    say(f"case: K={K} T={T}")
    scn = small_scenario(sc, n_samples=K, n_timesteps=T, n_drones=1)
    eng = make_engine(pkg, scn)
    nom0 = eng.get_nominal(0).copy()
    eng.step(1)
    check_instance(pkg, oracle, eng, scn, 0, nom0)
    eng.close()


def test_immediate_collision(pkg, oracle):
    """Ego starts inside an opponent's collision distance: every sample stops at t = 0."""
    sc = pkg.scenarios
    say("case: immediate collision")
    scn = small_scenario(sc, n_samples=200, n_timesteps=20, n_drones=2)
    for c in range(3):
        scn.drones[1].p[c] = scn.drones[0].p[c]
        scn.drones[1].v[c] = 0.0
    scn.drones[1].p[0] += 0.05
    eng = make_engine(pkg, scn)
    nom0 = eng.get_nominal(0).copy()
    eng.step(1)
    upd = check_instance(pkg, oracle, eng, scn, 0, nom0)
    assert np.all(eng.costs(0) > 900.0)
    assert upd["nu"] >= 1.0
    eng.close()


def test_multi_instance_world_one_step(pkg, oracle):
    sc = pkg.scenarios
    say("case: 4 drones in one world, 3 controlled (K=777, 2048, 1500; T=25, 40, 33)")
    scn = small_scenario(sc, n_samples=777, n_timesteps=25, n_drones=4, controlled=3, track="lemniscate")
    scn.instances[1] = sc.instance(2048, 40, 0, 1, sc.seed_mix(5, 1))
    scn.instances[2] = sc.instance(1500, 33, 0, 2, sc.seed_mix(5, 2))
    eng = make_engine(pkg, scn)
    noms = [eng.get_nominal(i).copy() for i in range(3)]
    eng.step(1)
    for i in range(3):
        check_instance(pkg, oracle, eng, scn, i, noms[i])
    eng.close()


def test_closed_loop_matches_oracle(pkg, oracle):
    sc = pkg.scenarios
    say("case: closed loop, 4 drones all controlled, K=1024 T=30, 6 steps")
    scn = small_scenario(sc, n_samples=1024, n_timesteps=30, n_drones=4, controlled=4, track="lemniscate")
    eng = make_engine(pkg, scn)
    n_steps = 6
    acts = eng.step(n_steps, advance=True)
    g_drones = eng.get_drones()
    drones = (pkg.abi.DroneState * 4)(*scn.drones)
    nominal = np.concatenate([np.tile(np.array(i.shift_fill[:], np.float32), (i.n_timesteps, 1)).reshape(-1)
                              for i in scn.instances]).astype(np.float32)
    steps = [0] * 4
    lams = [i.lambda_ for i in scn.instances]
    o_acts, _ = oracle.world_run(scn, drones, nominal, steps, lams, n_steps, threads=8)
    e_act = rel_err(acts, o_acts)
    g = np.array([list(d.p) + list(d.q) + list(d.v) + list(d.w) for d in g_drones])
    o = np.array([list(d.p) + list(d.q) + list(d.v) + list(d.w) for d in drones])
    e_state = rel_err(g, o)
    say(f"  closed loop after {n_steps} steps: actions max rel {e_act:.2e}, drone states max rel {e_state:.2e}")
    assert e_act <= 1e-4 and e_state <= 1e-4
    assert [d.gate for d in g_drones] == [d.gate for d in drones]
    eng.close()


def test_launch_shapes_agree(pkg):
    """Every launch shape (rollout block, persistent grid, samples per lane,
    timestep group, graph or stream, rollout structure) gives the same costs,
    noise and nominal."""
    sc = pkg.scenarios
    say("case: launch shapes on 3 uneven instances")
    scn = small_scenario(sc, n_samples=3000, n_timesteps=45, n_drones=3, controlled=3)
    scn.instances[1] = sc.instance(1111, 70, 0, 1, 99)
    scn.instances[2] = sc.instance(257, 12, 0, 2, 98)
    ref = None
    shapes = [dict(rollout_block=rb, rollout_grid=g, reduce_chunk=ch, noise_batch=nb, use_graph=ug, steps_per_graph=spg,
                   reduce_tgroup=tg, rollout_split=sp)
              for rb, g, ch, nb, ug, spg, tg, sp in [(128, 0, 256, 4, 1, 1, 25, 0), (64, 0, 64, 1, 0, 1, 256, 0),
                                                    (256, 7, 1024, 2, 1, 2, 7, 0), (128, 3, 4096, 4, 1, 3, 64, 0),
                                                    (64, 1, 192, 2, 0, 1, 1, 0), (128, 0, 1024, 1, 1, 1, 25, 1),
                                                    (128, 0, 512, 2, 0, 1, 16, 1), (128, 5, 1024, 4, 1, 2, 32, 1),
                                                    (128, 0, 1024, 1, 1, 1, 16, 2), (128, 0, 1024, 2, 0, 1, 16, 2),
                                                    (128, 9, 256, 4, 1, 2, 25, 2), (128, 0, 1024, 1, 1, 1, 16, 3),
                                                    (128, 0, 512, 2, 0, 1, 25, 3), (128, 7, 2048, 2, 1, 2, 32, 3)]]
    for shape in shapes:
        eng = make_engine(pkg, scn)
        eng.set_launch(**shape)
        eng.step(3, advance=True)
        res = ([eng.costs(i) for i in range(3)], [eng.noise(i) for i in range(3)],
               [eng.get_nominal(i) for i in range(3)], eng.actions())
        if ref is None:
            ref = res
        else:
            for a, b in zip(res[0], ref[0]):
                assert np.array_equal(a, b), shape
            for a, b in zip(res[1], ref[1]):
                assert np.array_equal(a, b), shape
            for a, b in zip(res[2], ref[2]):
                assert rel_err(a, b) <= REL, shape
        eng.close()
        say(f"  shape {shape} ok")


def test_deterministic_across_runs(pkg):
    sc = pkg.scenarios
    scn = small_scenario(sc, n_samples=2000, n_timesteps=40, n_drones=2, controlled=2)
    out = []
    for _ in range(2):
        eng = make_engine(pkg, scn)
        eng.step(4, advance=True)
        out.append(([eng.get_nominal(i) for i in range(2)], eng.actions(), [eng.costs(i) for i in range(2)]))
        eng.close()
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][1], out[1][1])


def test_full_size_four_drone_step(pkg, oracle):
    """The benchmark workload (4 drones, K=16384, T=100) checked sample by sample for one step."""
    sc = pkg.scenarios
    say("case: full size 4 x K=16384 x T=100, one step")
    scn = sc.c1_four_drones()
    eng = make_engine(pkg, scn)
    noms = [eng.get_nominal(i).copy() for i in range(4)]
    eng.step(1)
    for i in range(4):
        check_instance(pkg, oracle, eng, scn, i, noms[i])
    eng.close()


def test_single_controller_api(pkg, oracle):
    """dr_mppi_* (the reference getControl shape) equals the batch path."""
    sc = pkg.scenarios
    say("case: dr_mppi_step with 2 opponents")
    scn = small_scenario(sc, n_samples=900, n_timesteps=30, n_drones=3, controlled=1)
    ins = scn.instances[0]
    h = pkg.MPPI(scn.params, scn.cost, scn.tracks[0], ins.n_samples, ins.n_timesteps, ins.seed,
                 fill=list(ins.shift_fill), max_opponents=2)
    nom0 = np.tile(np.array(ins.shift_fill[:], np.float32), (ins.n_timesteps, 1))
    h.set_nominal(nom0)
    act = h.step(scn.drones[0], opponents_of(pkg, scn, 0))
    costs, noise, traj, upd = oracle_step(pkg, oracle, scn, 0, nom0)
    assert rel_err(act, upd["action"]) <= REL
    assert rel_err(h.get_nominal(), upd["shifted"]) <= REL
    assert h.diag()["step"] == 1
    h.close()


def test_rejects_non_unit_attitude(pkg):
    """Drone states enter with a unit attitude quaternion (the attitude recurrence renormalises to
    first order); a quaternion off by more than 1e-2 in |q|^2, or a non-finite value, is rejected
    with DR_EINVAL and leaves the engine's states unchanged."""
    sc = pkg.scenarios
    scn = small_scenario(sc, n_samples=64, n_timesteps=5, n_drones=2)
    eng = make_engine(pkg, scn)
    before = [list(d.p) + list(d.q) for d in eng.get_drones()]
    bad = (pkg.abi.DroneState * 2)(*scn.drones)
    bad[1].q[0] = 1.2
    with pytest.raises(pkg.EngineError, match="unit quaternion"):
        eng.set_drones(bad)
    bad[1].q[0] = 1.0
    bad[0].p[2] = float("nan")
    with pytest.raises(pkg.EngineError, match="not finite"):
        eng.set_drones(bad)
    assert [list(d.p) + list(d.q) for d in eng.get_drones()] == before
    ok = (pkg.abi.DroneState * 2)(*scn.drones)
    ok[1].q[0], ok[1].q[3] = 0.70710677, 0.70710677
    eng.set_drones(ok)
    eng.close()
