"""Launch shapes and their autotuners for the PR-MPPI, PR-StratRace, StratRace
and branching (StratRace, hidden obstacle, drone race) engines.

Bar: autotuning leaves the plans, knots, beliefs and world states as they were;
the step after it equals the step of an engine left at its default launch with
noise, per-rollout costs, branches and aggregates bit-exact and plans within
1e-6 relative (the FP64 weighted sums are split into other chunks, so the FP32
results can differ by one rounding).  The branching rollout gives the same bits
for every persistent grid size.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)))) if a.size else 0.0


def same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def pr_outputs(eng, n):
    return [(eng.noise(i), eng.pair_costs(i), eng.aggregates(i), eng.candidates(i), eng.get_nominals(i)) for i in range(n)]


def check_pr(out, ref):
    for o, r in zip(out, ref):
        assert same_bits(o[0], r[0]), "noise differs"
        np.testing.assert_array_equal(o[1], r[1])
        np.testing.assert_array_equal(o[2], r[2])
        for k in range(2):
            assert rel_err(o[3][k], r[3][k]) <= REL
            assert rel_err(o[4][k], r[4][k]) <= REL


def set_pr_plans(eng, n, T, dim, seed):
    rng = np.random.default_rng(seed)
    plans = []
    for i in range(n):
        nom = rng.normal(0, 0.6, (T, dim)).astype(np.float32)
        rob = rng.normal(0, 0.6, (T, dim)).astype(np.float32)
        eng.set_nominals(i, nom, rob)
        plans.append((nom, rob))
    return plans


@pytest.mark.parametrize("kind", ["pr", "prsr"])
def test_pr_autotune_restores_and_agrees(pkg, kind):
    say(f"[{kind} autotune]")
    sc = pkg.scenarios
    if kind == "pr":
        s = sc.c3_pr_four_drones(seed=2, n_samples=700, n_timesteps=33)
        make, n, T, dim = pkg.PREngine, 4, 33, 4
    else:
        s = sc.pr_sr_race(n_inst=2, n_samples=1024, n_timesteps=30, sampled=1)
        make, n, T, dim = pkg.PRSREngine, 2, 30, 2
    outs = []
    for tune in (False, True):
        eng = make(s)
        plans = set_pr_plans(eng, n, T, dim, 5)
        if tune:
            before = eng.get_drones() if kind == "pr" else eng.get_states()
            cfg, ms = eng.autotune(trials=2)
            say(f"  chosen {cfg} at {ms:.3f} ms/step")
            assert cfg["rollout_block"] in (0, 128, 256, 512) and cfg["reduce_chunk"] in (512, 1024, 2048)
            assert cfg["reduce_tgroup"] in (16, 32) and ms > 0
            got = eng.get_launch()
            assert (got.rollout_block, got.reduce_chunk, got.reduce_tgroup) == (
                cfg["rollout_block"], cfg["reduce_chunk"], cfg["reduce_tgroup"])
            for i in range(n):
                nom, rob = eng.get_nominals(i)
                assert same_bits(nom, plans[i][0]) and same_bits(rob, plans[i][1]), "plans not restored"
            after = eng.get_drones() if kind == "pr" else eng.get_states()
            assert bytes(memoryview(after)) == bytes(memoryview(before)), "world states changed"
        eng.step(1, advance=False)
        outs.append(pr_outputs(eng, n))
        eng.close()
    check_pr(outs[1], outs[0])
    say("  plans and worlds restored; next step identical rollouts, plans within 1e-6 of the default launch")


def test_sr_autotune_restores_and_agrees(pkg):
    say("[sr autotune]")
    s = pkg.scenarios.sr_race(n_inst=2, n_samples=1500, n_timesteps=25)
    outs = []
    for tune in (False, True):
        eng = pkg.SREngine(s)
        rng = np.random.default_rng(7)
        noms = [rng.normal(0, 0.6, (25, 2)).astype(np.float32) for _ in range(2)]
        for i in range(2):
            eng.set_nominal(i, noms[i])
        if tune:
            cfg, ms = eng.autotune(trials=2)
            say(f"  chosen {cfg} at {ms:.3f} ms/step")
            assert cfg["rollout_block"] in (64, 128, 256) and cfg["reduce_chunk"] in (512, 1024, 2048, 4096) and ms > 0
            for i in range(2):
                assert same_bits(eng.get_nominal(i), noms[i]), "nominal not restored"
        eng.step(1, advance=False)
        outs.append([(eng.noise(i), eng.costs(i), eng.get_nominal(i)) for i in range(2)])
        eng.close()
    for o, r in zip(outs[1], outs[0]):
        assert same_bits(o[0], r[0])
        np.testing.assert_array_equal(o[1], r[1])
        assert rel_err(o[2], r[2]) <= REL
    say("  nominal restored; next step identical costs, nominal within 1e-6 of the default launch")


def bm_scenario(pkg, kind):
    sc = pkg.scenarios
    if kind == "sr":
        return sc.bm_race(n_inst=2, n_samples=3000, n_timesteps=40, n_knots=10)
    if kind == "ho":
        return sc.ho_race(n_inst=2, n_samples=3000, n_timesteps=40, n_knots=10)
    return sc.drace_race(n_inst=2, n_samples=3000, n_timesteps=40, n_knots=10)


def bm_outputs(eng, n):
    return [(eng.noise(i), eng.costs_true(i), eng.branches(i), eng.costs(i), eng.min_costs(i), eng.knots(i), eng.nominal(i))
            for i in range(n)]


@pytest.mark.parametrize("kind", ["sr", "ho", "drace"])
def test_bm_grid_shapes_and_autotune(pkg, kind):
    say(f"[bm launch] {kind}")
    s = bm_scenario(pkg, kind)
    make = {"sr": pkg.BMEngine, "ho": pkg.HoBMEngine, "drace": pkg.DraceBMEngine}[kind]
    n = len(s.instances)
    ref = None
    for grid in (0, 3, 64, "tune"):
        eng = make(s)
        rng = np.random.default_rng(11)
        knots = []
        for i, ins in enumerate(s.instances):
            kn = rng.normal(0, 0.8, (3, ins.n_knots, 2)).astype(np.float32)
            eng.set_knots(i, kn)
            knots.append(kn)
        if grid == "tune":
            dense = [eng.nominal(i) for i in range(n)]
            beliefs = [eng.belief(i) for i in range(n)]
            cfg, ms = eng.autotune(trials=2)
            say(f"  chosen rollout_grid {cfg['rollout_grid']} at {ms:.3f} ms/step")
            assert cfg["rollout_grid"] >= 0 and ms > 0 and eng.get_launch().rollout_grid == cfg["rollout_grid"]
            for i in range(n):
                assert same_bits(eng.knots(i), knots[i]), "knots not restored"
                assert same_bits(eng.nominal(i), dense[i]), "dense plans not restored"
                assert same_bits(eng.belief(i), beliefs[i]), "belief changed"
        else:
            eng.set_launch(rollout_grid=grid)
        eng.step(1, advance=False)
        out = bm_outputs(eng, n)
        eng.close()
        if ref is None:
            ref = out
            continue
        for o, r in zip(out, ref):
            assert same_bits(o[0], r[0]), "knot noise differs"
            np.testing.assert_array_equal(o[1], r[1])
            np.testing.assert_array_equal(o[2][0], r[2][0])
            np.testing.assert_array_equal(o[2][1], r[2][1])
            np.testing.assert_array_equal(o[3], r[3])
            np.testing.assert_array_equal(o[4], r[4])
            assert same_bits(o[5], r[5]) and same_bits(o[6], r[6])
        say(f"  grid {grid}: identical to one workgroup per tile")


@pytest.mark.parametrize("refine", [0, 3])
def test_engine_autotune_restores_and_agrees(pkg, refine):
    """The batch engine's autotuner (rollout structure, noise batch, weighted-average chunk and
    step group, then the steps per graph; with the refinement on also the rollout over noise made
    ahead) on three uneven instances: the plans, step counters and drone states are as before it,
    and the next closed-loop step agrees with an engine left at its default launch -- per-sample
    costs bit-exact, plans and actions within 1e-6 relative (other reduce chunks can move the FP32
    update by one rounding)."""
    from test_gpu_parity import small_scenario
    sc = pkg.scenarios
    scn = small_scenario(sc, n_samples=512, n_timesteps=30, n_drones=4, controlled=3, seed=9, track="lemniscate")
    scn.instances[1] = sc.instance(1000, 57, 0, 1, sc.seed_mix(9, 1))
    scn.instances[2] = sc.instance(2048, 100, 0, 2, sc.seed_mix(9, 2))
    n = len(scn.instances)
    say(f"case: batch-engine autotune, 3 uneven instances, refinement {refine}")
    tuned = pkg.Engine(scn)
    ref = pkg.Engine(scn)
    for e in (tuned, ref):
        e.set_refine(refine, 0.05, 10)
        e.step(2, advance=True)
    noms = [tuned.get_nominal(i).copy() for i in range(n)]
    steps = [d["step"] for d in tuned.diag()]
    drones = [list(d.p) + list(d.v) for d in tuned.get_drones()]
    cfg, ms = tuned.autotune(trials=2)
    say(f"  tuned launch {cfg} ({ms:.3f} ms per step)")
    for i in range(n):
        assert same_bits(tuned.get_nominal(i), noms[i])
    assert [d["step"] for d in tuned.diag()] == steps
    assert [list(d.p) + list(d.v) for d in tuned.get_drones()] == drones
    a = tuned.step(1, advance=True)
    b = ref.step(1, advance=True)
    assert rel_err(a, b) <= REL
    for i in range(n):
        np.testing.assert_array_equal(tuned.costs(i), ref.costs(i))
        assert rel_err(tuned.get_nominal(i), ref.get_nominal(i)) <= REL
    tuned.close()
    ref.close()
