"""Argument validation of the C ABI.  Every case is rejected on the host
before any device call, so these run without a GPU."""
import ctypes as C

import pytest


def create(pkg, sc_mod, mutate):
    abi = pkg.abi
    lib = abi.load()
    tr = sc_mod.track_circle(0)
    tracks = (abi.Track * 1)(tr)
    inst = sc_mod.instance(128, 10, 0, 0, 1)
    world = [0]
    params, cost = sc_mod.quad_params(), sc_mod.cost_cfg()
    n_drones = 1
    ctx = dict(inst=inst, world=world, params=params, cost=cost, tracks=tracks, n_drones=n_drones, n_tracks=1)
    mutate(ctx)
    insts = (abi.Instance * 1)(ctx["inst"])
    w = (C.c_int32 * len(ctx["world"]))(*ctx["world"])
    h = C.c_void_p()
    rc = lib.dr_engine_create(C.byref(ctx["params"]), C.byref(ctx["cost"]), ctx["tracks"], ctx["n_tracks"], insts, 1, w,
                              ctx["n_drones"], C.byref(h))
    return rc, lib.dr_last_error().decode(), h


CASES = {
    "zero samples": lambda c: setattr(c["inst"], "n_samples", 0),
    "horizon too long": lambda c: setattr(c["inst"], "n_timesteps", 257),
    "zero horizon": lambda c: setattr(c["inst"], "n_timesteps", 0),
    "bad track index": lambda c: setattr(c["inst"], "track", 1),
    "bad drone index": lambda c: setattr(c["inst"], "drone", 3),
    "non-positive temperature": lambda c: setattr(c["inst"], "lambda_", 0.0),
    "non-positive mass": lambda c: setattr(c["params"], "mass", 0.0),
    "inverted thrust limits": lambda c: setattr(c["params"], "thrust_min", 50.0),
    "one-gate track": lambda c: setattr(c["tracks"][0], "n_gates", 1),
    "too many gates": lambda c: setattr(c["tracks"][0], "n_gates", 17),
    "zero gate radius": lambda c: setattr(c["tracks"][0].gates[2], "radius", 0.0),
    "zero proximity radius": lambda c: setattr(c["cost"], "opp_prox_radius", 0.0),
    "world of nine drones": lambda c: c.update(world=[0] * 9, n_drones=9),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_create_rejects(pkg, name):
    rc, msg, h = create(pkg, pkg.scenarios, CASES[name])
    assert rc == pkg.abi.DR_EINVAL, (rc, msg)
    assert msg
    assert not h.value


def test_null_handles_are_rejected(pkg):
    lib = pkg.abi.load()
    assert lib.dr_engine_step(None, 1, 0, None) == pkg.abi.DR_EINVAL
    assert lib.dr_engine_get_costs(None, 0, None, 0) == pkg.abi.DR_EINVAL
    assert lib.dr_mppi_step(None, None, None, 0, None) == pkg.abi.DR_EINVAL
    lib.dr_engine_destroy(None)
    lib.dr_mppi_destroy(None)


def create_pr(pkg, sc_mod, mutate):
    abi = pkg.abi
    lib = abi.load()
    tr = sc_mod.track_circle(0)
    tracks = (abi.Track * 1)(tr)
    import numpy as np
    inst = sc_mod.pr_instance(128, 10, 0, 0, 1, [0.5, 0.5], np.zeros((2, 1, 3), np.float32))
    ctx = dict(inst=inst, world=[0, 0], n_drones=2)
    mutate(ctx)
    insts = (abi.PrInstance * 1)(ctx["inst"])
    w = (C.c_int32 * len(ctx["world"]))(*ctx["world"])
    h = C.c_void_p()
    rc = lib.dr_pr_create(C.byref(sc_mod.quad_params()), C.byref(sc_mod.cost_cfg()), tracks, 1, insts, 1, w, ctx["n_drones"],
                          C.byref(h))
    return rc, lib.dr_last_error().decode(), h


PR_CASES = {
    "no branches": lambda c: setattr(c["inst"], "n_branches", 0),
    "nine branches": lambda c: setattr(c["inst"], "n_branches", 9),
    "sampled without models": lambda c: (setattr(c["inst"], "sampled", 1), setattr(c["inst"], "n_models", 0)),
    "seventeen models": lambda c: (setattr(c["inst"], "sampled", 1), setattr(c["inst"], "n_models", 17)),
    "sampled flag 2": lambda c: setattr(c["inst"], "sampled", 2),
    "negative belief": lambda c: c["inst"].belief.__setitem__(1, -0.1),
    "nan belief": lambda c: c["inst"].belief.__setitem__(0, float("nan")),
    "infinite safety weight": lambda c: setattr(c["inst"], "safety_weight", float("inf")),
    "zero samples": lambda c: setattr(c["inst"], "n_samples", 0),
    "horizon too long": lambda c: setattr(c["inst"], "n_timesteps", 257),
    "non-positive temperature": lambda c: setattr(c["inst"], "lambda_", -1.0),
    "bad drone index": lambda c: setattr(c["inst"], "drone", 2),
}


@pytest.mark.parametrize("name", sorted(PR_CASES))
def test_pr_create_rejects(pkg, name):
    rc, msg, h = create_pr(pkg, pkg.scenarios, PR_CASES[name])
    assert rc == pkg.abi.DR_EINVAL, (rc, msg)
    assert msg
    assert not h.value


def test_pr_null_handles_are_rejected(pkg):
    lib = pkg.abi.load()
    assert lib.dr_pr_step(None, 1, 0, None) == pkg.abi.DR_EINVAL
    assert lib.dr_pr_get_aggregates(None, 0, None, 0) == pkg.abi.DR_EINVAL
    assert lib.dr_pr_set_belief(None, 0, None, 0) == pkg.abi.DR_EINVAL
    assert lib.dr_pr_set_launch(None, None) == pkg.abi.DR_EINVAL
    lib.dr_pr_destroy(None)


# ---------------------------------------------------------------- branching engine (all environments)

def bm_create(pkg, kind, mutate_env=None, mutate_inst=None):
    abi, sc = pkg.abi, pkg.scenarios
    lib = abi.load()
    if kind == "stratrace":
        env, cost, fn = sc.sr_env(n_models=2), sc.bm_cost(), lib.dr_bm_create
    elif kind == "hiddenobs":
        env, cost, fn = sc.ho_env(), sc.ho_cost(), lib.dr_bmho_create
    else:
        env, cost, fn = sc.drace_env(), sc.drace_cost(), lib.dr_bmdrace_create
    ins = sc.bm_instance(256, 30, 1, n_knots=6)
    if mutate_env:
        mutate_env(env)
    if mutate_inst:
        mutate_inst(ins)
    arr = (abi.BmInstance * 1)(ins)
    h = C.c_void_p()
    rc = fn(C.byref(env), C.byref(cost), arr, 1, C.byref(h))
    return rc, lib.dr_last_error().decode(), h


INST_CASES = {
    "knots not increasing": lambda i: setattr(i, "knots", (C.c_int32 * 60)(0, 5, 5, 10, 15, 20)),
    "knot beyond the horizon": lambda i: i.knots.__setitem__(5, 30),
    "too many knots x steps": lambda i: (setattr(i, "n_timesteps", 256), setattr(i, "n_knots", 40)),
    "true model out of range": lambda i: setattr(i, "true_theta", 2),
    "negative belief": lambda i: i.init_belief.__setitem__(0, -0.1),
    "verification horizon beyond T": lambda i: (setattr(i, "n_verif_samples", 128), setattr(i, "verif_horizon", 31),
                                                setattr(i, "beta", 1e-5)),
    "verification beta of 1": lambda i: (setattr(i, "n_verif_samples", 128), setattr(i, "verif_horizon", 10), setattr(i, "beta", 1.0)),
    "zero temperature": lambda i: setattr(i, "lambda_", 0.0),
}


@pytest.mark.parametrize("kind", ["stratrace", "hiddenobs", "dronerace"])
@pytest.mark.parametrize("name", list(INST_CASES))
def test_branching_create_rejects_instances(pkg, kind, name):
    rc, msg, h = bm_create(pkg, kind, mutate_inst=INST_CASES[name])
    assert rc == pkg.abi.DR_EINVAL and msg and not h.value, (rc, msg)


def _nan(arr, i):
    arr[i] = float("nan")


ENV_CASES = {
    "hiddenobs": {
        "negative noise": lambda e: setattr(e, "pos_noise", -0.1),
        "zero step": lambda e: setattr(e, "dt", 0.0),
        "NaN geometry": lambda e: _nan(e.hidden_radius, 1),
        "repeated gate": lambda e: (e.gate_centers[1].__setitem__(0, 12.0), e.gate_centers[1].__setitem__(1, 7.0)),
        "no laps to win": lambda e: setattr(e, "n_win_laps", 0),
    },
    "dronerace": {
        "nine gates": lambda e: setattr(e, "n_gates", 9),
        "five racelines": lambda e: setattr(e, "n_racelines", 5),
        "raceline of a model missing": lambda e: setattr(e.opp_pid[1], "raceline_index", 2),
        "controlled agent 2": lambda e: setattr(e, "i_mppi", 2),
        "five obstacles": lambda e: setattr(e, "n_obstacles", 5),
        "NaN track point": lambda e: _nan(e.track_points[1][100], 0),
        "zero max speed": lambda e: e.max_speed.__setitem__(1, 0.0),
        "negative PD noise": lambda e: setattr(e.opp_pid[0], "action_noise", -1.0),
    },
}


@pytest.mark.parametrize("kind,name", [(k, n) for k, cases in ENV_CASES.items() for n in cases])
def test_branching_create_rejects_environments(pkg, kind, name):
    rc, msg, h = bm_create(pkg, kind, mutate_env=ENV_CASES[kind][name])
    assert rc == pkg.abi.DR_EINVAL and msg and not h.value, (rc, msg)


def test_cost_grad_wrappers_check_shapes_before_any_call(pkg):
    """Engine.cost_grad rejects a U whose shape is not (n, T_i, 4) before it
    calls the library (the engine object is never created: no GPU needed)."""
    import numpy as np
    sc = pkg.scenarios
    scn = sc.Scenario("v", [sc.track_circle(0)], [sc.instance(128, 10, 0, 0, 1)], [0], sc.start_states(sc.track_circle(0), 1),
                      1, "validation")
    eng = object.__new__(pkg.Engine)
    eng.sc, eng.n_inst, eng.h, eng.lib = scn, 1, None, None
    for shape in ((3, 9, 4), (3, 11, 4), (3, 10, 3), (10, 4), (0, 10, 4)):
        with pytest.raises(ValueError):
            eng.cost_grad(0, np.zeros(shape, np.float32))
    with pytest.raises(IndexError):
        eng.cost_grad(1, np.zeros((1, 10, 4), np.float32))
