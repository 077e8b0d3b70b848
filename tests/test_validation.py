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
