"""The wire codec of the reference's front-end (CPU; the library is loaded, no GPU call).

Headers are packed by drone_race_amd.wire in the dataclass field order of the front-end
(python-src/utils.py:52-66, 345-380, 443-453, 456-492, 504-532) and decoded by the
library (dr_wire_decode_*, restating env_stratrace.cpp:98-226 and config.cpp:5-115).
The byte layouts are also checked field by field at hand-computed offsets, and every
rejection the reference makes is checked to be a DR_EINVAL with its reason.
"""
import struct

import numpy as np
import pytest


@pytest.fixture(scope="module")
def wire(pkg):
    return pkg.wire


@pytest.fixture(scope="module")
def env_dict(wire):
    return wire.stratrace_env(n_models=2, seed=1234, true_theta=1)


def f32(x):
    return float(np.float32(x))


def test_env_header_layout_by_hand(wire, env_dict):
    b = wire.env_header(env_dict)
    NT = 2
    n_words = (3 + 2 + 3 + 3 + 4 + 4 + 3 + 3 + 4 + (1 + NT) + NT * 5 * 2 + 1 + 4 + 3 * (1 + 1024) + (1 + 512))
    assert len(b) == 4 * n_words
    i = lambda off: struct.unpack_from("<i", b, off)[0]  # noqa: E731
    f = lambda off: struct.unpack_from("<f", b, off)[0]  # noqa: E731
    assert (i(0), i(4), i(8), i(12), f(16)) == (0, 2, 1, 1, 2.0)          # header, ENV_STRATRACE, nModelFactors, modelSizes=[2]
    assert i(20) == 2 and i(32) == 2                                       # arenaMin, arenaMax lengths
    assert (i(44), i(48), f(52), i(56)) == (1, 2, f32(0.1), 1)             # nOppAgents, dim, dt, sendStates
    assert (f(60), f(64), f(68), f(72)) == (f32(0.4), 0.0, 0.0, 0.0)       # droneRadius, pos/speed/action noise
    assert (i(76), f(80), f(84)) == (2, 3.0, 2.0)                          # maxSpeed
    assert (i(88), f(92), f(96)) == (2, 2.0, 3.0)                          # maxAccel
    assert (i(100), f(104), i(108), i(112)) == (512, 2.0, 1, 1234)         # nTrackSamples, trackWidth, nWinLaps, seed
    assert (i(116), f(120), f(124)) == (2, 0.5, 0.5)                       # initBelief
    off = 128 + NT * 5 * 8
    assert (i(off), f(off + 4), f(off + 8), f(off + 12)) == (1, 10.0, 5.0, f32(1.9))  # trueTheta, kP, kV, maxOppLatDistFact
    assert f(off + 16) == f32(env_dict["track_length"]) and i(off + 20) == 1024    # trackLength, p_grid length
    assert (i(off + 24 + 4096), i(off + 28 + 2 * 4096), i(off + 32 + 3 * 4096)) == (1024, 1024, 512)  # dp, t, kappa lengths


def test_env_header_decodes(wire, env_dict):
    env, info = wire.decode_env(wire.env_header(env_dict))
    e = env_dict
    assert (info.seed, info.true_theta, info.send_states, info.n_models) == (1234, 1, 1, 2)
    np.testing.assert_array_equal(np.array(info.init_belief[:2]), [0.5, 0.5])
    np.testing.assert_array_equal(np.array(info.arena_min[:]), np.float32(e["arena_min"]))
    assert env.n_models == 2 and env.dt == f32(0.1) and env.drone_radius == f32(0.4)
    assert tuple(env.max_speed) == (3.0, 2.0) and tuple(env.max_accel) == (2.0, 3.0)
    assert env.track_width == 2.0 and env.n_win_laps == 1 and env.kp == 10.0 and env.kv == 5.0
    assert env.max_opp_lat_dist_fact == f32(1.9) and env.track_length == f32(e["track_length"])
    for th, (s1, s2, s3, sc, an) in enumerate(e["opps"]):
        o = env.opp[th]
        assert (o.s1, o.s2, o.s3, o.speed_scale, o.action_noise) == tuple(f32(v) for v in (s1, s2, s3, sc, an))
    np.testing.assert_array_equal(np.ctypeslib.as_array(env.p_grid), np.float32(e["p_grid"]))
    np.testing.assert_array_equal(np.ctypeslib.as_array(env.t_grid), np.float32(e["t_grid"]))
    np.testing.assert_array_equal(np.ctypeslib.as_array(env.kappa_grid), np.float32(e["kappa_grid"]))


def test_env_header_matches_scenario_env(pkg, wire, env_dict):
    """The decoded environment is the scenario builder's (same track construction)."""
    env, _ = wire.decode_env(wire.env_header(env_dict))
    ref = pkg.scenarios.sr_env(n_models=2)
    assert bytes(env) == bytes(ref)


def test_controller_headers_decode(pkg, wire):
    c = wire.decode_controller(wire.mppi_header(n_samples=4096, n_timesteps=80, n_knots=20))
    assert (c.kind, c.n_samples, c.n_timesteps, c.n_knots, c.use_splines) == (0, 4096, 80, 20, 1)
    assert list(c.knots[:20]) == pkg.scenarios.bm_knots(80, 20)
    assert (c.lambda_, c.sigma, c.min_confidence) == (10.0, 0.5, f32(0.95))
    k = c.cost
    assert (k.coll_dist_factor, k.boundary_cost, k.boundary_threshold_factor) == (f32(1.4), 50.0, f32(1.8))
    assert (k.outside_cost, k.win_cost, k.opp_win_cost, k.final_adv_weight, k.final_opp_adv_weight) == (1e7, 1e5, 1e5, 50.0, 0.0)
    assert (c.n_verif_samples, c.beta, c.verif_horizon, c.max_verif_eps) == (2 ** 17, f32(1e-5), 40, f32(0.001))
    p = wire.decode_controller(wire.prmppi_header(n_samples=65536, n_timesteps=60))
    assert (p.kind, p.n_samples, p.n_timesteps, p.P) == (1, 65536, 60, 9)
    assert (p.lambda_, p.sigma, p.gate_traversal_margin, p.safety_weight, p.delta) == (10.0, 0.5, f32(0.9), 1e4, f32(0.1))
    k = p.cost
    assert (k.coll_dist_factor, k.boundary_cost, k.boundary_threshold_factor, k.win_cost, k.opp_win_cost) == (f32(1.2), 20.0, f32(1.3),
                                                                                                           1e6, 1e5)
    # MPPI header layout: 22 words + the knot array
    assert len(wire.mppi_header(n_knots=20)) == 4 * (5 + 1 + 20 + 1 + 12 + 4)
    assert len(wire.prmppi_header()) == 4 * (4 + 13 + 1)


def test_state_header_decodes(pkg, wire, env_dict):
    s = wire.decode_state(wire.init_state_header(env_dict))
    ref = pkg.scenarios.sr_start_state(pkg.scenarios.sr_env(n_models=2))
    np.testing.assert_array_equal(np.array(s.pos[:]), np.array(ref.pos[:]))
    assert list(s.vel) == [0.0] * 4 and list(s.lat_dist) == [0.0, 0.0] and list(s.laps) == [0, 0]
    L = env_dict["track_length"]
    np.testing.assert_array_equal(np.array(s.s[:]), np.float32([26 / 512 * L, 102 / 512 * L]))


def _reject(pkg, fn, frame, needle):
    with pytest.raises(pkg.EngineError) as ei:
        fn(frame)
    assert needle in str(ei.value), str(ei.value)


def test_rejections(pkg, wire, env_dict):
    good = wire.env_header(env_dict)
    _reject(pkg, wire.decode_env, good[:-4], "underflow")
    _reject(pkg, wire.decode_env, good + b"\0\0\0\0", "trailing")
    _reject(pkg, wire.decode_env, struct.pack("<i", 1) + good[4:], "expected header")
    _reject(pkg, wire.decode_env, good[:4] + struct.pack("<i", 0) + good[8:], "environment kind")
    _reject(pkg, wire.decode_env, good[:8] + struct.pack("<i", 2) + good[12:], "model factors")
    _reject(pkg, wire.decode_env, good[:16] + struct.pack("<f", 5.0) + good[20:], "model size")
    e = dict(env_dict, n_track=500)
    _reject(pkg, wire.decode_env, wire.env_header(e), "track samples")
    e = dict(env_dict, n_opp=2)
    _reject(pkg, wire.decode_env, wire.env_header(e), "opponents")
    e = dict(env_dict, init_belief=np.array([1.0, 0.0, 0.0]))
    _reject(pkg, wire.decode_env, wire.env_header(e), "initBelief")
    e = dict(env_dict, true_theta=2)
    _reject(pkg, wire.decode_env, wire.env_header(e), "trueTheta")
    _reject(pkg, wire.decode_controller, wire.mppi_header(use_splines=False), "useSplines")
    m = wire.mppi_header(n_timesteps=30, n_knots=5)
    bad = m[:16] + struct.pack("<i", 6) + m[20:]  # nKnots 6 with 5 knots sent
    _reject(pkg, wire.decode_controller, bad, "number of knots")
    _reject(pkg, wire.decode_controller, m[:4] + struct.pack("<i", 7) + m[8:], "controller kind")
    _reject(pkg, wire.decode_state, wire.init_state_header(env_dict)[:-2], "underflow")


def test_state_frame_parser(wire):
    """decode_frame reads FullStateInfo in field order (utils.py:544-602)."""
    P = wire.Packer()
    P.i32(1)
    P.i32(7)
    for a in ([1, 2, 3, 4], [0] * 4, [5, 6], [0.1, -0.2], [0, 1], [0.3]):
        P.arr(a)
    P.arr([0.5, -0.5])
    P.arr([0.25, 0.75])
    P.i32(0)
    P.f32(0)
    P.f32(0)
    P.i32(1)
    P.f32(0)
    P.i32(2)
    for th in range(2):
        P.arr([0])
        P.arr([-1])
        P.arr([th])
        P.arr(np.zeros(3 * 4))
        P.arr(np.zeros(3 * 2))
        P.i32(3)
        P.i32(-1)
    d = wire.decode_frame(P.bytes(), 0)
    assert d["step"] == 7 and list(d["state"]["S"]) == [5, 6] and list(d["belief"]) == [0.25, 0.75]
    assert [p["predTheta"][0] for p in d["preds"]] == [0, 1] and d["preds"][1]["stopReason"] == 3
    with pytest.raises(ValueError):
        wire.decode_frame(P.bytes() + b"\0", 0)
    ev = wire.Packer()
    ev.i32(2)
    ev.i32(3)
    assert wire.decode_frame(ev.bytes(), 0) == {"type": 2, "event": 3}
