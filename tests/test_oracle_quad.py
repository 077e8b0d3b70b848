"""Pins of the oracle's Quad13 model, gate logic and running cost against
closed forms and physical invariants (the model has no reference code)."""
import math

import numpy as np
import pytest


def state(p=(0, 0, 5), q=(1, 0, 0, 0), v=(0, 0, 0), w=(0, 0, 0)):
    return np.array(list(p) + list(q) + list(v) + list(w), dtype=np.float64)


@pytest.fixture(scope="module")
def P(pkg):
    return pkg.scenarios.quad_params()


@pytest.mark.parametrize("f64", [0, 1])
def test_hover_is_an_equilibrium(oracle, P, f64):
    x = state()
    for _ in range(200):
        x, _, _ = oracle.quad_step(f64, P, x, [P.mass * P.gravity, 0, 0, 0])
    assert np.array_equal(x, state())


def test_free_fall_with_linear_drag(oracle, P):
    # v_{n+1} = v_n + dt (-g - c v_n)  =>  v_n = -(g/c) (1 - (1 - c dt)^n)
    x = state(p=(0, 0, 100))
    g, c, dt = P.gravity, P.drag, float(np.float32(P.dt))
    z = 100.0
    for n in range(1, 101):
        vprev = -(g / c) * (1 - (1 - c * dt) ** (n - 1))
        z += dt * vprev
        x, _, _ = oracle.quad_step(1, P, x, [0, 0, 0, 0])
        assert abs(x[9] - (-(g / c) * (1 - (1 - c * dt) ** n))) < 1e-9
        assert abs(x[2] - z) < 1e-9


@pytest.mark.parametrize("f64,tol", [(0, 1e-6), (1, 1e-13)])
def test_quaternion_stays_unit(oracle, P, f64, tol):
    rng = np.random.default_rng(0)
    x = state()
    for _ in range(300):
        u = [P.mass * P.gravity, *rng.uniform(-6, 6, 3)]
        x, _, _ = oracle.quad_step(f64, P, x, u)
        assert abs(np.linalg.norm(x[3:7]) - 1.0) <= tol


def test_zero_rate_keeps_attitude(oracle, P):
    q0 = np.array([0.9, 0.1, -0.3, 0.2])
    q0 /= np.linalg.norm(q0)
    x = state(q=q0)
    for _ in range(50):
        x, _, _ = oracle.quad_step(1, P, x, [5.0, 0, 0, 0])
    assert np.max(np.abs(x[3:7] - q0)) < 1e-14


def test_body_rate_first_order_tracking(oracle, P):
    # w_n = u + (w_0 - u)(1 - k)^n with k = dt * rate_gain
    k = float(np.float32(P.dt * P.rate_gain))
    u = np.array([1.0, -2.0, 0.5])
    w0 = np.array([0.3, 0.1, -0.4])
    x = state(w=w0)
    for n in range(1, 30):
        x, _, _ = oracle.quad_step(1, P, x, [P.mass * P.gravity, *u])
        assert np.max(np.abs(x[10:13] - (u + (w0 - u) * (1 - k) ** n))) < 1e-12


def test_yaw_rate_keeps_hover_and_turns_by_euler_angle(oracle, P):
    # rotation about body z leaves the thrust axis vertical; each step turns the
    # attitude by 2 atan(dt/2 * wz) once the body rate has converged
    wz = 2.0
    x = state(w=(0, 0, wz))
    for _ in range(40):
        x, _, _ = oracle.quad_step(1, P, x, [P.mass * P.gravity, 0, 0, wz])
    assert np.max(np.abs(x[7:10])) < 1e-12
    yaw = 2 * math.atan2(x[6], x[3])
    expect = 40 * 2 * math.atan(float(np.float32(0.5 * np.float32(P.dt))) * wz)
    assert abs(math.remainder(yaw - expect, 2 * math.pi)) < 1e-12


def test_thrust_direction_rolls_with_attitude(oracle, P):
    phi = 0.3
    q = (math.cos(phi / 2), math.sin(phi / 2), 0.0, 0.0)  # roll about x: b3 = (0, -sin phi, cos phi)
    x = state(q=q)
    f = 12.0
    x1, _, _ = oracle.quad_step(1, P, x, [f, 0, 0, 0])
    dt = float(np.float32(P.dt))
    a = (x1[7:10] - x[7:10]) / dt
    inv_m = float(np.float32(1.0 / np.float32(P.mass)))
    np.testing.assert_allclose(a, [0.0, -f * inv_m * math.sin(phi), f * inv_m * math.cos(phi) - P.gravity], atol=1e-12)


def test_controls_are_clamped(oracle, P):
    x = state()
    x1, _, _ = oracle.quad_step(1, P, x, [1e3, 50, -50, 0])
    x2, _, _ = oracle.quad_step(1, P, x, [P.thrust_max, P.rate_max, -P.rate_max, 0])
    assert np.array_equal(x1, x2)
    x3, _, _ = oracle.quad_step(1, P, x, [-5.0, 0, 0, 0])
    x4, _, _ = oracle.quad_step(1, P, x, [P.thrust_min, 0, 0, 0])
    assert np.array_equal(x3, x4)


def one_gate_track(pkg, center=(0, 0, 0), normal=(1, 0, 0), radius=1.0):
    abi = pkg.abi
    tr = abi.Track()
    tr.n_gates = 2
    tr.n_win_laps = 100
    tr.ground_z = -100.0
    for i, c in enumerate([center, (50.0, 0.0, 0.0)]):
        for d in range(3):
            tr.gates[i].center[d] = c[d]
            tr.gates[i].normal[d] = normal[d]
        tr.gates[i].radius = radius
    return tr


def crossing(oracle, P, tr, p0, v, gate=1, laps=0, f64=1):
    # a drag-free, gravity-compensated step moves p by dt*v exactly enough for the test
    x = state(p=p0, v=v)
    P2 = type(P)(**{f: getattr(P, f) for f, _ in P._fields_})
    P2.drag = 0.0
    _, g, l = oracle.quad_step(f64, P2, x, [P.mass * P.gravity, 0, 0, 0], track=tr, margin=0.95, gate=gate, laps=laps)
    return g, l


def test_gate_crossing_cases(oracle, P, pkg):
    tr = one_gate_track(pkg)
    dt = P.dt
    # next gate is gate 0 (current gate 1); crossing at y=0.5 inside radius*margin
    assert crossing(oracle, P, tr, (-0.05, 0.5, 0), (5.0, 0, 0)) == (0, 1)   # passing gate 0 counts a lap
    # both directions count
    assert crossing(oracle, P, tr, (0.05, 0.5, 0), (-5.0, 0, 0)) == (0, 1)
    # outside radius*margin (0.95): no crossing at y = 0.96
    assert crossing(oracle, P, tr, (-0.05, 0.96, 0), (5.0, 0, 0)) == (1, 0)
    # inside radius but outside the margin band is a miss, inside the band a hit
    assert crossing(oracle, P, tr, (-0.05, 0.94, 0), (5.0, 0, 0)) == (0, 1)
    # segment that stops short of the plane
    assert crossing(oracle, P, tr, (-0.2, 0.0, 0), (5.0, 0, 0)) == (1, 0)
    # motion parallel to the gate plane
    assert crossing(oracle, P, tr, (0.0, 0.0, 0), (0.0, 5.0, 0)) == (1, 0)
    # gate 1 is next when the current gate is 0: no lap increment
    tr2 = one_gate_track(pkg)
    assert crossing(oracle, P, tr2, (49.95, 0.0, 0), (5.0, 0, 0), gate=0, laps=3) == (1, 3)
    assert dt > 0


def test_progress_closed_form(oracle, pkg):
    tr = one_gate_track(pkg)
    at_next = state(p=(50.0, 0, 0))
    at_gate = state(p=(0.0, 0, 0))
    mid = state(p=(25.0, 0, 0))
    assert oracle.progress(1, tr, at_next, 0, 2) == pytest.approx(0 + 2 * 2 + 0.8, abs=1e-7)
    assert oracle.progress(1, tr, at_gate, 0, 0) == pytest.approx(0.0, abs=1e-7)
    assert oracle.progress(1, tr, mid, 0, 0) == pytest.approx(0.4, abs=1e-7)


def test_running_cost_closed_forms(oracle, pkg, P):
    abi = pkg.abi
    K = pkg.scenarios.cost_cfg()
    tr = one_gate_track(pkg)
    tr.ground_z = 0.0
    r, rp = P.drone_radius, K.opp_prox_radius
    danger = K.boundary_threshold_factor * r

    def opp(p, v=(0, 0, 0)):
        o = abi.Opponent()
        for d in range(3):
            o.p[d], o.v[d] = p[d], v[d]
        return o

    x = state(p=(0, 0, 4))
    # no opponents, high up: zero cost
    c, t = oracle.running_cost(1, P, K, tr, x, 0, 0, [], 0.1)
    assert (c, t) == (0.0, 0)
    # one opponent at distance 1 (predicted there at tau = 0.5 from p - v tau)
    c, t = oracle.running_cost(1, P, K, tr, x, 0, 0, [opp((1.0, -0.5, 4), (0, 1.0, 0))], 0.5)
    assert t == 0 and c == pytest.approx(K.opp_prox_cost * (1 - 1.0 / rp) ** 2, rel=1e-6)
    # opponent inside the collision distance: proximity + near-boundary + outside, terminal
    d = 0.1
    c, t = oracle.running_cost(1, P, K, tr, x, 0, 0, [opp((d, 0, 4))], 0.0)
    bd = d - r
    expect = K.opp_prox_cost * (1 - d / rp) ** 2 + K.boundary_cost * (1 - bd / danger) ** 2 + K.outside_cost
    assert t == 1 and c == pytest.approx(expect, rel=1e-6)
    # near the ground, outside the collision distance: near-boundary only
    z = 0.2
    c, t = oracle.running_cost(1, P, K, tr, state(p=(0, 0, z)), 0, 0, [], 0.0)
    assert t == 0 and c == pytest.approx(K.boundary_cost * (1 - z / danger) ** 2, rel=1e-6)
    # below the collision distance from the ground
    c, t = oracle.running_cost(1, P, K, tr, state(p=(0, 0, 0.1)), 0, 0, [], 0.0)
    assert t == 1 and c == pytest.approx(K.boundary_cost * (1 - 0.1 / danger) ** 2 + K.outside_cost, rel=1e-6)
    # winner
    c, t = oracle.running_cost(1, P, K, tr, x, 0, tr.n_win_laps, [], 0.0)
    assert t == 2 and c == -K.win_cost


def test_inverse_square_root_of_the_renormalisation(oracle):
    """The FP32 1/sqrt of the attitude renormalisation (bit-pattern estimate + three Newton steps)
    against libm in FP64: within 3 units in the last place over a dense sweep of [0.25, 4] (where
    |q + dt/2 q w|^2 lies) and random normal numbers over 2^-60 .. 2^60; 1/sqrt(1) exactly 1, and
    exact powers of four give exact results (1/sqrt(4) = 0.5)."""
    rng = np.random.default_rng(3)
    xs = np.concatenate([np.linspace(0.25, 4.0, 20001), 2.0 ** rng.uniform(-60, 60, 5000)]).astype(np.float32)
    got = np.array([oracle.inv_sqrt32(float(x)) for x in xs], np.float32)
    ref = 1.0 / np.sqrt(xs.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    err = np.abs(got.astype(np.float64) - ref) / ulp
    assert err.max() <= 3.0, err.max()
    assert oracle.inv_sqrt32(1.0) == 1.0
    assert oracle.inv_sqrt32(4.0) == 0.5 and oracle.inv_sqrt32(0.25) == 2.0
