"""Workloads of the benchmark configurations, built from seeded Philox draws.

Choices the configurations leave open, fixed here:
  quadrotor    mass 1 kg, g 9.81, linear drag 0.1/s, body-rate gain 20/s
               (tau 50 ms), thrust in [0, 30] N, rates in [-6, 6] rad/s,
               radius 0.15 m, Euler dt 0.02 s
  gates        altitude 4 m before jitter; the jitter applies to x, y and z;
               normal = horizontal track tangent rotated by the yaw jitter
  lemniscate   Bernoulli, x = a cos s / (1 + sin^2 s), y = a sin s cos s /
               (1 + sin^2 s), a = 30 m, gates at s = 2 pi (i + 1/2) / n
  start        drones hover 35% of the way from the last gate to gate 0,
               spread 2.5 m apart across the approach, gate index n - 1
  nominal      hover thrust (m g) and zero rates; the shift fills the freed
               last step with the hover control (the reference fills zero,
               mppi_controller.cu:430-436; shift_fill = 0 restores that)
  costs        MPPIConfig defaults where a field exists (config.h:116-154);
               opponent proximity radius 2 m, weight 50; final progress
               weight 10
"""
import math

import numpy as np

from . import abi
from .philox import uniforms, normals, STREAM_TRACK, STREAM_BRANCH

HOVER_FILL = None  # set from the quad params below


def quad_params():
    return abi.QuadParams(mass=1.0, gravity=9.81, drag=0.1, rate_gain=20.0, thrust_min=0.0, thrust_max=30.0,
                          rate_max=6.0, drone_radius=0.15, dt=0.02)


def hover_control(p=None):
    p = p or quad_params()
    return [p.mass * p.gravity, 0.0, 0.0, 0.0]


def cost_cfg(opp_prox_radius=2.0):
    return abi.CostCfg(gate_traversal_margin=0.95, coll_dist_factor=1.0, boundary_cost=10.0,
                       boundary_threshold_factor=1.5, outside_cost=1000.0, win_cost=1000.0, opp_win_cost=1000.0,
                       final_adv_weight=10.0, final_opp_adv_weight=0.0, opp_prox_radius=opp_prox_radius,
                       opp_prox_cost=50.0, _pad=0.0)


def _make_track(centers, tangents, radius, n_win_laps=100, ground_z=0.0):
    tr = abi.Track()
    n = len(centers)
    assert 2 <= n <= abi.MAX_GATES
    tr.n_gates = n
    tr.n_win_laps = n_win_laps
    tr.ground_z = ground_z
    for i in range(n):
        g = tr.gates[i]
        for d in range(3):
            g.center[d] = float(centers[i][d])
            g.normal[d] = float(tangents[i][d])
        g.radius = float(radius)
    return tr


def track_circle(seed, n_gates=8, radius=20.0, jitter=1.0, gate_radius=1.0, altitude=4.0, salt=0):
    """Gates at angles 2 pi i / n on a circle, centres + U(-jitter, jitter) per axis."""
    u = uniforms(seed, 3 * n_gates, STREAM_TRACK, salt)
    centers, tangents = [], []
    for i in range(n_gates):
        a = 2.0 * math.pi * i / n_gates
        c = np.array([radius * math.cos(a), radius * math.sin(a), altitude])
        c = c + jitter * (2.0 * u[3 * i:3 * i + 3] - 1.0)
        centers.append(c)
        tangents.append([-math.sin(a), math.cos(a), 0.0])
    return _make_track(centers, tangents, gate_radius)


def track_lemniscate(seed, n_gates=12, half_width=30.0, jitter=1.5, yaw_jitter=0.3, gate_radius=1.0, altitude=4.0, salt=0):
    """Bernoulli lemniscate figure-eight; position jitter U(-j, j) per axis, yaw jitter U(-y, y)."""
    u = uniforms(seed, 4 * n_gates, STREAM_TRACK, salt)
    a = half_width
    centers, tangents = [], []
    for i in range(n_gates):
        s = 2.0 * math.pi * (i + 0.5) / n_gates
        den = 1.0 + math.sin(s) ** 2
        x = a * math.cos(s) / den
        y = a * math.sin(s) * math.cos(s) / den
        h = 1e-6
        den2 = 1.0 + math.sin(s + h) ** 2
        dx = (a * math.cos(s + h) / den2 - x) / h
        dy = (a * math.sin(s + h) * math.cos(s + h) / den2 - y) / h
        yaw = math.atan2(dy, dx) + yaw_jitter * (2.0 * u[4 * i + 3] - 1.0)
        c = np.array([x, y, altitude]) + jitter * (2.0 * u[4 * i:4 * i + 3] - 1.0)
        centers.append(c)
        tangents.append([math.cos(yaw), math.sin(yaw), 0.0])
    return _make_track(centers, tangents, gate_radius)


def start_states(track, n_drones, spacing=2.5, frac=0.35):
    """Hovering drones on the approach to gate 0, spread across it."""
    n = track.n_gates
    g_prev = np.array(track.gates[n - 1].center[:])
    g0 = np.array(track.gates[0].center[:])
    base = g_prev + frac * (g0 - g_prev)
    d = g0 - g_prev
    lat = np.array([-d[1], d[0], 0.0])
    lat /= max(np.linalg.norm(lat), 1e-9)
    out = (abi.DroneState * n_drones)()
    for j in range(n_drones):
        off = (j - 0.5 * (n_drones - 1)) * spacing
        p = base + off * lat
        s = out[j]
        for i in range(3):
            s.p[i] = float(p[i])
            s.v[i] = 0.0
            s.w[i] = 0.0
        s.q[0], s.q[1], s.q[2], s.q[3] = 1.0, 0.0, 0.0, 0.0
        s.gate = n - 1
        s.laps = 0
    return out


def instance(n_samples, n_timesteps, track, drone, seed, lam=1.0, sigma=(2.0, 0.5, 0.5, 0.5), fill=None, adapt=0):
    fill = hover_control() if fill is None else fill
    return abi.Instance(n_samples=n_samples, n_timesteps=n_timesteps, track=track, drone=drone, seed=seed,
                        lambda_=lam, sigma=(abi.C.c_float * 4)(*sigma), shift_fill=(abi.C.c_float * 4)(*fill),
                        adapt_lambda=adapt)


class Scenario:
    """Everything dr_engine_create needs, plus the run length."""

    def __init__(self, name, tracks, instances, drone_world, drones, steps, description):
        self.name = name
        self.tracks = tracks
        self.instances = instances
        self.drone_world = drone_world
        self.drones = drones
        self.steps = steps
        self.description = description
        self.params = quad_params()
        self.cost = cost_cfg()

    @property
    def sample_steps_per_step(self):
        return sum(i.n_samples * i.n_timesteps for i in self.instances)


def c0_single(seed=0, n_samples=1024, n_timesteps=50, steps=500):
    """Plain MPPI, 1 quadrotor, K=1024, T=50, 8 gates on r=20 m, U(-1,1) jitter, seed 0."""
    tr = track_circle(seed, 8, 20.0, 1.0)
    drones = start_states(tr, 1)
    inst = [instance(n_samples, n_timesteps, 0, 0, seed_mix(seed, 0))]
    return Scenario("c0_single_quad_k1024_t50", [tr], inst, [0], drones, steps,
                    "1 quadrotor, K=1024, T=50, 8-gate circle r=20 m, seed 0")


def c1_four_drones(seed=1, n_samples=16384, n_timesteps=100, steps=1000):
    """Plain MPPI, 4 drones racing, K=16384, T=100, 12-gate lemniscate a=30 m, seed 1."""
    tr = track_lemniscate(seed, 12, 30.0, 1.5, 0.3)
    drones = start_states(tr, 4)
    inst = [instance(n_samples, n_timesteps, 0, j, seed_mix(seed, j)) for j in range(4)]
    return Scenario("c1_4drones_k16384_t100", [tr], inst, [0, 0, 0, 0], drones, steps,
                    "4 drones, one controller each, K=16384, T=100, 12-gate lemniscate a=30 m, seed 1")


def c2_batch64(seed=3, n_inst=64, steps=200):
    """64 single-drone controllers: K log-uniform [1024, 32768], T uniform {30..150},
    each on its own seeded 8-12 gate circle of radius U(15, 40) m."""
    u = uniforms(seed, 4 * n_inst, STREAM_TRACK, salt=7)
    tracks, inst, worlds, drones_all = [], [], [], []
    for i in range(n_inst):
        k = int(round(math.exp(math.log(1024) + u[4 * i] * (math.log(32768) - math.log(1024)))))
        t = 30 + int(u[4 * i + 1] * 121)
        ng = 8 + int(u[4 * i + 2] * 5)
        r = 15.0 + 25.0 * u[4 * i + 3]
        tr = track_circle(seed_mix(seed, 1000 + i), ng, r, 1.0)
        tracks.append(tr)
        inst.append(instance(k, t, i, i, seed_mix(seed, i)))
        worlds.append(i)
        drones_all.append(start_states(tr, 1)[0])
    drones = (abi.DroneState * n_inst)(*drones_all)
    return Scenario("c2_batch64_uneven", tracks, inst, worlds, drones, steps,
                    "64 single-drone controllers, K~logU[1024,32768], T~U{30..150}, own 8-12 gate circle r~U(15,40)")


def c4_eight_drones(n_samples=4096, n_timesteps=100, seed=1):
    """8 drones, one controller each, on the 12-gate lemniscate of c1 (seed 1)."""
    tr = track_lemniscate(seed, 12, 30.0, 1.5, 0.3)
    drones = start_states(tr, 8, spacing=2.5)
    inst = [instance(n_samples, n_timesteps, 0, j, seed_mix(seed, 100 + j)) for j in range(8)]
    return Scenario(f"c4_8drones_k{n_samples}_t{n_timesteps}", [tr], inst, [0] * 8, drones, 500,
                    f"8 drones, one controller each, K={n_samples}, T={n_timesteps}, 12-gate lemniscate a=30 m, seed {seed}")


def pr_instance(n_samples, n_timesteps, track, drone, seed, belief, vel_shift, lam=1.0, sigma=(2.0, 0.5, 0.5, 0.5), fill=None,
                adapt=0, sampled=0, n_models=0, safety_weight=10000.0):
    """One PR-MPPI controller: belief[B], vel_shift[B][opponent slot][3] (opponent slots are the
    other drones of the world in increasing drone index)."""
    fill = hover_control() if fill is None else fill
    ins = abi.PrInstance(n_samples=n_samples, n_timesteps=n_timesteps, track=track, drone=drone, seed=seed, lambda_=lam,
                         sigma=(abi.C.c_float * 4)(*sigma), shift_fill=(abi.C.c_float * 4)(*fill), adapt_lambda=adapt,
                         n_branches=len(belief), sampled=sampled, n_models=n_models, safety_weight=safety_weight)
    for b, p in enumerate(belief):
        ins.belief[b] = float(p)
    vs = np.asarray(vel_shift, dtype=np.float32)
    for b in range(vs.shape[0]):
        for j in range(vs.shape[1]):
            for c in range(3):
                ins.vel_shift[b][j][c] = float(vs[b, j, c])
    return ins


def branch_draws(seed, i, n_opp, b_min=3, b_max=6):
    """B ~ U{b_min..b_max}, belief ~ Dirichlet(1) (normalised Exp(1) draws) and per-branch
    opponent-velocity shifts ~ N(0, 1) m/s per opponent and axis, from Philox stream 2."""
    u = uniforms(seed, 1 + abi.MAX_BRANCHES, STREAM_BRANCH, salt=2 * i)
    B = b_min + int(u[0] * (b_max - b_min + 1))
    e = -np.log1p(-u[1:1 + B])
    belief = (e / e.sum()).astype(np.float32)
    z = normals(seed, B * max(n_opp, 1) * 3, STREAM_BRANCH, salt=2 * i + 1)
    shift = z.reshape(B, max(n_opp, 1), 3)[:, :n_opp, :].astype(np.float32)
    return B, belief, shift


class PrScenario(Scenario):
    """Scenario of PR-MPPI controllers (dr_pr_create)."""


def c3_pr_four_drones(seed=2, n_samples=8192, n_timesteps=80, steps=500, sampled=0, n_models=0):
    """PR-MPPI, 4 drones on the 12-gate lemniscate (a = 30 m) generated with the seed; each
    controller holds B ~ U{3..6} opponent-velocity branches with a Dirichlet(1) belief; K samples
    evaluated under every branch, T steps."""
    tr = track_lemniscate(seed, 12, 30.0, 1.5, 0.3)
    drones = start_states(tr, 4)
    inst = []
    for j in range(4):
        B, belief, shift = branch_draws(seed, j, 3)
        inst.append(pr_instance(n_samples, n_timesteps, 0, j, seed_mix(seed, 200 + j), belief, shift, sampled=sampled,
                                n_models=n_models))
    return PrScenario(f"c3_pr_4drones_k{n_samples}_t{n_timesteps}", [tr], inst, [0, 0, 0, 0], drones, steps,
                      f"PR-MPPI, 4 drones, B~U{{3..6}} branches, Dirichlet(1) belief, N(0,1) velocity shifts, "
                      f"K={n_samples} per branch, T={n_timesteps}, 12-gate lemniscate a=30 m, seed {seed}")


# ---------------------------------------------------------------------------
# StratRace2D (the reference's compiled 2-D race)

def sr_track_grids(track_fn, n=512, with_dp=False):
    """Track tables as the reference's StratRaceEnvironmentConfig.__post_init__ builds them
    (python-src/utils.py:382-428): polyline of track_fn(k / n), arc length, periodic cubic
    splines x(s), y(s) on a uniform arc-length grid; unit tangent; signed curvature
    (and the first derivatives dp when with_dp)."""
    from scipy.interpolate import CubicSpline
    q = np.array([track_fn(k / n) for k in range(n)], dtype=np.float64)
    d = np.linalg.norm(np.diff(q, axis=0, append=q[0:1]), axis=1)
    s_k = np.zeros(n + 1)
    s_k[1:] = np.cumsum(d)
    L = s_k[-1]
    qp = np.vstack([q, q[0:1]])
    cx = CubicSpline(s_k, qp[:, 0], bc_type="periodic")
    cy = CubicSpline(s_k, qp[:, 1], bc_type="periodic")
    s = np.linspace(0.0, L, n, endpoint=False)
    p = np.column_stack([cx(s), cy(s)])
    dx, dy = cx(s, 1), cy(s, 1)
    dp = np.column_stack([dx, dy])
    t = dp / np.maximum(np.linalg.norm(dp, axis=1, keepdims=True), 1e-6)
    ddx, ddy = cx(s, 2), cy(s, 2)
    kappa = (dx * ddy - dy * ddx) / np.maximum((dx ** 2 + dy ** 2) ** 1.5, 1e-6)
    if with_dp:
        return L, p, t, kappa, dp
    return L, p, t, kappa


def sr_opp_preset(drone_radius, s1="default", s2="default", s3="default", speed_scale=1.0, action_noise=0.1):
    """OppConfig.preset of python-src/utils.py:290-342 (one opponent)."""
    v1 = {"insensitive": 0.0, "default": 2.5 * drone_radius, "conservative": 4 * drone_radius}
    v2 = {"infinite": 0.0, "default": 1 / (4 * drone_radius) ** 2, "insensitive": 1 / (1e-5 * drone_radius) ** 2}
    v3 = {"no-block": 0.0, "default": 20 / drone_radius}
    return abi.SrOppConfig(s1=v1.get(s1, s1), s2=v2.get(s2, s2), s3=v3.get(s3, s3), speed_scale=speed_scale,
                           action_noise=action_noise)


def sr_env(radius=6.0, drone_radius=0.4, n_models=2, action_noise=0.1, pos_noise=0.0, speed_noise=0.0,
           env_action_noise=0.0):
    """The reference's stratRaceEnv (python-src/main.py:864-916): round track of radius 6 m,
    drone radius 0.4, max speed (3, 2), max accel (2, 3), track width 2, one lap, kV = 5,
    max opponent lateral factor 1.9, dt 0.1; opponent models = the first n_models presets."""
    L, p, t, kappa = sr_track_grids(lambda f: radius * np.array([math.cos(2 * math.pi * f), math.sin(2 * math.pi * f)]))
    e = abi.SrEnv()
    e.dt, e.drone_radius = 0.1, drone_radius
    e.pos_noise, e.speed_noise, e.action_noise = pos_noise, speed_noise, env_action_noise
    e.max_speed[0], e.max_speed[1] = 3.0, 2.0
    e.max_accel[0], e.max_accel[1] = 2.0, 3.0
    e.track_width, e.n_win_laps = 2.0, 1
    e.kp, e.kv, e.max_opp_lat_dist_fact = 10.0, 5.0, 1.9
    e.track_length = L
    presets = [dict(s1="insensitive", s3="default"), dict(s1="insensitive", s3="no-block"),
               dict(s1="default", s3="default"), dict(s1="conservative", s3="no-block")]
    e.n_models = n_models
    for m in range(n_models):
        e.opp[m] = sr_opp_preset(drone_radius, action_noise=action_noise, **presets[m])
    pf, tf, kf = p.astype(np.float32), t.astype(np.float32), kappa.astype(np.float32)
    for i in range(abi.SR_TRACK_SAMPLES):
        e.p_grid[i][0], e.p_grid[i][1] = float(pf[i, 0]), float(pf[i, 1])
        e.t_grid[i][0], e.t_grid[i][1] = float(tf[i, 0]), float(tf[i, 1])
        e.kappa_grid[i] = float(kf[i])
    return e


def sr_cost():
    """The dense MPPIConfig of stratRaceEnv (python-src/main.py:920-942)."""
    return abi.SrCost(coll_dist_factor=1.3, boundary_cost=50.0, boundary_threshold_factor=1.7, outside_cost=1e6,
                      win_cost=1e6, opp_win_cost=1e5, final_adv_weight=500.0, final_opp_adv_weight=0.0)


def sr_start_state(env, start=(0.05, 0.2)):
    """Ego and opponent at the grid samples round(512 * start), at rest (main.py:908-916)."""
    idx = [int(round(f * abi.SR_TRACK_SAMPLES)) for f in start]
    s = abi.SrState()
    for a, i in enumerate(idx):
        s.pos[2 * a], s.pos[2 * a + 1] = env.p_grid[i][0], env.p_grid[i][1]
        s.s[a] = float(np.float32(i / abi.SR_TRACK_SAMPLES * env.track_length))
    return s


class SrScenario:
    def __init__(self, name, env, cost, instances, states, steps, description):
        self.name, self.env, self.cost = name, env, cost
        self.instances, self.states, self.steps, self.description = instances, states, steps, description

    @property
    def sample_steps_per_step(self):
        return sum(i.n_samples * i.n_timesteps for i in self.instances)


def sr_instance(n_samples, n_timesteps, seed, theta=0, lam=10.0, sigma=1.5, adapt=0):
    return abi.SrInstance(n_samples=n_samples, n_timesteps=n_timesteps, true_theta=theta, adapt_lambda=adapt, seed=seed,
                          lambda_=lam, sigma=sigma)


def sr_race(n_inst=1, n_samples=65536, n_timesteps=70, seed=42, steps=300, **env_kw):
    """The reference's StratRace2D plain-MPPI race (K = 2^16, T = 70, temperature 10, sigma 1.5),
    n_inst independent worlds; instance i follows opponent model i % n_models."""
    env = sr_env(**env_kw)
    inst = [sr_instance(n_samples, n_timesteps, seed_mix(seed, 300 + i), theta=i % env.n_models) for i in range(n_inst)]
    st = [sr_start_state(env) for _ in range(n_inst)]
    return SrScenario(f"sr_race_{n_inst}x_k{n_samples}_t{n_timesteps}", env, sr_cost(), inst, st, steps,
                      f"StratRace2D, {n_inst} worlds, K={n_samples}, T={n_timesteps}, round track r=6 m")


def prsr_instance(n_samples, n_timesteps, seed, belief, lam=10.0, sigma=0.5, safety_weight=1e4, sampled=0, n_draws=0,
                  true_theta=0, adapt=0, fill=(0.0, 0.0)):
    ins = abi.PrsrInstance(n_samples=n_samples, n_timesteps=n_timesteps, sampled=sampled, n_draws=n_draws, seed=seed,
                           lambda_=lam, sigma=sigma, safety_weight=safety_weight, adapt_lambda=adapt, true_theta=true_theta)
    for b, p in enumerate(belief):
        ins.belief[b] = float(p)
    ins.shift_fill[0], ins.shift_fill[1] = fill
    return ins


def pr_sr_cost():
    """The PRMPPIConfig of stratRaceEnv (python-src/main.py:985-1004): collision factor 1.2,
    boundary cost 20 (factor 1.3), win 1e6, opponent win 1e5, terminal advance weight 500."""
    return abi.SrCost(coll_dist_factor=1.2, boundary_cost=20.0, boundary_threshold_factor=1.3, outside_cost=0.0,
                      win_cost=1e6, opp_win_cost=1e5, final_adv_weight=500.0, final_opp_adv_weight=0.0)


def pr_sr_race(n_inst=1, n_samples=65536, n_timesteps=60, seed=42, sampled=1, delta=0.1, steps=300, **env_kw):
    """The reference's StratRace2D PR-MPPI race: K = 2^16, T = 60, temperature 10, sigma 0.5,
    safety weight 1e4, P = ceil((1 - delta) / delta) models drawn from the belief (uniform over the
    environment's models); world i follows model i % n_models."""
    env = sr_env(**env_kw)
    P = int(math.ceil((1 - delta) / delta))
    belief = [1.0 / env.n_models] * env.n_models
    inst = [prsr_instance(n_samples, n_timesteps, seed_mix(seed, 400 + i), belief, sampled=sampled, n_draws=P if sampled else 0,
                          true_theta=i % env.n_models) for i in range(n_inst)]
    st = [sr_start_state(env) for _ in range(n_inst)]
    return SrScenario(f"pr_sr_race_{n_inst}x_k{n_samples}_t{n_timesteps}", env, pr_sr_cost(), inst, st, steps,
                      f"PR-MPPI on StratRace2D, {n_inst} worlds, K={n_samples}, T={n_timesteps}, P={P if sampled else env.n_models}")


def bm_knots(n_timesteps, n_knots):
    """Default knot times of MPPIConfig (python-src/utils.py:496-497): tau_m = m * floor((T - 2) / (M - 1))."""
    return [m * ((n_timesteps - 2) // (n_knots - 1)) for m in range(n_knots)]


def bm_instance(n_samples, n_timesteps, seed, n_knots=20, knots=None, lam=10.0, sigma=0.5, min_confidence=0.95, true_theta=0,
                adapt=1, belief=(0.5, 0.5), n_verif=0, verif_horizon=40, beta=1e-5, max_verif_eps=0.001):
    """One branching spline MPPI controller; n_verif > 0 adds the plan verification (the
    defaults are mainGate's: 2^17 samples, beta 1e-5, horizon 40, maxVerifEps 0.001,
    python-src/main.py:1022-1026)."""
    tau = bm_knots(n_timesteps, n_knots) if knots is None else list(knots)
    ins = abi.BmInstance(n_samples=n_samples, n_timesteps=n_timesteps, n_knots=len(tau), seed=seed, lambda_=lam, sigma=sigma,
                         min_confidence=min_confidence, true_theta=true_theta, adapt_lambda=adapt, n_verif_samples=n_verif,
                         verif_horizon=min(verif_horizon, n_timesteps) if n_verif else 0, beta=beta if n_verif else 0.0,
                         max_verif_eps=max_verif_eps if n_verif else 0.0)
    for m, t in enumerate(tau):
        ins.knots[m] = int(t)
    for b, p in enumerate(belief):
        ins.init_belief[b] = float(p)
    return ins


def bm_cost():
    """The spline MPPIConfig of stratRaceEnv (python-src/main.py:944-968): collision factor 1.4,
    boundary cost 50 (factor 1.8), outside 1e7, win 1e5, opponent win 1e5, terminal advance weight 50."""
    return abi.SrCost(coll_dist_factor=1.4, boundary_cost=50.0, boundary_threshold_factor=1.8, outside_cost=1e7,
                      win_cost=1e5, opp_win_cost=1e5, final_adv_weight=50.0, final_opp_adv_weight=0.0)


def bm_race(n_inst=1, n_samples=65536, n_timesteps=80, n_knots=20, seed=42, steps=300, n_models=2, n_verif=0, **env_kw):
    """The reference's default branching spline MPPI race on StratRace2D (stratRaceEnv with
    splines, main.py:944-968): K = 2^16, T = 80, 20 knots, temperature 10 (adapted), sigma 0.5,
    confidence 0.95, two opponent models with a uniform initial belief; world i follows model
    i % n_models; n_verif > 0 adds mainGate's plan verification (main.py:1022-1026)."""
    env = sr_env(n_models=n_models, **env_kw)
    belief = [1.0 / n_models] * n_models
    inst = [bm_instance(n_samples, n_timesteps, seed_mix(seed, 500 + i), n_knots=n_knots, true_theta=i % n_models, belief=belief,
                        n_verif=n_verif) for i in range(n_inst)]
    st = [sr_start_state(env) for _ in range(n_inst)]
    tag = f"_v{n_verif}" if n_verif else ""
    return SrScenario(f"bm_race_{n_inst}x_k{n_samples}_t{n_timesteps}_m{n_knots}{tag}", env, bm_cost(), inst, st, steps,
                      f"branching spline MPPI on StratRace2D, {n_inst} worlds, K={n_samples}, T={n_timesteps}, {n_knots} knots, "
                      f"{n_models} opponent models")


# ------------------------------------------------------------------ HiddenObs

def ho_env(pos_noise=0.0, speed_noise=0.0, action_noise=0.0, n_win_laps=1):
    """The HiddenObs course of hiddenObsEnv (python-src/main.py:705-781): a quarter-annulus
    corridor, two diagonal double half-planes, three rectangles, three gates and two hidden
    circular obstacles (model 0: the bottom one at (9, 5.5), model 1: the top one at (9, 8.5))."""
    min_r, max_r, diag_w, middle_w = 3.0, 5.0, 1.5, 4.0
    hid_w, hid_r, small_w, small_off = 4.0, 0.8, 4.0, 1.5
    corr_h, small_h, small_hoff, hid_h = 3.0, 1.0, 0.0, 1.5
    avg_r, neg_avg_r = (min_r + max_r) / 2, (max_r - min_r) / 2
    maxx, maxy = max_r + diag_w + middle_w + diag_w, corr_h + max_r + diag_w
    e = abi.HoEnv()
    e.arena_min[:] = [0.0, 0.0]
    e.arena_max[:] = [maxx + 1, maxy]
    e.dt, e.drone_radius, e.max_speed, e.max_accel, e.n_win_laps = 0.1, 0.3, 2.0, 3.0, n_win_laps
    e.pos_noise, e.speed_noise, e.action_noise = pos_noise, speed_noise, action_noise
    for g, (c, v) in enumerate(zip([(maxx, corr_h + avg_r), (neg_avg_r, corr_h), (max_r, corr_h + avg_r)], [(1, 0), (0, 1), (1, 0)])):
        e.gate_centers[g][:] = c
        e.gate_vectors[g][:] = v
        e.gate_radius[g] = (max_r - min_r) / 2
    rects = [(max_r - min_r, 0, maxx, corr_h), (max_r, corr_h, maxx, corr_h + min_r - diag_w),
             (max_r + small_off, corr_h + avg_r + small_hoff - small_h / 2, max_r + small_off + small_w, corr_h + avg_r + small_hoff + small_h / 2)]
    for r, box in enumerate(rects):
        e.rect_obstacles[r][:] = box
    e.hidden_centers[0][:] = [max_r + hid_w, avg_r + corr_h - hid_h]
    e.hidden_centers[1][:] = [max_r + hid_w, avg_r + corr_h + hid_h]
    e.hidden_radius[:] = [hid_r, hid_r]
    e.dbl_hp_limits[:] = [max_r, max_r + diag_w, max_r + diag_w + middle_w, maxx]
    e.dbl_hp_lambdas[:] = [corr_h, -corr_h - max_r - min_r, maxx + corr_h + max_r, maxx - corr_h - min_r]
    e.annulus_center[:] = [max_r, corr_h]
    e.annulus_radius[:] = [min_r, max_r]
    return e


def ho_start_state(env):
    """At rest at ((max_r - min_r) / 2, 1.5 drone radii), before gate 1 (main.py:783)."""
    s = abi.HoState()
    s.pos[:] = [1.0, env.drone_radius * 1.5]
    return s


def ho_cost():
    """The spline MPPIConfig of hiddenObsEnv (python-src/main.py:811-833): collision factor 1.2,
    boundary cost 10 (factor 1.4), outside 1e7, win 1e6, terminal advance weight 500."""
    return abi.SrCost(coll_dist_factor=1.2, boundary_cost=10.0, boundary_threshold_factor=1.4, outside_cost=1e7,
                      win_cost=1e6, opp_win_cost=0.0, final_adv_weight=500.0, final_opp_adv_weight=0.0)


def ho_race(n_inst=1, n_samples=65536, n_timesteps=80, n_knots=20, seed=2, steps=300, n_verif=0, **env_kw):
    """The reference's branching spline MPPI on HiddenObs (hiddenObsEnv with splines,
    main.py:811-833): K = 2^16, T = 80, 20 knots, temperature 10 (adapted), sigma 0.8,
    confidence 0.95, uniform initial belief over which hidden obstacle exists; world i has
    obstacle i % 2."""
    env = ho_env(**env_kw)
    inst = [bm_instance(n_samples, n_timesteps, seed_mix(seed, 700 + i), n_knots=n_knots, lam=10.0, sigma=0.8, true_theta=i % 2,
                        belief=(0.5, 0.5), n_verif=n_verif) for i in range(n_inst)]
    st = [ho_start_state(env) for _ in range(n_inst)]
    tag = f"_v{n_verif}" if n_verif else ""
    return SrScenario(f"ho_race_{n_inst}x_k{n_samples}_t{n_timesteps}_m{n_knots}{tag}", env, ho_cost(), inst, st, steps,
                      f"branching spline MPPI on HiddenObs, {n_inst} worlds, K={n_samples}, T={n_timesteps}, {n_knots} knots")


# ------------------------------------------------------------------ DroneRace

def drace_env(round_obstacle=True, action_noise=0.1, pid_noise=1.0, i_mppi=0, repulsion=0.0):
    """The DroneRace course of tinyGateEnv (python-src/main.py:40-161): a 25 m straight with
    two gates ((25, 0) and (0.25, 0), normal +x, radius 1), one round obstacle of radius 0.25
    at its middle (or a 5 m x 1 m box), and two racelines passing above and below it
    (y = +-1.5 sin^2 ramps); the opponent models are PD controllers (kp 5, kd 20, action
    noise 1) on raceline 0 ("Top") and 1 ("Bottom")."""
    length = 25.0 if round_obstacle else 30.0
    height, obs_size, race_delay, margin, height_factor = 0.5, 0.2, 0.1, 0.05, 3.0
    e = abi.DraceEnv()
    e.arena_min[:] = [-0.1 * length, -2 * height * height_factor]
    e.arena_max[:] = [1.1 * length, 2 * height * height_factor]
    e.dt, e.drone_radius, e.action_noise = 0.1, 0.75, action_noise
    e.max_speed[:] = [2.05, 2.5]
    e.max_accel[:] = [3.0, 3.0]
    e.n_win_laps, e.target_distance, e.n_gates = 1, 0.05, 2
    e.gate_centers[0][:] = [length, 0.0]
    e.gate_centers[1][:] = [length * 0.01, 0.0]
    for g in range(2):
        e.gate_vectors[g][:] = [1.0, 0.0]
        e.gate_radius[g] = 1.0
    if round_obstacle:
        e.n_round_obstacles = 1
        e.round_centers[0][:] = [length / 2, 0.0]
        e.round_radius[0] = height * 0.5
    else:
        e.n_obstacles = 1
        e.obstacles[0][:] = [length * (0.5 - obs_size / 2), -height, length * (0.5 + obs_size / 2), height]
    e.n_racelines, e.n_models, e.i_mppi = 2, 2, i_mppi
    for m in range(2):
        e.opp_pid[m] = abi.DracePid(kp=5.0, kd=20.0, repulsion_factor=repulsion, repulsion_power=2.0, repulsion_dist_fact=5.0,
                                    raceline_index=m, action_noise=pid_noise)

    def lift(x):
        xn = x / length
        if xn > 1.0:
            return 0.0
        if xn > 0.5:
            xn = 1 - xn
        if xn < 0.5 - obs_size / 2 - race_delay:
            return 0.0
        if xn < 0.5 - obs_size / 2.0 - margin:
            return np.sin(np.pi / 2.0 * (xn - (0.5 - obs_size / 2 - race_delay)) / (race_delay - margin)) ** 2
        return 1.0

    xs = np.linspace(0.0, length * 1.1, abi.DRACE_SAMPLES)
    ys = np.array([lift(x) for x in xs]) * height * height_factor
    for i in range(abi.DRACE_SAMPLES):
        e.track_points[0][i][:] = [xs[i], ys[i]]
        e.track_points[1][i][:] = [xs[i], -ys[i]]
    return e


def drace_start_state(env, start=(0.1, 0.02)):
    """Both agents on raceline 0 at samples int(512 s), at rest, progress s on both racelines,
    next gate 0 (main.py:132-138)."""
    st = abi.DraceState()
    for a, s0 in enumerate(start):
        i = int(s0 * abi.DRACE_SAMPLES)
        st.pos[2 * a], st.pos[2 * a + 1] = env.track_points[0][i][0], env.track_points[0][i][1]
        for r in range(abi.DRACE_MAX_RACELINES):
            st.s[a][r] = s0 if r < env.n_racelines else 0.0
        st.gates[a] = 1
    return st


def drace_cost():
    """The spline MPPIConfig of tinyGateEnv (main.py:140-161): collision factor 1.1, boundary
    cost 0 (factor 1.7), outside 1e6, win 1e6, opponent win 1000 (MPPIConfig's default),
    terminal advance weight 200."""
    return abi.SrCost(coll_dist_factor=1.1, boundary_cost=0.0, boundary_threshold_factor=1.7, outside_cost=1e6, win_cost=1e6,
                      opp_win_cost=1000.0, final_adv_weight=200.0, final_opp_adv_weight=0.0)


def drace_race(n_inst=1, n_samples=65536, n_timesteps=60, n_knots=10, seed=42, steps=300, n_verif=0, **env_kw):
    """The reference's branching spline MPPI on DroneRace (tinyGateEnv with splines): K = 2^16,
    T = 60, 10 knots, temperature 10 (adapted), sigma 2, confidence 0.95, uniform belief over the
    opponent's two PD models; world i's opponent follows model i % 2."""
    env = drace_env(**env_kw)
    inst = [bm_instance(n_samples, n_timesteps, seed_mix(seed, 900 + i), n_knots=n_knots, lam=10.0, sigma=2.0, true_theta=i % 2,
                        belief=(0.5, 0.5), n_verif=n_verif) for i in range(n_inst)]
    st = [drace_start_state(env) for _ in range(n_inst)]
    tag = f"_v{n_verif}" if n_verif else ""
    return SrScenario(f"drace_race_{n_inst}x_k{n_samples}_t{n_timesteps}_m{n_knots}{tag}", env, drace_cost(), inst, st, steps,
                      f"branching spline MPPI on DroneRace, {n_inst} worlds, K={n_samples}, T={n_timesteps}, {n_knots} knots")


def seed_mix(seed, i):
    """64-bit Philox key for instance i of a scenario seeded `seed`."""
    return ((int(seed) * 0x9E3779B97F4A7C15) ^ (0xA5A5A5A5 + 0x632BE59BD9B4E019 * (i + 1))) & 0xFFFFFFFFFFFFFFFF
