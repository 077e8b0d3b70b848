"""Wire protocol and simulation session of the reference's ZMQ front-end (StratRace2D build).

Frames follow python-src/protocol.py:13-116 (BytePacker / ByteUnpacker): little-endian int32
and float32; an array is an int32 length followed by float32 values. The three header frames
the front-end sends (protocol.py:225-229) are the environment config, the controller config
and the initial state, each packed field by field in dataclass order (utils.py:52-66,
345-380, 443-453, 456-492, 504-532); the engine answers with STATE frames (FullStateInfo,
utils.py:544-602), an EVENT and DONE (engine.cpp:266-352).

`Session` runs SimulationEngine on this library's engines (dr_session_*); `serve` is the
socket loop of main.cpp:9-44 over any object with recv() -> bytes and send(bytes), e.g. a
pyzmq PAIR socket. The header builders below reproduce what the front-end's `stratRaceEnv`
(python-src/main.py:864-1000) sends, for tests and for clients without the front-end.
"""
import ctypes as C
import math
import struct

import numpy as np

from . import abi
from .abi import check


class Packer:
    """BytePacker's encoding (protocol.py:13-30)."""

    def __init__(self):
        self.buf = bytearray()

    def i32(self, x):
        self.buf += struct.pack("<i", int(x))

    def f32(self, x):
        self.buf += struct.pack("<f", float(np.float32(x)))

    def arr(self, a):
        a = np.asarray(a, dtype=np.float32).ravel()
        self.i32(a.size)
        self.buf += a.tobytes()

    def bytes(self):
        return bytes(self.buf)


class Unpacker:
    """ByteUnpacker's decoding (protocol.py:72-116)."""

    def __init__(self, data):
        self.buf = memoryview(bytes(data))
        self.off = 0

    def _take(self, n):
        if self.off + n > len(self.buf):
            raise ValueError(f"buffer underflow: need {n} bytes, have {len(self.buf) - self.off}")
        chunk = self.buf[self.off:self.off + n]
        self.off += n
        return chunk

    def i32(self):
        return struct.unpack("<i", self._take(4))[0]

    def f32(self):
        return struct.unpack("<f", self._take(4))[0]

    def arr(self):
        n = self.i32()
        if n < 0:
            raise ValueError(f"negative array length {n}")
        return np.frombuffer(self._take(4 * n), dtype=np.float32, count=n).copy()

    def finished(self):
        if self.off != len(self.buf):
            raise ValueError(f"{len(self.buf) - self.off} trailing bytes left")


# --------------------------------------------------------------------------- headers

OPP_PRESETS = [dict(s1="insensitive", s3="default"), dict(s1="insensitive", s3="no-block"),
               dict(s1="default", s3="default"), dict(s1="conservative", s3="no-block")]


def stratrace_env(n_models=2, first_model=0, seed=42, true_theta=0, init_belief=None, send_states=True, drone_radius=0.4,
                  action_noise=0.1, pos_noise=0.0, speed_noise=0.0, env_action_noise=0.0):
    """The fields of `stratRaceEnv`'s StratRaceEnvironmentConfig (main.py:864-906; defaults of
    utils.py:345-380) as a dict, track tables in float64 as __post_init__ computes them."""
    from .scenarios import sr_opp_preset, sr_track_grids
    L, p, t, kappa, dp = sr_track_grids(lambda f: 6 * np.array([math.cos(2 * math.pi * f), math.sin(2 * math.pi * f)]), with_dp=True)
    opps = [sr_opp_preset(drone_radius, action_noise=action_noise, **OPP_PRESETS[first_model + m]) for m in range(n_models)]
    width = 2.0
    return dict(
        n_models=n_models, arena_min=p.min(axis=0) - 2 * width, arena_max=p.max(axis=0) + 2 * width, n_opp=1, dim=2, dt=0.1,
        send_states=send_states, drone_radius=drone_radius, pos_noise=pos_noise, speed_noise=speed_noise,
        action_noise=env_action_noise, max_speed=[3.0, 2.0], max_accel=[2.0, 3.0], n_track=512, track_width=width, n_win_laps=1,
        seed=seed, init_belief=np.full(n_models, 1.0 / n_models) if init_belief is None else np.asarray(init_belief, float),
        opps=[(o.s1, o.s2, o.s3, o.speed_scale, o.action_noise) for o in opps], true_theta=true_theta, kp=10.0, kv=5.0,
        lat_fact=1.9, track_length=L, p_grid=p, dp_grid=dp, t_grid=t, kappa_grid=kappa)


def env_header(e):
    """encodeConfig(envConfig, MSG_HEADER) for a StratRace2D config (field order utils.py:52-66, 345-380;
    actionDim, nTrueModels and trackFunction are not sent)."""
    w = Packer()
    w.i32(abi.MSG_HEADER)
    w.i32(abi.ENV_STRATRACE)
    w.i32(1)                       # nModelFactors
    w.arr([e["n_models"]])         # modelSizes
    w.arr(e["arena_min"])
    w.arr(e["arena_max"])
    w.i32(e["n_opp"])
    w.i32(e["dim"])
    w.f32(e["dt"])
    w.i32(1 if e["send_states"] else 0)
    w.f32(e["drone_radius"])
    w.f32(e["pos_noise"])
    w.f32(e["speed_noise"])
    w.f32(e["action_noise"])
    w.arr(e["max_speed"])
    w.arr(e["max_accel"])
    w.i32(e["n_track"])
    w.f32(e["track_width"])
    w.i32(e["n_win_laps"])
    w.i32(e["seed"])
    w.arr(e["init_belief"])
    for s1, s2, s3, speed_scale, noise in e["opps"]:  # OppConfig: s1_i, s2_i, s3_i, speedScale, actionNoise
        for v in (s1, s2, s3, speed_scale, noise):
            w.arr([v])
    w.i32(e["true_theta"])
    w.f32(e["kp"])
    w.f32(e["kv"])
    w.f32(e["lat_fact"])
    w.f32(e["track_length"])
    w.arr(e["p_grid"])
    w.arr(e["dp_grid"])
    w.arr(e["t_grid"])
    w.arr(e["kappa_grid"])
    return w.bytes()


def mppi_header(n_samples=2 ** 16, n_timesteps=80, n_knots=20, knots=None, lam=10.0, sigma=0.5, margin=0.95, coll=1.4, boundary=50.0,
                factor=1.8, outside=1e7, win=1e5, opp_win=1e5, final_adv=50.0, final_opp_adv=0.0, min_conf=0.95, n_verif=2 ** 17,
                beta=1e-5, verif_horizon=40, max_verif_eps=0.001, use_splines=True):
    """encodeConfig(MPPIConfig, MSG_HEADER) (utils.py:456-492); defaults = stratRaceEnv's spline
    setting (main.py:944-968) with mainGate's verification settings (main.py:1022-1026)."""
    tau = [m * ((n_timesteps - 2) // (n_knots - 1)) for m in range(n_knots)] if knots is None else list(knots)
    w = Packer()
    w.i32(abi.MSG_HEADER)
    w.i32(abi.CONT_MPPI)
    w.i32(n_samples)
    w.i32(n_timesteps)
    w.i32(len(tau))
    w.arr(tau)
    w.i32(1 if use_splines else 0)
    for v in (lam, sigma, margin, coll, boundary, factor, outside, win, opp_win, final_adv, final_opp_adv, min_conf):
        w.f32(v)
    w.i32(n_verif)
    w.f32(beta)
    w.i32(verif_horizon)
    w.f32(max_verif_eps)
    return w.bytes()


def prmppi_header(n_samples=2 ** 16, n_timesteps=60, lam=10.0, sigma=0.5, margin=0.9, coll=1.2, boundary=20.0, factor=1.3, win=1e6,
                  opp_win=1e5, final_adv=500.0, final_opp_adv=0.0, safety=1e4, min_safe=0.0, delta=0.1, P=0):
    """encodeConfig(PRMPPIConfig, MSG_HEADER) (utils.py:504-532); defaults = stratRaceEnv's PR
    setting (main.py:970-993); P = ceil((1 - delta) / delta) when 0 (utils.py:537-538)."""
    if P == 0:
        P = math.ceil((1 - delta) / delta)
    w = Packer()
    w.i32(abi.MSG_HEADER)
    w.i32(abi.CONT_PRMPPI)
    w.i32(n_samples)
    w.i32(n_timesteps)
    for v in (lam, sigma, margin, coll, boundary, factor, win, opp_win, final_adv, final_opp_adv, safety, min_safe, delta):
        w.f32(v)
    w.i32(P)
    return w.bytes()


def init_state_header(e, start=(0.05, 0.2)):
    """encodeConfig(initState, MSG_HEADER) of stratRaceEnv's initial state (main.py:908-916):
    agents at the grid samples round(512 * start), at rest; latDistTarget is not sent."""
    idx = np.round(np.asarray(start) * e["n_track"]).astype(int)
    w = Packer()
    w.i32(abi.MSG_HEADER)
    w.arr(e["p_grid"][idx])
    w.arr(np.zeros((2, 2)))
    w.arr(idx.astype(float) / e["n_track"] * e["track_length"])
    w.arr(np.zeros(2))
    w.arr(np.zeros(2))
    return w.bytes()


# --------------------------------------------------------------------------- frames from the engine

def decode_frame(frame, kind):
    """One frame of the engine, as the front-end reads it (protocol.py:233-278): a dict with
    'type' and, for STATE, the FullStateInfo fields (utils.py:544-602)."""
    u = Unpacker(frame)
    t = u.i32()
    out = {"type": t}
    if t == abi.MSG_STATE:
        out["step"] = u.i32()
        out["state"] = {k: u.arr() for k in ("pos", "vel", "S", "latDist", "laps", "latDistTarget")}
        out["egoAction"] = u.arr()
        if kind == abi.CONT_MPPI:
            out["belief"] = u.arr()
            out["failCount"] = u.i32()
            out["epsilon"] = u.f32()
            out["epsilonPartial"] = u.f32()
            out["useNewPlan"] = u.i32()
            out["certifiedLoss"] = u.f32()
            n = u.i32()
            out["preds"] = [dict(initPredTheta=u.arr(), branchTime=u.arr(), predTheta=u.arr(), fullPos=u.arr(), egoActions=u.arr(),
                                 stopReason=u.i32(), stopTime=u.i32()) for _ in range(n)]
        else:
            out["prevPos"] = u.arr()
            out["belief"] = u.arr()
            out["useNomPlan"] = u.i32()
            out["resetNom"] = u.i32()
            n = u.i32()
            out["preds"] = [dict(fullPos=u.arr(), egoActions=u.arr(), stopReason=u.i32(), stopTime=u.i32()) for _ in range(n)]
    elif t == abi.MSG_EVENT:
        out["event"] = u.i32()
    u.finished()
    return out


def _buf(b):
    return C.cast(C.c_char_p(b), C.c_void_p), len(b)


def decode_env(frame):
    """dr_wire_decode_env: (SrEnv, WireEnvInfo)."""
    lib = abi.load()
    env, info = abi.SrEnv(), abi.WireEnvInfo()
    p, n = _buf(frame)
    check(lib.dr_wire_decode_env(p, n, C.byref(env), C.byref(info)))
    return env, info


def decode_controller(frame):
    lib = abi.load()
    c = abi.WireController()
    p, n = _buf(frame)
    check(lib.dr_wire_decode_controller(p, n, C.byref(c)))
    return c


def decode_state(frame):
    lib = abi.load()
    s = abi.SrState()
    p, n = _buf(frame)
    check(lib.dr_wire_decode_state(p, n, C.byref(s)))
    return s


class Session:
    """SimulationEngine on this library's engines (dr_session_*): iterate to get the frames
    SimulationEngine::run sends, as (msg_type, bytes)."""

    def __init__(self, env_frame, cont_frame, state_frame, max_steps=0):
        self.lib = abi.load()
        self._frames = (bytes(env_frame), bytes(cont_frame), bytes(state_frame))
        h = C.c_void_p()
        args = []
        for f in self._frames:
            args += list(_buf(f))
        check(self.lib.dr_session_create(*args, int(max_steps), C.byref(h)))
        self.h = h
        self.kind = decode_controller(cont_frame).kind

    def next_frame(self):
        p = C.POINTER(C.c_uint8)()
        n = C.c_size_t()
        t = C.c_int32()
        check(self.lib.dr_session_next(self.h, C.byref(p), C.byref(n), C.byref(t)))
        if t.value < 0:
            return None
        return t.value, C.string_at(p, n.value)

    def __iter__(self):
        while True:
            f = self.next_frame()
            if f is None:
                return
            yield f

    def close(self):
        if self.h:
            self.lib.dr_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve(sock, max_steps=0):
    """runEngine (main.cpp:9-44) over `sock` (recv() -> bytes, send(bytes)): the three header
    frames in, every frame of the session out. Returns the last event type."""
    env_frame = sock.recv()
    cont_frame = sock.recv()
    state_frame = sock.recv()
    sess = Session(env_frame, cont_frame, state_frame, max_steps)
    event = None
    try:
        for t, frame in sess:
            if t == abi.MSG_EVENT:
                event = struct.unpack("<i", frame[4:8])[0]
            sock.send(frame)
    finally:
        sess.close()
    return event
