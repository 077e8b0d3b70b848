"""GPU test of the simulation session behind the wire protocol (SimulationEngine::run,
engine.cpp:314-352): every frame the session sends is rebuilt independently from an engine
handle stepped in parallel and from the oracle's display predictions (engine.cpp:96-257),
and must match byte for byte.
"""
import copy
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def state_part(P, step, st, prev, action, env, true_theta, oracle):
    P.i32(1)
    P.i32(step)
    P.arr(st.pos[:])
    P.arr(st.vel[:])
    P.arr(st.s[:])
    P.arr(st.lat_dist[:])
    P.arr([float(v) for v in st.laps])
    P.arr([oracle.sr_opp_target(0, env, true_theta, prev)])
    P.arr(action)


def tail_frames(wire, step, ev, max_steps):
    out = []
    if ev:
        P = wire.Packer()
        P.i32(2)
        P.i32(ev - 1)
        out.append(P.bytes())
    elif step == max_steps:
        P = wire.Packer()
        P.i32(2)
        P.i32(3)
        out.append(P.bytes())
    if out:
        P = wire.Packer()
        P.i32(3)
        out.append(P.bytes())
    return out


def test_mppi_session_frames(pkg, oracle):
    wire, abi = pkg.wire, pkg.abi
    say("[session] branching spline MPPI with plan verification (2^17 samples), K=1024 T=30 M=8, 6 steps")
    e = wire.stratrace_env(n_models=2, seed=77, true_theta=1)
    env_f, cont_f, st_f = wire.env_header(e), wire.mppi_header(n_samples=1024, n_timesteps=30, n_knots=8), wire.init_state_header(e)
    sess = wire.Session(env_f, cont_f, st_f, max_steps=6)
    frames = list(sess)
    sess.close()

    env, info = wire.decode_env(env_f)
    cont = wire.decode_controller(cont_f)
    x0 = wire.decode_state(st_f)
    ins = pkg.scenarios.bm_instance(1024, 30, 77, knots=list(cont.knots[:8]), lam=cont.lambda_, sigma=cont.sigma,
                                    min_confidence=cont.min_confidence, true_theta=1, adapt=1, belief=list(info.init_belief[:2]),
                                    n_verif=cont.n_verif_samples, verif_horizon=min(cont.verif_horizon, 30), beta=cont.beta,
                                    max_verif_eps=cont.max_verif_eps)
    eng = pkg.BMEngine(pkg.scenarios.SrScenario("session", env, cont.cost, [ins], [x0], 6, ""))
    T, NT = 30, 2

    def mppi_frame(step, st, prev, action, belief, nominal):
        P = wire.Packer()
        state_part(P, step, st, prev, action, env, 1, oracle)
        P.arr(belief)
        d = eng.diag()[0]
        P.i32(d.fail_count)
        P.f32(d.epsilon)
        P.f32(d.epsilon_partial)
        P.i32(d.use_new_plan)
        P.f32(d.certified_loss)
        P.i32(NT)
        for theta in range(NT):
            pos, act, inf = oracle.sr_predict(env, st, theta, nominal, T, True, belief, cont.min_confidence)
            P.arr([inf[0]])
            P.arr([inf[1]])
            P.arr([inf[2]])
            P.arr(pos)
            P.arr(act)
            P.i32(inf[3])
            P.i32(inf[4])
        return P.bytes()

    want = [mppi_frame(0, x0, x0, [0.0, 0.0], eng.belief(0), np.zeros((NT + 1, T, 2), np.float32))]
    st = copy.deepcopy(x0)
    for step in range(1, 7):
        prev = st
        act = eng.step(1, advance=True)[0]
        st = copy.deepcopy(eng.get_states()[0])
        want.append(mppi_frame(step, st, prev, act, eng.belief(0), eng.nominal(0)))
        ev = int(eng.events()[0])
        want += tail_frames(wire, step, ev, 6)
        if ev:
            break
    eng.close()
    assert [t for t, _ in frames] == [wire.Unpacker(w).i32() for w in want]
    for n, ((t, got), w) in enumerate(zip(frames, want)):
        assert got == w, f"frame {n} (type {t}) differs"
    d = wire.decode_frame(frames[-3][1], abi.CONT_MPPI)
    say(f"  {len(frames)} frames identical; step {d['step']} belief {d['belief']} predicted branches "
        f"{[int(p['predTheta'][0]) for p in d['preds']]}; verification: failures {d['failCount']} useNewPlan {d['useNewPlan']} "
        f"certified loss {d['certifiedLoss']:.3e} epsilon {d['epsilon']:.3e}")


def test_prmppi_session_frames(pkg, oracle):
    wire, abi = pkg.wire, pkg.abi
    say("[session] PR-MPPI, K=512 T=20 P=3, 5 steps")
    e = wire.stratrace_env(n_models=2, seed=91, true_theta=0)
    env_f, cont_f, st_f = wire.env_header(e), wire.prmppi_header(n_samples=512, n_timesteps=20, P=3), wire.init_state_header(e)
    sess = wire.Session(env_f, cont_f, st_f, max_steps=5)
    frames = list(sess)
    sess.close()

    env, info = wire.decode_env(env_f)
    cont = wire.decode_controller(cont_f)
    x0 = wire.decode_state(st_f)
    ins = pkg.scenarios.prsr_instance(512, 20, 91, list(info.init_belief[:2]), lam=cont.lambda_, sigma=cont.sigma,
                                      safety_weight=cont.safety_weight, sampled=1, n_draws=3, true_theta=0, adapt=1)
    eng = pkg.PRSREngine(pkg.scenarios.SrScenario("session", env, cont.cost, [ins], [x0], 5, ""))
    T, NT = 20, 2

    def pr_frame(step, st, prev, action):
        P = wire.Packer()
        state_part(P, step, st, prev, action, env, 0, oracle)
        P.arr(prev.pos[:])
        P.arr(eng.belief(0, NT))
        d = eng.diag()[0]
        P.i32(d.use_nom_plan)
        P.i32(d.reset_nom)
        P.i32(NT + 1)
        nom, rob = eng.get_nominals(0)
        for j in range(NT + 1):
            pos, act, inf = oracle.sr_predict(env, prev, min(j, NT - 1), nom if j < NT else rob, T, False)
            P.arr(pos)
            P.arr(act)
            P.i32(inf[3])
            P.i32(inf[4])
        return P.bytes()

    want = [pr_frame(0, x0, x0, [0.0, 0.0])]
    st = copy.deepcopy(x0)
    for step in range(1, 6):
        prev = st
        act = eng.step(1, advance=True)[0]
        st = copy.deepcopy(eng.get_states()[0])
        want.append(pr_frame(step, st, prev, act))
        ev = int(eng.events()[0])
        want += tail_frames(wire, step, ev, 5)
        if ev:
            break
    eng.close()
    assert len(frames) == len(want)
    for n, ((t, got), w) in enumerate(zip(frames, want)):
        assert got == w, f"frame {n} (type {t}) differs"
    d = wire.decode_frame(frames[-3][1], abi.CONT_PRMPPI)
    say(f"  {len(frames)} frames identical; step {d['step']} useNomPlan {d['useNomPlan']} resetNom {d['resetNom']}")


def test_serve_over_socket_and_silent_states(pkg):
    """serve() moves the frames over a socket object; sendStates = 0 leaves EVENT and DONE only."""
    wire = pkg.wire
    say("[session] serve over a socket object; sendStates off")
    e = wire.stratrace_env(n_models=2, seed=5)

    class Sock:
        def __init__(self, frames):
            self.inbox, self.sent = list(frames), []

        def recv(self):
            return self.inbox.pop(0)

        def send(self, b):
            self.sent.append(bytes(b))

    frames = (wire.env_header(e), wire.mppi_header(n_samples=256, n_timesteps=20, n_knots=5), wire.init_state_header(e))
    s = Sock(frames)
    assert wire.serve(s, max_steps=3) == pkg.abi.EVT_TRUNCATED
    sess = wire.Session(*frames, max_steps=3)
    assert s.sent == [f for _, f in sess]
    sess.close()
    assert [wire.Unpacker(f).i32() for f in s.sent] == [1, 1, 1, 1, 2, 3]
    quiet = Sock((wire.env_header(dict(e, send_states=False)),) + frames[1:])
    wire.serve(quiet, max_steps=3)
    assert [wire.Unpacker(f).i32() for f in quiet.sent] == [2, 3]
