"""ORACLE — test infrastructure only.

Algorithmic FP32 work per rollout sample-step of the benchmark workload
(4 drones, K=16384, T=100, first control step): dynamics, gate logic and
costs counted by the oracle's op-counting scalar on the first `n` samples
of each instance, plus the per-step noise cost read off the code
(Box-Muller: log 20, sincos 21, 5 more per pair, two pairs; 4 sigma
scalings: 96 flops).  Prints the figure bench.py uses."""
import ctypes as C
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
NOISE_FLOPS_PER_STEP = 2 * (20 + 21 + 5) + 4


def package():
    name = "drone_race_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = os.path.join(ROOT, "drone-race_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"), submodule_search_locations=[d])
    m = importlib.util.module_from_spec(spec)
    sys.modules[name] = m
    spec.loader.exec_module(m)
    return m


def count(scn, n=1024):
    sys.path.insert(0, HERE)
    import pyoracle
    pkg = package()
    abi = pkg.abi
    L = pyoracle.lib(abi)
    L.or_count_rollout_ops.argtypes = [C.POINTER(abi.QuadParams), C.POINTER(abi.CostCfg), C.POINTER(abi.Track),
                                       C.POINTER(abi.DroneState), C.POINTER(abi.Opponent), C.c_int,
                                       C.POINTER(C.c_float), C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.POINTER(C.c_float), C.POINTER(C.c_longlong)]
    flops = 0
    steps = 0
    for ins in scn.instances:
        opp = []
        for d in range(len(scn.drone_world)):
            if d != ins.drone and scn.drone_world[d] == scn.drone_world[ins.drone]:
                o = abi.Opponent()
                for c in range(3):
                    o.p[c], o.v[c] = scn.drones[d].p[c], scn.drones[d].v[c]
                opp.append(o)
        oa = (abi.Opponent * max(1, len(opp)))(*opp)
        nom = np.tile(np.array(ins.shift_fill[:], np.float32), (ins.n_timesteps, 1)).reshape(-1)
        sig = np.array(ins.sigma[:], np.float32)
        out = C.c_longlong()
        rc = L.or_count_rollout_ops(C.byref(scn.params), C.byref(scn.cost), C.byref(scn.tracks[ins.track]),
                                    C.byref(scn.drones[ins.drone]), oa, len(opp),
                                    nom.ctypes.data_as(C.POINTER(C.c_float)), n, ins.n_timesteps,
                                    ins.seed & 0xFFFFFFFF, ins.seed >> 32, 0,
                                    sig.ctypes.data_as(C.POINTER(C.c_float)), C.byref(out))
        assert rc == 0
        flops += out.value
        steps += n * ins.n_timesteps
    return flops / steps + NOISE_FLOPS_PER_STEP


def count_refine(scn, n_grad=3, ck=10):
    """Algorithmic FP32 flops of one adjoint-refinement launch over every controller of the
    workload, counted by the oracle's op-counting scalar on each controller's hover plan from its
    start state: n_grad iterates of J and dJ/dU (forward with checkpoints, segment recompute,
    reverse) plus the J of the last iterate, plus the projected step (fma and two-sided clamp,
    4 flops per control value) of each iterate."""
    sys.path.insert(0, HERE)
    import pyoracle
    pkg = package()
    abi = pkg.abi
    L = pyoracle.lib(abi)
    L.or_count_cost_grad_ops.argtypes = [C.POINTER(abi.QuadParams), C.POINTER(abi.CostCfg), C.POINTER(abi.Track),
                                         C.POINTER(abi.DroneState), C.POINTER(abi.Opponent), C.c_int, C.POINTER(C.c_float),
                                         C.c_int, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    total = 0
    for ins in scn.instances:
        opp = []
        for d in range(len(scn.drone_world)):
            if d != ins.drone and scn.drone_world[d] == scn.drone_world[ins.drone]:
                o = abi.Opponent()
                for c in range(3):
                    o.p[c], o.v[c] = scn.drones[d].p[c], scn.drones[d].v[c]
                opp.append(o)
        oa = (abi.Opponent * max(1, len(opp)))(*opp)
        T = ins.n_timesteps
        U = np.tile(np.array(ins.shift_fill[:], np.float32), (T, 1)).reshape(-1)
        fwd, full = C.c_longlong(), C.c_longlong()
        rc = L.or_count_cost_grad_ops(C.byref(scn.params), C.byref(scn.cost), C.byref(scn.tracks[ins.track]),
                                      C.byref(scn.drones[ins.drone]), oa, len(opp), U.ctypes.data_as(C.POINTER(C.c_float)), T, ck,
                                      C.byref(fwd), C.byref(full))
        assert rc == 0
        total += n_grad * (full.value + 4 * T * 4) + fwd.value
    return total


# Noise cost per StratRace2D rollout-step read off the code: one Box-Muller
# pair (log 20, sincos 21, 5 more) and two sigma scalings for the control
# noise of a (sample, step) = 48; the environment's opponent-action noise is
# drawn four normals at a time (two pairs) = 92 per step.
SR_CONTROL_NOISE = (20 + 21 + 5) + 2
SR_ENV_NOISE = 2 * (20 + 21 + 5)


def count_sr(pkg, n=256):
    """FP32 flops per rollout-step of the three StratRace2D workloads at their first control step
    (oracle op counter over n samples, early stops included), plus the noise per rollout-step:
    plain (control noise per step), PR-MPPI (control noise shared by the 2*P rollouts of a
    sample), branching (the NB*M knot-noise pairs of a sample shared by its model rollouts)."""
    sys.path.insert(0, HERE)
    import pyoracle
    abi = pkg.abi
    L = pyoracle.lib(abi)
    L.or_count_sr_ops.restype = C.c_longlong
    L.or_count_sr_ops.argtypes = [C.c_int, C.POINTER(abi.SrEnv), C.POINTER(abi.SrCost), C.POINTER(abi.SrState), C.c_int,
                                  C.POINTER(C.c_float), C.c_int, C.c_int, C.c_uint64, C.POINTER(abi.BmInstance),
                                  C.POINTER(C.c_float), C.POINTER(C.c_longlong)]
    fp = C.POINTER(C.c_float)
    sc = pkg.scenarios
    out = {}
    steps = C.c_longlong()
    s1 = sc.sr_race(n_inst=1)
    i1 = s1.instances[0]
    nom = np.zeros(i1.n_timesteps * 2, np.float32)
    f = L.or_count_sr_ops(0, C.byref(s1.env), C.byref(s1.cost), C.byref(s1.states[0]), i1.true_theta, nom.ctypes.data_as(fp), n,
                          i1.n_timesteps, i1.seed, None, None, C.byref(steps))
    out["sr_race"] = f / steps.value + SR_CONTROL_NOISE + SR_ENV_NOISE
    s2 = sc.pr_sr_race(n_inst=1)
    i2 = s2.instances[0]
    nom = np.zeros(i2.n_timesteps * 2, np.float32)
    f = L.or_count_sr_ops(1, C.byref(s2.env), C.byref(s2.cost), C.byref(s2.states[0]), 0, nom.ctypes.data_as(fp), n,
                          i2.n_timesteps, i2.seed, None, None, C.byref(steps))
    out["pr_sr_race"] = f / steps.value + SR_CONTROL_NOISE / (2 * i2.n_draws) + SR_ENV_NOISE
    s3 = sc.bm_race(n_inst=1)
    i3 = s3.instances[0]
    NT = s3.env.n_models
    kn = np.zeros((NT + 1) * i3.n_knots * 2, np.float32)
    b = np.zeros(abi.SR_MAX_MODELS, np.float32)
    b[:NT] = i3.init_belief[:NT]
    f = L.or_count_sr_ops(2, C.byref(s3.env), C.byref(s3.cost), C.byref(s3.states[0]), 0, kn.ctypes.data_as(fp), n,
                          i3.n_timesteps, i3.seed, C.byref(i3), b.ctypes.data_as(fp), C.byref(steps))
    knot_noise = (NT + 1) * i3.n_knots * SR_CONTROL_NOISE * n / steps.value
    out["bm_race"] = f / steps.value + knot_noise + SR_ENV_NOISE
    out["bm_race_steps_fraction"] = steps.value / (n * NT * i3.n_timesteps)
    return out


def count_ho(pkg, n=256):
    """FP32 flops per rollout-step of the branching HiddenObs workload at its first control step
    (oracle op counter over n samples, early stops included) plus the knot noise of a sample shared
    by its model rollouts; the course has no environment noise."""
    sys.path.insert(0, HERE)
    import pyoracle
    oracle = pyoracle.Oracle(pkg.abi)
    s = pkg.scenarios.ho_race(n_inst=1)
    ins = s.instances[0]
    kn = np.zeros(3 * ins.n_knots * 2, np.float32)
    f, steps = oracle.count_ho(s.env, s.cost, s.states[0], ins, kn, np.float32(ins.init_belief[:2]), n)
    knot_noise = 3 * ins.n_knots * SR_CONTROL_NOISE * n / steps
    return {"ho_race": f / steps + knot_noise, "ho_race_steps_fraction": steps / (n * 2 * ins.n_timesteps)}


def count_drace(pkg, n=256):
    """FP32 flops per rollout-step of the branching DroneRace workload at its first control step
    (oracle op counter over n samples) plus the knot noise of a sample shared by its model
    rollouts and the environment noise per step (the opponent's PD noise pair and the action
    noise of both agents: three Box-Muller pairs)."""
    sys.path.insert(0, HERE)
    import pyoracle
    oracle = pyoracle.Oracle(pkg.abi)
    s = pkg.scenarios.drace_race(n_inst=1)
    ins = s.instances[0]
    NT = s.env.n_models
    kn = np.zeros((NT + 1) * ins.n_knots * 2, np.float32)
    f, steps = oracle.count_drace(s.env, s.cost, s.states[0], ins, kn, np.float32(ins.init_belief[:NT]), n)
    knot_noise = (NT + 1) * ins.n_knots * SR_CONTROL_NOISE * n / steps
    env_noise = 3 * (20 + 21 + 5)
    return {"drace_race": f / steps + knot_noise + env_noise, "drace_race_steps_fraction": steps / (n * NT * ins.n_timesteps)}


if __name__ == "__main__":
    pkg = package()
    print(f"{count(pkg.scenarios.c1_four_drones()):.2f} FP32 flops per sample-step (4 drones, K=16384, T=100)")
    print(f"{count(pkg.scenarios.c4_eight_drones(4096)):.2f} FP32 flops per sample-step (8 drones, K=4096, T=100)")
    print(f"{count_refine(pkg.scenarios.c4_eight_drones(4096))} FP32 flops per refinement launch (8 drones, T=100, 3 iterates)")
    for k, v in {**count_sr(pkg), **count_ho(pkg), **count_drace(pkg)}.items():
        print(f"{k}: {v:.2f}")
