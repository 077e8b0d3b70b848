"""Benchmark of the MPPI control step on one MI355X.

Headline (the metric: control steps/s of batched adjoint-refined MPPI against
plain MPPI at equal mean trajectory cost): 8 drones racing on a 12-gate
lemniscate (half-width 30 m), one controller each, T=100, FP32, closed loop
(each control step ends with every drone integrated with its action).
  adjoint_k4096  MPPI at K=4096 followed by 3 projected gradient steps (0.05,
                 checkpoints every 10) on the weighted nominal: `value`
  plain_k16384   plain MPPI at K=16384: the comparison (`vs_baseline` =
                 value / its control steps/s, measured in the same run)
  plain_k4096    plain MPPI at K=4096: what the gradient steps add
A "step" is one control step of the whole batch; control steps/s count every
controller (8 per batch step).  Quality: the mean trajectory cost of the
executed closed loop over 1000 steps and 4 seeds per leg.

Also reported: plain MPPI on 4 drones at K=16384 (c1), the 64-instance batch
against per-instance launches (c2), PR-MPPI (c3), the single-quadrotor
workload's 500 steps on host and GPU (c0), the reference's own StratRace2D
controllers, and the oracle's OpenMP CPU runs of the same workloads.

Output: one JSON line (driver contract) with a roofline object for the
dominant kernel of the headline step and the CPU run of the headline workload
as the baseline.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "control steps/s, batched+adjoint vs plain MPPI at equal mean trajectory cost"

# FP32 flops per rollout sample-step (dynamics, gates and costs counted by the
# oracle's op-counting scalar on the first control step of this workload, plus
# 96 for the Philox Box-Muller noise).  tests/test_bench_constants.py keeps it
# equal to oracle/count_ops.py.
FLOPS_PER_SAMPLE_STEP = 292.07
# The same for the 8-drone workload at K=4096 (seven opponents in every running cost), and the
# oracle-counted flops of one adjoint-refinement launch over its 8 controllers (oracle/count_ops.py
# count_refine); tests/test_bench_constants.py keeps both equal to the counter.
C4_FLOPS_PER_SAMPLE_STEP = 356.12
C4_REFINE_FLOPS = 2420640
# The noise part of the per-sample-step figures (Box-Muller and sigma scaling; oracle/count_ops.py
# NOISE_FLOPS_PER_STEP).  With the noise rows made ahead (rollout_split 4) the refinement launch
# makes them for the next step and the rollout reads them back.
NOISE_FLOPS_PER_SAMPLE_STEP = 96
# HBM bytes per rollout sample-step the rollout kernel must move: the float4
# noise row it writes (read back once by the weighted average).
ROLLOUT_BYTES_PER_SAMPLE_STEP = 16.0
# FP32 flops per rollout-step of the StratRace2D, HiddenObs and DroneRace workloads (oracle op
# counter at the first control step, plus noise; oracle/count_ops.py count_sr, count_ho and
# count_drace).  The branching
# rollouts stop at terminal events: their launches are priced by the rollout-steps they ran, which
# the engine reports per rollout.
SR_FLOPS_PER_STEP = {"sr_race": 1044.30, "pr_sr_race": 1010.14, "bm_race": 1513.46, "ho_race": 689.45, "drace_race": 1320.55}
PEAK_FP32_TFLOPS = 157.3
# Without packed FP32 (this build compiles with it off) a lane retires one FP32 FMA per cycle: half of the published figure.
PEAK_FP32_UNPACKED_TFLOPS = PEAK_FP32_TFLOPS / 2
PEAK_HBM_GBS = 8000.0


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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(pkg, scn, budget_s):
    """The oracle's FP32 restatement, OpenMP over samples, on this host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle

    path, flags = pyoracle.timing_build()
    oracle = pyoracle.Oracle(pkg.abi, path)
    threads = min(16, os.cpu_count() or 1)
    drones = (pkg.abi.DroneState * len(scn.drone_world))(*scn.drones)
    nominal = np.concatenate([np.tile(np.array(i.shift_fill[:], np.float32), (i.n_timesteps, 1)).reshape(-1)
                              for i in scn.instances]).astype(np.float32)
    steps = [0] * len(scn.instances)
    lams = [i.lambda_ for i in scn.instances]
    t0 = time.perf_counter()
    oracle.world_run(scn, drones, nominal, steps, lams, 1, threads)
    one = time.perf_counter() - t0
    n = max(1, int(budget_s / max(one, 1e-6)))
    t0 = time.perf_counter()
    oracle.world_run(scn, drones, nominal, steps, lams, n, threads)
    dt = time.perf_counter() - t0
    n_inst = len(scn.instances)
    return {"value": n_inst * n / dt, "unit": "control steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} closed-loop batch steps ({n_inst} controllers each) of the same workload after 1 warm-up "
                      f"step, oracle FP32 restatement built {flags}, OpenMP {threads} threads, {dt:.1f} s",
            "sample_steps_per_s": n * scn.sample_steps_per_step / dt}


def cpu_stratrace(pkg, budget_s=12.0):
    """The oracle's FP32 restatement of the three StratRace2D controllers at the reference sizes,
    OpenMP over samples on this host: one control step (rollouts + reductions + update) from the
    start state, repeated within about budget_s / 3 seconds each."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle

    path, flags = pyoracle.timing_build()
    oracle = pyoracle.Oracle(pkg.abi, path)
    threads = min(16, os.cpu_count() or 1)
    sc = pkg.scenarios
    out = {}

    def timed(fn, work):
        t0 = time.perf_counter()
        fn()
        one = time.perf_counter() - t0
        n = max(1, int(budget_s / 3 / max(one, 1e-6)))
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        dt = time.perf_counter() - t0
        return {"control_steps_per_s": n / dt, "rollout_steps_per_s": n * work / dt, "cores": threads, "kind": "port",
                "sample": f"{n} control steps after 1 warm-up step, oracle FP32 built {flags}, OpenMP {threads} threads, "
                          f"{dt:.1f} s"}

    s1 = sc.sr_race(n_inst=1)
    i1 = s1.instances[0]
    zero = np.zeros((i1.n_timesteps, 2), np.float32)

    def sr_step():
        costs, noise = oracle.sr_rollout(0, s1.env, s1.cost, s1.states[0], i1.true_theta, zero, i1.n_samples, i1.n_timesteps, i1.seed,
                                         0, i1.sigma, threads=threads)
        oracle.sr_update(0, costs, noise, zero, i1.lambda_, s1.env.max_accel[0])
    out["sr_race"] = timed(sr_step, i1.n_samples * i1.n_timesteps)

    s2 = sc.pr_sr_race(n_inst=1)
    i2 = s2.instances[0]
    z2 = np.zeros((i2.n_timesteps, 2), np.float32)
    belief = np.array(i2.belief[:s2.env.n_models], np.float32)

    def prsr_step():
        oracle.prsr_step(0, s2.env, s2.cost, s2.states[0], i2, belief, 0, [i2.lambda_] * 3, z2, z2, threads=threads, want_noise=False,
                         want_pairs=False)
    out["pr_sr_race"] = timed(prsr_step, 2 * i2.n_draws * i2.n_samples * i2.n_timesteps)

    s3 = sc.bm_race(n_inst=1)
    i3 = s3.instances[0]
    kn = np.zeros((s3.env.n_models + 1, i3.n_knots, 2), np.float32)
    b3 = np.array(i3.init_belief[:s3.env.n_models], np.float32)

    def bm_step():
        oracle.bm_step(0, s3.env, s3.cost, s3.states[0], i3, b3, 0, i3.lambda_, kn, threads=threads, want_noise=False)
    out["bm_race"] = timed(bm_step, s3.env.n_models * i3.n_samples * i3.n_timesteps)

    s4 = sc.ho_race(n_inst=1)
    i4 = s4.instances[0]
    k4 = np.zeros((3, i4.n_knots, 2), np.float32)
    b4 = np.array(i4.init_belief[:2], np.float32)

    def ho_step():
        oracle.bm_step(0, s4.env, s4.cost, s4.states[0], i4, b4, 0, i4.lambda_, k4, threads=threads, want_noise=False)
    out["ho_race"] = timed(ho_step, 2 * i4.n_samples * i4.n_timesteps)

    s5 = sc.drace_race(n_inst=1)
    i5 = s5.instances[0]
    k5 = np.zeros((3, i5.n_knots, 2), np.float32)
    b5 = np.array(i5.init_belief[:2], np.float32)

    def drace_step():
        oracle.bm_step(0, s5.env, s5.cost, s5.states[0], i5, b5, 0, i5.lambda_, k5, threads=threads, want_noise=False)
    out["drace_race"] = timed(drace_step, 2 * i5.n_samples * i5.n_timesteps)
    return out


def batch_vs_per_instance(pkg, steps=50, warmup=5):
    """64 uneven controllers (K ~ logU[1024, 32768], T ~ U{30..150}): one
    batched engine (one graph replay per control step of all 64) against 64
    single-controller engines stepped one after the other with plain stream
    launches and a synchronisation per controller, the way the reference's
    getControl runs (mppi_controller.cu:250-466)."""
    sc = pkg.scenarios
    scn = sc.c2_batch64()
    n = len(scn.instances)
    eng = pkg.Engine(scn)
    eng.step(warmup, advance=True)
    t0 = time.perf_counter()
    eng.step(steps, advance=True)
    batched = n * steps / (time.perf_counter() - t0)
    eng.close()
    singles = []
    for i, ins in enumerate(scn.instances):
        ins1 = sc.instance(ins.n_samples, ins.n_timesteps, 0, 0, ins.seed)
        one = sc.Scenario(f"c2_{i}", [scn.tracks[i]], [ins1], [0], (pkg.abi.DroneState * 1)(scn.drones[i]), 1, "")
        e = pkg.Engine(one)
        e.set_launch(use_graph=0)
        singles.append(e)
    for e in singles:
        e.step(warmup, advance=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        for e in singles:
            e.step(1, advance=True)
    per_instance = n * steps / (time.perf_counter() - t0)
    for e in singles:
        e.close()
    return {"workload": scn.name, "batched_control_steps_per_s": batched,
            "per_instance_control_steps_per_s": per_instance, "speedup": batched / per_instance,
            "sample_steps_per_batch_step": scn.sample_steps_per_step, "steps": steps}


# The three legs of the adjoint-vs-plain comparison: name, samples per controller, gradient steps.
C4_LEGS = (("plain_k16384", 16384, 0), ("plain_k4096", 4096, 0), ("adjoint_k4096", 4096, 3))
C4_GRAD_STEP, C4_CKPT = 0.05, 10


def start_progress(scn):
    """Race progress of every controlled drone at the start (gate + n_gates laps + 0.8 of the way from
    the last gate passed to the next, env_dronerace.h:374-394), for the progress made in a run."""
    import math
    out = []
    for ins in scn.instances:
        d, tr = scn.drones[ins.drone], scn.tracks[ins.track]
        nxt = 0 if d.gate + 1 == tr.n_gates else d.gate + 1
        a, b = tr.gates[d.gate].center, tr.gates[nxt].center
        g2 = sum((b[c] - d.p[c]) ** 2 for c in range(3))
        c2 = sum((b[c] - a[c]) ** 2 for c in range(3))
        out.append(d.gate + tr.n_gates * d.laps + 0.8 * (1.0 - math.sqrt(g2 / c2)))
    return out


def c4_speed(pkg, steps, warmup, autotune=True, launch=None, legs=None):
    """Throughput of each leg on 8 drones (12-gate lemniscate, seed 1), closed loop, graph replays:
    autotuned launch shape, `warmup` untimed steps, then exactly `steps` timed batch steps between two
    synchronisations; then the same number of steps again with HIP events between the kernels for
    the per-kernel device time."""
    import torch
    sc = pkg.scenarios
    out = {}
    for name, K, refine in C4_LEGS:
        if legs and name not in legs:
            continue
        scn = sc.c4_eight_drones(K)
        n = len(scn.instances)
        eng = pkg.Engine(scn)
        eng.set_refine(refine, C4_GRAD_STEP, C4_CKPT)
        tuned = None
        if launch:
            eng.set_launch(**launch)
        elif autotune:
            tuned, _ = eng.autotune(trials=5)
        used = {f: getattr(eng.get_launch(), f) for f, _ in pkg.abi.LaunchCfg._fields_}
        eng.step(warmup, advance=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.step_async(steps, advance=True)
        eng.synchronize()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        n_timed = max(steps, 20)
        eng.set_timing(True)
        eng.step(n_timed, advance=True)
        eng.set_timing(False)
        kt = {k: v / n_timed for k, v in eng.last_timing().items()}
        eng.close()
        out[name] = {"samples": K, "grad_steps": refine, "control_steps_per_s": n * steps / dt, "ms_per_batch_step": 1e3 * dt / steps,
                     "kernel_ms_per_step": kt, "launch": used, "autotuned": tuned is not None,
                     "sample_steps_per_s": scn.sample_steps_per_step * steps / dt}
        log(f"bench: {name}: {out[name]['control_steps_per_s']:.0f} control steps/s ({out[name]['ms_per_batch_step']:.4f} ms per "
            f"batch step of {n}); device ms per step {json.dumps({k: round(v, 4) for k, v in kt.items()})}")
    return out


def c4_quality(pkg, launches, seeds=(1, 2, 3, 4), n_steps=1000):
    """Mean trajectory cost of each leg over n_steps closed-loop control steps from the start of the
    8-drone race, for each seed (the seed draws the track's gate jitter and the controllers' noise
    keys): per controller, the running cost of every executed state (undiscounted, the engine's own
    cost function on the device) minus the final progress weight times the progress made (gates
    passed, fractional), i.e. the controllers' objective evaluated on the trajectory actually
    flown.  Lower is better.  Reported: the mean over seeds and controllers, the per-seed means and
    their spread, gates passed and executed states in collision."""
    import numpy as np
    sc = pkg.scenarios
    out = {}
    for name, K, refine in C4_LEGS:
        per_seed, gates, colls, costs = [], [], [], []
        for seed in seeds:
            scn = sc.c4_eight_drones(K, seed=seed)
            eng = pkg.Engine(scn)
            eng.set_launch(**launches[name])
            eng.set_refine(refine, C4_GRAD_STEP, C4_CKPT)
            eng.set_exec_cost(True)
            p0 = start_progress(scn)
            eng.step(n_steps, advance=True)
            d = eng.diag()
            eng.close()
            w = scn.cost.final_adv_weight
            prog = [x["exec_progress"] - a for x, a in zip(d, p0)]
            J = [x["exec_cost"] - w * g for x, g in zip(d, prog)]
            per_seed.append(float(np.mean(J)))
            gates += prog
            costs += J
            colls += [x["exec_collisions"] for x in d]
        out[name] = {"mean_traj_cost": float(np.mean(costs)), "per_seed_mean_cost": per_seed,
                     "seed_std": float(np.std(per_seed, ddof=1)) if len(per_seed) > 1 else 0.0,
                     "gates_passed_mean": float(np.mean(gates)), "gates_passed_min": float(np.min(gates)),
                     "collision_states_mean": float(np.mean(colls)), "controllers": len(costs)}
        log(f"bench: quality {name}: mean trajectory cost {out[name]['mean_traj_cost']:.2f} (per seed "
            f"{[round(x, 2) for x in per_seed]}), gates passed {out[name]['gates_passed_mean']:.2f}, "
            f"collision states {out[name]['collision_states_mean']:.2f}")
    out["seeds"] = list(seeds)
    out["steps"] = n_steps
    return out


def vector_roofline(kernel, flops, ms, note):
    """FP32 vector (VALU) roofline of one kernel: algorithmic flops per launch over its HIP-event time.
    peak = the published FP32 vector figure, which counts packed FP32 (two lanes per instruction);
    this build compiles with packed FP32 off, so the unpacked ceiling is half of it."""
    achieved = flops / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    return {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
            "peak_unpacked": PEAK_FP32_UNPACKED_TFLOPS, "frac_unpacked": achieved / PEAK_FP32_UNPACKED_TFLOPS, "traffic": None,
            "kernel": kernel, "flops_per_launch": flops, "ms_per_launch": ms, "note": note}


def c4_rooflines(pkg, adj):
    """Rooflines of the adjoint leg's kernels from its per-kernel HIP-event times; the dominant one
    (largest device time per step) is the headline roofline.  With the noise rows made ahead
    (rollout_split 4), the rollout's flops exclude the noise and read its rows back, and the
    refinement launch (8 latency-bound refinement workgroups + the next step's noise rows on the
    CUs they leave idle) counts the noise flops and writes the rows."""
    sc = pkg.scenarios
    scn = sc.c4_eight_drones(4096)
    ss = scn.sample_steps_per_step
    kt = adj["kernel_ms_per_step"]
    ahead = adj.get("launch", {}).get("rollout_split") == 4
    roll_flops = (C4_FLOPS_PER_SAMPLE_STEP - (NOISE_FLOPS_PER_SAMPLE_STEP if ahead else 0)) * ss
    ref_flops = C4_REFINE_FLOPS + (NOISE_FLOPS_PER_SAMPLE_STEP * ss if ahead else 0)
    r = {"k_rollout": vector_roofline("k_rollout", roll_flops, kt["rollout_ms"],
                                      "oracle-counted flops per sample-step (8 drones: 7 opponents in the cost)" +
                                      ("; the noise rows are read back (made by the previous refinement launch)" if ahead else "")),
         "k_refine": vector_roofline("k_refine", ref_flops, kt["refine_ms"],
                                     "oracle-counted flops of the one-lane adjoint (3 iterates of J and dJ/dU + the last J) per "
                                     "launch of 8 controllers, bound by its serial per-step dependency chain, not by issue" +
                                     ("; plus the next step's noise rows (96 flops per sample-step, Philox integer work "
                                      "uncounted) made by the launch's other workgroups on the CUs the refinement leaves idle"
                                      if ahead else ""))}
    if ahead:
        r["k_refine"]["parts"] = {"refine_flops": C4_REFINE_FLOPS, "noise_flops": NOISE_FLOPS_PER_SAMPLE_STEP * ss,
                                  "noise_bytes_written": ROLLOUT_BYTES_PER_SAMPLE_STEP * ss}
    wavg_bytes = ROLLOUT_BYTES_PER_SAMPLE_STEP * ss
    ms = kt["reduce_ms"]
    ach = wavg_bytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    r["k_wavg"] = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS, "traffic": None,
                   "kernel": "k_wavg", "bytes_per_launch": wavg_bytes, "ms_per_launch": ms,
                   "note": "reads every noise row once (16 B per sample-step)"}
    dom = max(("k_rollout", kt["rollout_ms"]), ("k_wavg", kt["reduce_ms"]), ("k_refine", kt["refine_ms"]), key=lambda x: x[1])[0]
    tfile = os.path.join(ROOT, "profiles", "traffic_c4_8drones_k4096_t100.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tr = json.load(f)
        if tr.get("launch", {}).get("rollout_split") == adj.get("launch", {}).get("rollout_split"):
            for k in r:
                if k in tr:
                    r[k]["traffic"] = tr[k]
    return dom, r


def cpu_c4(pkg, budget_s):
    """The adjoint leg on the host: the oracle's FP32 restatement of MPPI at K=4096 followed by the
    same 3 projected gradient steps (FP32 adjoint, lowest J kept), closed loop, 8 drones, OpenMP
    over samples, timed build (-O3) of the oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle

    path, flags = pyoracle.timing_build()
    oracle = pyoracle.Oracle(pkg.abi, path)
    threads = min(16, os.cpu_count() or 1)
    scn = pkg.scenarios.c4_eight_drones(4096)
    n_inst = len(scn.instances)
    drones = (pkg.abi.DroneState * len(scn.drone_world))(*scn.drones)
    nominal = np.concatenate([np.tile(np.array(i.shift_fill[:], np.float32), (i.n_timesteps, 1)).reshape(-1)
                              for i in scn.instances]).astype(np.float32)
    steps, lams = [0] * n_inst, [i.lambda_ for i in scn.instances]
    refine = (3, C4_GRAD_STEP, C4_CKPT)
    t0 = time.perf_counter()
    oracle.world_run(scn, drones, nominal, steps, lams, 1, threads, refine=refine)
    one = time.perf_counter() - t0
    n = max(1, int(budget_s / max(one, 1e-6)))
    t0 = time.perf_counter()
    oracle.world_run(scn, drones, nominal, steps, lams, n, threads, refine=refine)
    dt = time.perf_counter() - t0
    return {"value": n_inst * n / dt, "unit": "control steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} closed-loop batch steps ({n_inst} controllers each: MPPI K=4096 T=100 + 3 gradient steps) after "
                      f"1 warm-up step, oracle FP32 restatement built {flags}, OpenMP {threads} threads, {dt:.1f} s",
            "sample_steps_per_s": n * scn.sample_steps_per_step / dt}


def c0_runs(pkg, n_steps=500):
    """The single-quadrotor workload (K=1024, T=50, 8-gate circle r=20 m, seed 0) for its 500 control
    steps: on the host (oracle FP32 restatement, timed build, OpenMP) and on the GPU (graph replays)."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    path, flags = pyoracle.timing_build()
    oracle = pyoracle.Oracle(pkg.abi, path)
    threads = min(16, os.cpu_count() or 1)
    scn = pkg.scenarios.c0_single()
    ins = scn.instances[0]
    drones = (pkg.abi.DroneState * 1)(*scn.drones)
    nominal = np.tile(np.array(ins.shift_fill[:], np.float32), (ins.n_timesteps, 1)).reshape(-1)
    t0 = time.perf_counter()
    oracle.world_run(scn, drones, nominal, [0], [ins.lambda_], n_steps, threads)
    cpu_dt = time.perf_counter() - t0
    cpu_gates = drones[0].gate + scn.tracks[0].n_gates * drones[0].laps
    eng = pkg.Engine(scn)
    eng.step(1, advance=True)
    eng.close()
    eng = pkg.Engine(scn)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step_async(n_steps, advance=True)
    eng.synchronize()
    gpu_dt = time.perf_counter() - t0
    g = eng.get_drones()[0]
    eng.close()
    return {"workload": scn.name, "steps": n_steps,
            "cpu": {"control_steps_per_s": n_steps / cpu_dt, "seconds": cpu_dt, "cores": threads, "kind": "port",
                    "build": flags, "final_gate_count": cpu_gates},
            "gpu": {"control_steps_per_s": n_steps / gpu_dt, "seconds": gpu_dt,
                    "final_gate_count": g.gate + scn.tracks[0].n_gates * g.laps}}


def tune_launch(eng, autotune):
    """Launch shape the engine's autotuner keeps (it restores the plans and state it steps), with the
    ms per step it measured; None with autotuning off (the engine's default launch)."""
    if not autotune:
        return None
    cfg, ms = eng.autotune(trials=3)
    log(f"bench: autotuned launch {cfg} ({ms:.3f} ms/step)")
    return {"launch": cfg, "ms_per_step": ms}


def pr_mppi(pkg, steps=100, warmup=10, autotune=True, launch=None):
    """PR-MPPI, 4 drones, B ~ U{3..6} opponent-velocity branches per controller,
    Dirichlet(1) beliefs, K = 8192 samples evaluated under every branch for
    both plans, T = 80, closed loop, graph replays; launch shape given, or autotuned."""
    sc = pkg.scenarios
    scn = sc.c3_pr_four_drones()
    n = len(scn.instances)
    eng = pkg.PREngine(scn)
    if launch:
        eng.set_launch(**launch)
        tuned = {"launch": launch, "ms_per_step": None}
    else:
        tuned = tune_launch(eng, autotune)
    eng.step(warmup, advance=True)
    t0 = time.perf_counter()
    eng.step(steps, advance=True)
    elapsed = time.perf_counter() - t0
    eng.set_timing(True)
    eng.step(10, advance=True)
    ms = eng.last_timing() / 10
    eng.set_timing(False)
    diag = eng.diag()
    eng.close()
    rollouts = sum(2 * i.n_branches * i.n_samples for i in scn.instances)
    ss = sum(2 * i.n_branches * i.n_samples * i.n_timesteps for i in scn.instances)
    return {"workload": scn.name, "autotuned": tuned, "branches": [i.n_branches for i in scn.instances],
            "control_steps_per_s": n * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "rollouts_per_step": rollouts, "rollout_sample_steps_per_s": ss * steps / elapsed,
            "kernel_ms_per_step": {"rollout": float(ms[0]), "aggregate": float(ms[1]), "weighted_avg": float(ms[2]),
                                   "update_choice_advance": float(ms[3]), "step": float(ms[4])},
            "use_nom_plan": [int(d.use_nom_plan) for d in diag], "steps": steps}


def sr_roofline(key, rollout_steps, rollout_ms, kernel):
    """FP32 roofline of a StratRace2D rollout kernel: counted flops per rollout-step times the
    rollout-steps of one launch, over its HIP-event time."""
    return vector_roofline(kernel, SR_FLOPS_PER_STEP[key] * rollout_steps, rollout_ms,
                           "FP32 vector (VALU) work, no MFMA; oracle-counted flops per rollout-step")


def sr_race(pkg, steps=50, warmup=5, n_inst=1, autotune=True):
    """Plain MPPI on the reference's StratRace2D race at its dense setting
    (K = 2^16, T = 70, temperature 10, sigma 1.5, round track r = 6 m), closed
    loop with the environment's noise, graph replays; n_inst worlds."""
    sc = pkg.scenarios
    scn = sc.sr_race(n_inst=n_inst)
    eng = pkg.SREngine(scn)
    tuned = tune_launch(eng, autotune)
    eng.step(warmup, advance=True)
    t0 = time.perf_counter()
    eng.step(steps, advance=True)
    elapsed = time.perf_counter() - t0
    eng.set_timing(True)
    eng.step(5, advance=True)
    ms = eng.last_timing() / 5
    eng.close()
    ss = scn.sample_steps_per_step
    return {"workload": scn.name, "autotuned": tuned, "control_steps_per_s": n_inst * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "sample_steps_per_s": ss * steps / elapsed,
            "kernel_ms_per_step": {"rollout": float(ms[0]), "weighted_avg": float(ms[1]), "update_world": float(ms[2]),
                                   "step": float(ms[3])},
            "roofline": sr_roofline("sr_race", ss, float(ms[0]), "k_sr_rollout"), "steps": steps}


def pr_sr_race(pkg, steps=30, warmup=3, autotune=True):
    """PR-MPPI on the reference's StratRace2D race at its PR setting (K = 2^16, T = 60,
    P = 9 models drawn from the belief, temperature 10, sigma 0.5), closed loop, one world."""
    sc = pkg.scenarios
    scn = sc.pr_sr_race(n_inst=1)
    eng = pkg.PRSREngine(scn)
    tuned = tune_launch(eng, autotune)
    eng.step(warmup, advance=True)
    t0 = time.perf_counter()
    eng.step(steps, advance=True)
    elapsed = time.perf_counter() - t0
    eng.set_timing(True)
    eng.step(5, advance=True)
    ms = eng.last_timing() / 5
    eng.close()
    ins = scn.instances[0]
    rollouts = 2 * ins.n_draws * ins.n_samples
    return {"workload": scn.name, "autotuned": tuned, "control_steps_per_s": steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "rollouts_per_step": rollouts, "rollout_env_steps_per_s": rollouts * ins.n_timesteps * steps / elapsed,
            "kernel_ms_per_step": {"rollout": float(ms[0]), "aggregate": float(ms[1]), "weighted_avg": float(ms[2]),
                                   "update_choice_world": float(ms[3]), "step": float(ms[4])},
            "roofline": sr_roofline("pr_sr_race", rollouts * ins.n_timesteps, float(ms[0]), "k_prsr_rollout"), "steps": steps}


def timed_branching_steps(eng, n=5):
    """n more closed-loop steps, one at a time, with HIP-event timing: the mean device ms of each
    phase and the mean rollout-steps the rollout launch ran (the recording rollout is the timed one)."""
    import numpy as np
    eng.set_record_lengths(True)
    eng.set_timing(True)
    ms = np.zeros(6)
    ran = 0
    for _ in range(n):
        eng.step(1, advance=True)
        ms += eng.last_timing()
        ran += sum(int(eng.rollout_lengths(i).sum()) for i in range(eng.n_inst))
    eng.set_timing(False)
    eng.set_record_lengths(False)
    return ms / n, ran / n


def bm_race(pkg, steps=30, warmup=3, n_verif=0, n_inst=1, autotune=True):
    """Branching spline MPPI (the reference's default MPPI controller) on its StratRace2D race at
    the spline setting (K = 2^16, T = 80, 20 knots, two opponent models, uniform initial belief,
    temperature 10 adapted, sigma 0.5), closed loop with belief updates, n_inst worlds batched in
    one graph (control steps/s counts every world's step); n_verif > 0
    adds the plan verification at mainGate's setting (2^17 samples, horizon 40, beta 1e-5,
    maxVerifEps 0.001)."""
    sc = pkg.scenarios
    scn = sc.bm_race(n_inst=n_inst, n_verif=n_verif)
    eng = pkg.BMEngine(scn)
    tuned = tune_launch(eng, autotune)
    eng.step(warmup, advance=True)
    t0 = time.perf_counter()
    eng.step(steps, advance=True)
    elapsed = time.perf_counter() - t0
    ms, ran = timed_branching_steps(eng)
    d = eng.diag()[0]
    eng.close()
    ins = scn.instances[0]
    rollouts = n_inst * scn.env.n_models * ins.n_samples
    return {"workload": scn.name, "autotuned": tuned, "control_steps_per_s": n_inst * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "rollouts_per_step": rollouts, "rollout_env_steps_per_s": rollouts * ins.n_timesteps * steps / elapsed,
            "kernel_ms_per_step": {"rollout": float(ms[0]), "aggregate_min": float(ms[1]), "weighted_avg": float(ms[2]),
                                   "update_verify": float(ms[3]), "choice_world": float(ms[4]), "step": float(ms[5])},
            "final_branch": int(d.branch), "use_new_plan": int(d.use_new_plan), "certified_loss": float(d.certified_loss),
            "rollout_steps_per_launch": ran, "rollout_steps_fraction": ran / (rollouts * ins.n_timesteps),
            "roofline": sr_roofline("bm_race", ran, float(ms[0]), "k_bm_rollout"),
            "verification_samples": n_verif, "steps": steps}


def ho_race(pkg, steps=30, warmup=3, autotune=True):
    """Branching spline MPPI on the reference's HiddenObs course at its spline setting
    (hiddenObsEnv, main.py:705-833: K = 2^16, T = 80, 20 knots, temperature 10 adapted,
    sigma 0.8, uniform belief over which hidden obstacle exists), closed loop, one world."""
    sc = pkg.scenarios
    scn = sc.ho_race(n_inst=1)
    eng = pkg.HoBMEngine(scn)
    tuned = tune_launch(eng, autotune)
    eng.step(warmup, advance=True)
    t0 = time.perf_counter()
    eng.step(steps, advance=True)
    elapsed = time.perf_counter() - t0
    ms, ran = timed_branching_steps(eng)
    d = eng.diag()[0]
    st = eng.get_states()[0]
    belief = eng.belief(0)
    events = int(eng.events()[0])
    eng.close()
    ins = scn.instances[0]
    rollouts = 2 * ins.n_samples
    return {"workload": scn.name, "autotuned": tuned, "control_steps_per_s": steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "rollouts_per_step": rollouts, "rollout_env_steps_per_s": rollouts * ins.n_timesteps * steps / elapsed,
            "kernel_ms_per_step": {"rollout": float(ms[0]), "aggregate_min": float(ms[1]), "weighted_avg": float(ms[2]),
                                   "update_verify": float(ms[3]), "choice_world": float(ms[4]), "step": float(ms[5])},
            "final_branch": int(d.branch), "final_state": {"pos": [float(st.pos[0]), float(st.pos[1])], "gates": int(st.gates),
                                                           "laps": int(st.laps), "event": events},
            "final_belief": [float(b) for b in belief],
            "rollout_steps_per_launch": ran, "rollout_steps_fraction": ran / (rollouts * ins.n_timesteps),
            "roofline": sr_roofline("ho_race", ran, float(ms[0]), "k_bm_rollout<HoPolicy>"),
            "steps": steps}


def drace_race(pkg, steps=30, warmup=3, autotune=True):
    """Branching spline MPPI on the reference's DroneRace course at its spline setting
    (tinyGateEnv, main.py:40-161: K = 2^16, T = 60, 10 knots, temperature 10 adapted, sigma 2,
    two PD opponent models on the upper and lower raceline, uniform belief), closed loop, one
    world."""
    sc = pkg.scenarios
    scn = sc.drace_race(n_inst=1)
    eng = pkg.DraceBMEngine(scn)
    tuned = tune_launch(eng, autotune)
    eng.step(warmup, advance=True)
    t0 = time.perf_counter()
    eng.step(steps, advance=True)
    elapsed = time.perf_counter() - t0
    ms, ran = timed_branching_steps(eng)
    d = eng.diag()[0]
    st = eng.get_states()[0]
    belief = eng.belief(0)
    eng.close()
    ins = scn.instances[0]
    rollouts = scn.env.n_models * ins.n_samples
    return {"workload": scn.name, "autotuned": tuned, "control_steps_per_s": steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "rollouts_per_step": rollouts, "rollout_env_steps_per_s": rollouts * ins.n_timesteps * steps / elapsed,
            "kernel_ms_per_step": {"rollout": float(ms[0]), "aggregate_min": float(ms[1]), "weighted_avg": float(ms[2]),
                                   "update_verify": float(ms[3]), "choice_world": float(ms[4]), "step": float(ms[5])},
            "final_branch": int(d.branch), "final_state": {"pos": [float(v) for v in st.pos], "gates": list(st.gates),
                                                           "laps": list(st.laps)},
            "final_belief": [float(b) for b in belief],
            "rollout_steps_per_launch": ran, "rollout_steps_fraction": ran / (rollouts * ins.n_timesteps),
            "roofline": sr_roofline("drace_race", ran, float(ms[0]), "k_bm_rollout<DracePolicy>"),
            "steps": steps}


def record_lengths_run(pkg, env, steps=4):
    """`steps` closed-loop steps from the start of one branching race (bm: StratRace2D, ho:
    HiddenObs, drace: DroneRace; one world at the reference sizes) with rollout-length recording on
    from the first step: the rollout-steps each rollout launch ran, so that counter totals over
    those launches can be put per rollout-step."""
    sc = pkg.scenarios
    scn, cls = {"bm": (sc.bm_race, pkg.BMEngine), "ho": (sc.ho_race, pkg.HoBMEngine),
                "drace": (sc.drace_race, pkg.DraceBMEngine)}[env]
    scn = scn(n_inst=1)
    eng = cls(scn)
    eng.set_record_lengths(True)
    ran, slots = [], 0
    for _ in range(steps):
        eng.step(1, advance=True)
        lens = [eng.rollout_lengths(i) for i in range(eng.n_inst)]
        ran.append(sum(int(a.sum()) for a in lens))
        slots = sum(int(a.size) for a in lens) * scn.instances[0].n_timesteps
    eng.set_record_lengths(False)
    eng.close()
    return {"workload": scn.name, "env": env, "steps": steps, "rollout_steps_per_launch": ran, "rollout_slots_per_launch": slots}


def session_race(pkg, steps=30):
    """The simulation session behind the wire protocol (SimulationEngine::run) at the reference's
    settings: the branching spline MPPI header (K = 2^16, T = 80, 20 knots) and the PR-MPPI header
    (K = 2^16, T = 60, P = 9), two opponent models; per control step getControl + dynStep + the
    STATE frame with its display predictions, in process (no socket)."""
    wire = pkg.wire
    e = wire.stratrace_env(n_models=2, seed=42)
    out = {}
    for name, cont in (("mppi", wire.mppi_header()), ("prmppi", wire.prmppi_header())):
        sess = wire.Session(wire.env_header(e), cont, wire.init_state_header(e), max_steps=steps)
        first = sess.next_frame()  # step-0 STATE
        t0 = time.perf_counter()
        n_state, n_bytes, last = 0, 0, None
        for t, frame in sess:
            if t == 1:
                n_state += 1
                n_bytes += len(frame)
            last = t
        elapsed = time.perf_counter() - t0
        sess.close()
        out[name] = {"control_steps_per_s": n_state / elapsed, "ms_per_step": 1e3 * elapsed / max(n_state, 1),
                     "state_frame_bytes": n_bytes / max(n_state, 1), "steps": n_state, "first_frame_bytes": len(first[1]),
                     "last_frame_type": last}
    return out


def plain_run(pkg, scn, steps, warmup, launch=None, autotune=True):
    """Plain MPPI on a workload: launch shape (given or autotuned), `warmup` steps, exactly `steps`
    timed closed-loop batch steps between two synchronisations, then per-kernel device time from HIP
    events over as many steps, and the rollout kernel's roofline."""
    import torch
    n_inst = len(scn.instances)
    eng = pkg.Engine(scn)
    tuned = None
    if launch:
        eng.set_launch(**launch)
    elif autotune:
        tuned, ms = eng.autotune(trials=5)
        log(f"bench: {scn.name}: autotuned launch {tuned} ({ms:.3f} ms/step)")
    used = {f: getattr(eng.get_launch(), f) for f, _ in pkg.abi.LaunchCfg._fields_}
    eng.step(warmup, advance=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step_async(steps, advance=True)
    eng.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n_timed = min(max(steps, 20), 200)
    eng.set_timing(True)
    eng.step(n_timed, advance=True)
    eng.set_timing(False)
    kt = {k: v / n_timed for k, v in eng.last_timing().items()}
    eng.close()
    ss = scn.sample_steps_per_step
    flops = FLOPS_PER_SAMPLE_STEP * ss
    roof = vector_roofline("k_rollout", flops, kt["rollout_ms"], "oracle-counted flops per sample-step (dynamics, gates, costs, noise)")
    tfile = os.path.join(ROOT, "profiles", f"traffic_{scn.name}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            roof["traffic"] = json.load(f).get("rollout_hbm_bytes_per_launch")
    return {"workload": scn.name, "control_steps_per_s": n_inst * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "sample_steps_per_s": ss * steps / elapsed, "kernel_ms_per_step": kt, "launch": used, "autotuned": tuned is not None,
            "roofline": roof, "steps": steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--launch", default="", help="launch shape of the plain 4-drone run as JSON (skips its autotuner)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=8.0)
    ap.add_argument("--no-quality", action="store_true", help="skip the closed-loop trajectory-cost runs of the three legs")
    ap.add_argument("--quality-steps", type=int, default=1000)
    ap.add_argument("--quality-seeds", type=int, default=4)
    ap.add_argument("--no-batch64", action="store_true")
    ap.add_argument("--no-pr", action="store_true")
    ap.add_argument("--no-sr", action="store_true")
    ap.add_argument("--plain-only", action="store_true", help="run only plain MPPI on 4 drones, K=16384 (for profiling)")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the throughput of the three 8-drone legs (for profiling)")
    ap.add_argument("--c4-legs", default="", help="comma-separated subset of the 8-drone legs to run with --adjoint-only")
    ap.add_argument("--c4-launch", default="", help="launch shape of the 8-drone legs as JSON (skips their autotuner)")
    ap.add_argument("--pr-only", action="store_true", help="run only the PR-MPPI section (for profiling)")
    ap.add_argument("--pr-launch", default="", help="launch shape of the PR-MPPI section as JSON (skips its autotuner)")
    ap.add_argument("--sr-only", action="store_true", help="run only the StratRace2D section (for profiling)")
    ap.add_argument("--bm-only", action="store_true", help="run only the branching spline MPPI section (for profiling)")
    ap.add_argument("--session-only", action="store_true", help="run only the wire-protocol session section")
    ap.add_argument("--ho-only", action="store_true", help="run only the HiddenObs and DroneRace sections (for profiling)")
    ap.add_argument("--record-lengths", default="", choices=["", "bm", "ho", "drace"],
                    help="only run 4 closed-loop steps of one branching race (bm: StratRace2D, ho: HiddenObs, drace: "
                         "DroneRace; one world, reference sizes) with rollout-length recording on, and print the "
                         "rollout-steps each rollout launch ran (the program profiled for the instruction-mix counters)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 512 MiB buffer 3 times (reference kernel for the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.empty(512 << 20, dtype=torch.uint8, device="cuda:0").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel(), "copies": 3}))
        return
    if args.record_lengths:
        print(json.dumps(record_lengths_run(package(), args.record_lengths)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (each controller batch runs on a "
              f"single MI355X; sharding across GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)

    import torch

    pkg = package()
    sc = pkg.scenarios
    torch.cuda.set_device(0)
    if args.session_only:
        print(json.dumps({"session": session_race(pkg, steps=min(args.steps, 30))}), flush=True)
        return
    if args.ho_only:
        print(json.dumps({"ho_race": ho_race(pkg, steps=min(args.steps, 30), warmup=args.warmup, autotune=not args.no_autotune),
                          "drace_race": drace_race(pkg, steps=min(args.steps, 30), warmup=args.warmup, autotune=not args.no_autotune)}), flush=True)
        return
    if args.bm_only:
        print(json.dumps({"bm_race": bm_race(pkg, steps=min(args.steps, 30), warmup=args.warmup, autotune=not args.no_autotune),
                          "bm_race_verified": bm_race(pkg, steps=min(args.steps, 30), warmup=args.warmup, n_verif=2 ** 17, autotune=not args.no_autotune)}),
              flush=True)
        return
    if args.sr_only:
        print(json.dumps({"sr_race": sr_race(pkg, steps=args.steps, warmup=args.warmup, autotune=not args.no_autotune),
                          "pr_sr_race": pr_sr_race(pkg, steps=min(args.steps, 30), warmup=args.warmup, autotune=not args.no_autotune),
                          "bm_race": bm_race(pkg, steps=min(args.steps, 30), warmup=args.warmup, autotune=not args.no_autotune),
                          "bm_race_verified": bm_race(pkg, steps=min(args.steps, 30), warmup=args.warmup, n_verif=2 ** 17, autotune=not args.no_autotune)}),
              flush=True)
        return
    if args.pr_only:
        launch = json.loads(args.pr_launch) if args.pr_launch else None
        print(json.dumps({"pr_mppi": pr_mppi(pkg, steps=args.steps, warmup=args.warmup, autotune=not args.no_autotune, launch=launch)}),
              flush=True)
        return
    if args.plain_only:
        launch = json.loads(args.launch) if args.launch else None
        print(json.dumps({"c1": plain_run(pkg, sc.c1_four_drones(), args.steps, args.warmup, launch, not args.no_autotune)}), flush=True)
        return
    if args.adjoint_only:
        legs = [x for x in args.c4_legs.split(",") if x] or None
        launch = json.loads(args.c4_launch) if args.c4_launch else None
        print(json.dumps({"c4": c4_speed(pkg, args.steps, args.warmup, not args.no_autotune, launch, legs)}), flush=True)
        return

    log("bench: 8 drones, K=16384 plain / K=4096 plain / K=4096 + 3 adjoint gradient steps: throughput ...")
    legs = c4_speed(pkg, args.steps, args.warmup, not args.no_autotune)
    adj, base = legs["adjoint_k4096"], legs["plain_k16384"]
    dom, roofs = c4_rooflines(pkg, adj)
    scn4 = sc.c4_eight_drones(4096)
    out = {
        "metric": METRIC,
        "value": adj["control_steps_per_s"],
        "unit": "control steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": adj["ms_per_batch_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": adj["control_steps_per_s"] / base["control_steps_per_s"],
        "vs_baseline_ref": "plain MPPI at K=16384 on the same 8-drone workload, measured in this run (no published figure "
                           "exists); the comparison the metric names",
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "c4_8drones_k4096_t100_adjoint3", "controllers": len(scn4.instances), "samples": 4096, "horizon": 100,
                   "grad_steps": 3, "grad_step": C4_GRAD_STEP, "ckpt_every": C4_CKPT, "closed_loop": True, "launch": adj["launch"],
                   "autotuned": adj["autotuned"], "baseline_workload": "c4_8drones_k16384_t100 (plain MPPI)"},
        "legs": legs,
        "speedup_vs_plain_k16384": adj["control_steps_per_s"] / base["control_steps_per_s"],
        "speedup_plain_k4096_vs_plain_k16384": legs["plain_k4096"]["control_steps_per_s"] / base["control_steps_per_s"],
        "roofline": roofs[dom],
        "rooflines": roofs,
        "sample_steps_per_s": adj["sample_steps_per_s"],
    }
    if not args.no_quality:
        log(f"bench: trajectory cost of the three legs, {args.quality_steps} closed-loop steps x {args.quality_seeds} seeds ...")
        q = c4_quality(pkg, {k: v["launch"] for k, v in legs.items()}, seeds=tuple(range(1, 1 + args.quality_seeds)),
                       n_steps=args.quality_steps)
        out["quality"] = q
        ja, jb = q["adjoint_k4096"]["mean_traj_cost"], q["plain_k16384"]["mean_traj_cost"]
        out["cost_ratio"] = ja / jb if jb != 0 else None
        out["equal_cost"] = bool(ja <= jb)
        out["cost_ratio_note"] = ("mean trajectory cost of the adjoint leg over that of plain K=16384; costs are negative "
                                  "(progress rewards), so >= 1 means the adjoint leg is at least as good; equal_cost: its "
                                  "mean cost is <= plain K=16384's")
    log("bench: plain MPPI, 4 drones, K=16384, T=100 ...")
    launch = json.loads(args.launch) if args.launch else None
    out["c1"] = plain_run(pkg, sc.c1_four_drones(), args.steps, args.warmup, launch, not args.no_autotune)
    log(f"bench: c1 {json.dumps(out['c1'])}")
    if not args.no_batch64:
        log("bench: 64 uneven controllers, batched vs per-instance ...")
        out["batch64"] = batch_vs_per_instance(pkg)
        out["batch64_speedup"] = out["batch64"]["speedup"]
        log(f"bench: batch64 speedup {out['batch64']['speedup']:.2f}x")
    if not args.no_pr:
        log("bench: PR-MPPI, 4 drones, K=8192 per branch, T=80 ...")
        out["pr_mppi"] = pr_mppi(pkg, autotune=not args.no_autotune)
        log(f"bench: pr_mppi {json.dumps(out['pr_mppi'])}")
    if not args.no_sr:
        log("bench: StratRace2D plain MPPI, K=65536, T=70 ...")
        out["sr_race"] = sr_race(pkg, autotune=not args.no_autotune)
        log("bench: PR-MPPI on StratRace2D, K=65536, T=60, P=9 ...")
        out["pr_sr_race"] = pr_sr_race(pkg, autotune=not args.no_autotune)
        log("bench: branching spline MPPI on StratRace2D, K=65536, T=80, 20 knots ...")
        out["bm_race"] = bm_race(pkg, autotune=not args.no_autotune)
        log("bench: branching spline MPPI, 8 worlds in one batch ...")
        out["bm_race_8x"] = bm_race(pkg, n_inst=8, autotune=not args.no_autotune)
        log("bench: branching spline MPPI with plan verification (2^17 samples, horizon 40) ...")
        out["bm_race_verified"] = bm_race(pkg, n_verif=2 ** 17, autotune=not args.no_autotune)
        log("bench: branching spline MPPI on HiddenObs, K=65536, T=80, 20 knots ...")
        out["ho_race"] = ho_race(pkg, autotune=not args.no_autotune)
        log("bench: branching spline MPPI on DroneRace, K=65536, T=60, 10 knots ...")
        out["drace_race"] = drace_race(pkg, autotune=not args.no_autotune)
        log("bench: wire-protocol session (MPPI and PR-MPPI headers at the reference settings) ...")
        out["session"] = session_race(pkg)
    if not args.no_cpu:
        log("bench: CPU baseline of the headline workload (adjoint leg on the host) ...")
        out["cpu_baseline"] = cpu_c4(pkg, args.cpu_budget)
        log(f"bench: CPU baseline {out['cpu_baseline']['value']:.2f} control steps/s")
        log("bench: single-quadrotor workload, 500 steps on host and GPU ...")
        out["c0"] = c0_runs(pkg)
        log(f"bench: c0 {json.dumps(out['c0'])}")
        out["c1"]["cpu_baseline"] = cpu_baseline(pkg, sc.c1_four_drones(), args.cpu_budget / 2)
        if not args.no_sr:
            log("bench: CPU baselines of the StratRace2D controllers ...")
            out["cpu_baseline_stratrace"] = cpu_stratrace(pkg, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
