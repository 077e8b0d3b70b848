"""ORACLE — test infrastructure only.

ctypes binding of oracle/liboracle.so (the CPU restatement in oracle.h /
oracle.cpp).  Imported by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg as the checker; never by the product.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def timing_build():
    """Path of the -O3 build used to time the CPU baseline: x86-64-v4 (AVX-512) when this host's CPU has
    AVX-512 F/BW/CD/DQ/VL, else x86-64-v3; and the flag set it was built with."""
    flags = set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
    except OSError:
        pass
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags
    name = "liboracle_t4.so" if v4 else "liboracle_t3.so"
    return os.path.join(HERE, name), ("-O3 -march=x86-64-v4" if v4 else "-O3 -march=x86-64-v3") + " -ffp-contract=off -fopenmp"


_libs = {}


def lib(abi, path=None):
    """Load liboracle.so (or another build of the same sources at `path`); `abi` is the engine's ctypes
    struct module."""
    global _lib
    path = path or LIB_PATH
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    F, D, I, U = C.POINTER(C.c_float), C.POINTER(C.c_double), C.c_int, C.c_uint32
    L.or_philox4x32_10.argtypes = [C.POINTER(U), C.POINTER(U), C.POINTER(U)]
    L.or_normals4.argtypes = [U, U, U, U, U, U, F]
    L.or_log32.restype = C.c_float
    L.or_log32.argtypes = [C.c_float]
    L.or_exp32.restype = C.c_float
    L.or_exp32.argtypes = [C.c_float]
    L.or_sincos_turns.argtypes = [U, F, F]
    L.or_inv_sqrt32.restype = C.c_float
    L.or_inv_sqrt32.argtypes = [C.c_float]
    L.or_quad_step.argtypes = [I, C.POINTER(abi.QuadParams), C.POINTER(abi.Track), C.c_float, D, C.POINTER(I),
                               C.POINTER(I), D]
    L.or_progress.restype = C.c_double
    L.or_progress.argtypes = [I, C.POINTER(abi.Track), D, I, I]
    L.or_running_cost.restype = C.c_double
    L.or_running_cost.argtypes = [I, C.POINTER(abi.QuadParams), C.POINTER(abi.CostCfg), C.POINTER(abi.Track), D, I, I,
                                  C.POINTER(abi.Opponent), I, C.c_double, C.POINTER(I)]
    L.or_rollout.argtypes = [I, C.POINTER(abi.QuadParams), C.POINTER(abi.CostCfg), C.POINTER(abi.Track),
                             C.POINTER(abi.DroneState), C.POINTER(abi.Opponent), I, F, I, I, U, U, U, F, D, F, D, I]
    L.or_mppi_update.argtypes = [I, C.POINTER(abi.QuadParams), D, F, F, I, I, C.c_float, F, D, D, D, D, D, D]
    L.or_world_run.argtypes = [C.POINTER(abi.QuadParams), C.POINTER(abi.CostCfg), C.POINTER(abi.Track),
                               C.POINTER(abi.Instance), I, C.POINTER(C.c_int32), C.POINTER(abi.DroneState), I, F,
                               C.POINTER(U), F, I, I, F, F, I, C.c_float, I]
    L.or_cost_grad.argtypes = [I, C.POINTER(abi.QuadParams), C.POINTER(abi.CostCfg), C.POINTER(abi.Track),
                               C.POINTER(abi.DroneState), C.POINTER(abi.Opponent), I, F, I, I, I, D, D,
                               C.POINTER(C.c_uint64), I]
    L.or_pr_step.argtypes = [I, C.POINTER(abi.QuadParams), C.POINTER(abi.CostCfg), C.POINTER(abi.Track),
                             C.POINTER(abi.DroneState), C.POINTER(abi.Opponent), I, C.POINTER(abi.PrInstance), F, U, F, F,
                             F, I, F, D, D, C.POINTER(I), D, D]
    L.or_pr_choose.argtypes = [I, C.POINTER(abi.QuadParams), C.POINTER(abi.CostCfg), C.POINTER(abi.Track),
                               C.POINTER(abi.DroneState), C.POINTER(abi.Opponent), I, C.POINTER(abi.PrInstance), F,
                               C.POINTER(I), D, F, D, D, D, D]
    L.or_sr_env_step.argtypes = [I, C.POINTER(abi.SrEnv), I, F, C.POINTER(abi.SrState), C.c_uint64, U, U, U, I]
    L.or_sr_project.restype = C.c_double
    L.or_sr_project.argtypes = [I, C.POINTER(abi.SrEnv), C.c_double, C.c_double, C.c_double, D]
    L.or_sr_opp_action.argtypes = [I, C.POINTER(abi.SrEnv), I, C.POINTER(abi.SrState), D]
    L.or_sr_state_cost.restype = C.c_double
    L.or_sr_state_cost.argtypes = [I, C.POINTER(abi.SrEnv), C.POINTER(abi.SrCost), C.POINTER(abi.SrState)]
    L.or_sr_rollout.argtypes = [I, C.POINTER(abi.SrEnv), C.POINTER(abi.SrCost), C.POINTER(abi.SrState), I, F, I, I,
                                C.c_uint64, U, C.c_float, I, D, F]
    L.or_sr_update.argtypes = [I, D, F, F, I, I, C.c_float, C.c_float, F, D, D, D, D, D]
    L.or_prsr_step.argtypes = [I, C.POINTER(abi.SrEnv), C.POINTER(abi.SrCost), C.POINTER(abi.SrState), C.POINTER(abi.PrsrInstance), F,
                               U, F, F, F, I, F, D, D, C.POINTER(I), D, D]
    L.or_prsr_choose.argtypes = [I, C.POINTER(abi.SrEnv), C.POINTER(abi.SrCost), C.POINTER(abi.SrState), C.POINTER(abi.PrsrInstance),
                                 F, C.POINTER(I), D, U, F, D, D, D, D]
    L.or_bm_basis.argtypes = [I, C.POINTER(I), I, I, D]
    L.or_sr_update_belief.argtypes = [I, C.POINTER(abi.SrEnv), D, D, D]
    L.or_sr_env_step_belief.argtypes = [I, C.POINTER(abi.SrEnv), I, F, C.POINTER(abi.SrState), F, C.c_uint64, U, U, U, I, U]
    L.or_bm_step.argtypes = [I, C.POINTER(abi.SrEnv), C.POINTER(abi.SrCost), C.POINTER(abi.SrState), C.POINTER(abi.BmInstance), F,
                             U, C.c_float, F, F, I, F, D, C.POINTER(I), C.POINTER(I), D, D, D, D, D, D, C.POINTER(I), C.POINTER(I), D]
    L.or_sr_predict.argtypes = [C.POINTER(abi.SrEnv), C.POINTER(abi.SrState), I, F, I, I, F, C.c_float, F, F, C.POINTER(I)]
    L.or_sr_opp_target.restype = C.c_double
    L.or_sr_opp_target.argtypes = [I, C.POINTER(abi.SrEnv), I, C.POINTER(abi.SrState)]
    L.or_sr_opp_target_goal.restype = C.c_double
    L.or_sr_opp_target_goal.argtypes = [I, C.POINTER(abi.SrEnv), I, C.POINTER(abi.SrState), C.POINTER(C.c_double)]
    L.or_ho_prim.restype = C.c_double
    L.or_ho_prim.argtypes = [I, I, D]
    L.or_ho_eval.restype = C.c_double
    L.or_ho_eval.argtypes = [I, I, C.POINTER(abi.HoEnv), C.POINTER(abi.SrCost), C.POINTER(abi.HoState), I]
    L.or_ho_env_step.argtypes = [I, C.POINTER(abi.HoEnv), I, F, C.POINTER(abi.HoState), F, C.c_uint64, U, U, U, U]
    L.or_bmho_step.argtypes = [I, C.POINTER(abi.HoEnv), C.POINTER(abi.SrCost), C.POINTER(abi.HoState), C.POINTER(abi.BmInstance), F,
                               U, C.c_float, F, F, I, F, D, C.POINTER(I), C.POINTER(I), D, D, D, D, D, D, C.POINTER(I), C.POINTER(I), D]
    L.or_count_ho_ops.restype = C.c_longlong
    L.or_count_ho_ops.argtypes = [C.POINTER(abi.HoEnv), C.POINTER(abi.SrCost), C.POINTER(abi.HoState), F, I, C.POINTER(abi.BmInstance), F,
                                  C.POINTER(C.c_longlong)]
    L.or_drace_project.restype = C.c_double
    L.or_drace_project.argtypes = [I, F, C.c_double, C.c_double, C.c_double]
    L.or_drace_pid.argtypes = [I, C.POINTER(abi.DraceEnv), C.POINTER(abi.DraceState), I, I, D]
    L.or_drace_eval.restype = C.c_double
    L.or_drace_eval.argtypes = [I, I, C.POINTER(abi.DraceEnv), C.POINTER(abi.SrCost), C.POINTER(abi.DraceState)]
    L.or_drace_env_step.argtypes = [I, C.POINTER(abi.DraceEnv), I, F, C.POINTER(abi.DraceState), F, C.c_uint64, U, U, U, U, I]
    L.or_bmdrace_step.argtypes = [I, C.POINTER(abi.DraceEnv), C.POINTER(abi.SrCost), C.POINTER(abi.DraceState), C.POINTER(abi.BmInstance),
                                  F, U, C.c_float, F, F, I, F, D, C.POINTER(I), C.POINTER(I), D, D, D, D, D, D, C.POINTER(I),
                                  C.POINTER(I), D]
    L.or_count_drace_ops.restype = C.c_longlong
    L.or_count_drace_ops.argtypes = [C.POINTER(abi.DraceEnv), C.POINTER(abi.SrCost), C.POINTER(abi.DraceState), F, I,
                                     C.POINTER(abi.BmInstance), F, C.POINTER(C.c_longlong)]
    _libs[path] = L
    if path == LIB_PATH:
        _lib = L
    return L


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    def __init__(self, abi, path=None):
        self.abi = abi
        self.L = lib(abi, path)

    def philox(self, ctr, key):
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        o = (C.c_uint32 * 4)()
        self.L.or_philox4x32_10(c, k, o)
        return list(o)

    def normals4(self, key, ctr):
        o = np.empty(4, dtype=np.float32)
        self.L.or_normals4(key[0], key[1], ctr[0], ctr[1], ctr[2], ctr[3], _f(o))
        return o

    def log32(self, x):
        return self.L.or_log32(x)

    def exp32(self, x):
        return self.L.or_exp32(x)

    def inv_sqrt32(self, x):
        return self.L.or_inv_sqrt32(x)

    def sincos(self, k):
        s, c = C.c_float(), C.c_float()
        self.L.or_sincos_turns(k, C.byref(s), C.byref(c))
        return s.value, c.value

    def quad_step(self, f64, params, x, u, track=None, margin=0.95, gate=0, laps=0):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        u = np.ascontiguousarray(u, dtype=np.float64)
        g, l = C.c_int(gate), C.c_int(laps)
        self.L.or_quad_step(int(f64), C.byref(params), C.byref(track) if track is not None else None, margin, _d(x),
                            C.byref(g), C.byref(l), _d(u))
        return x, g.value, l.value

    def progress(self, f64, track, x, gate, laps):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return self.L.or_progress(int(f64), C.byref(track), _d(x), gate, laps)

    def running_cost(self, f64, params, cost, track, x, gate, laps, opponents, tau):
        x = np.ascontiguousarray(x, dtype=np.float64)
        opp = (self.abi.Opponent * max(1, len(opponents)))(*opponents)
        term = C.c_int()
        c = self.L.or_running_cost(int(f64), C.byref(params), C.byref(cost), C.byref(track), _d(x), gate, laps, opp,
                                   len(opponents), tau, C.byref(term))
        return c, term.value

    def rollout(self, f64, params, cost, track, ego, opponents, nominal, K, T, seed, step, sigma, want_noise=True,
                threads=8):
        nominal = np.ascontiguousarray(nominal, dtype=np.float32).reshape(-1)
        sig = np.ascontiguousarray(sigma, dtype=np.float32)
        costs = np.empty(K, dtype=np.float64)
        noise = np.empty(T * K * 4, dtype=np.float32) if want_noise else None
        traj = C.c_double()
        opp = (self.abi.Opponent * max(1, len(opponents)))(*opponents)
        rc = self.L.or_rollout(int(f64), C.byref(params), C.byref(cost), C.byref(track), C.byref(ego), opp,
                               len(opponents), _f(nominal), K, T, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF, step,
                               _f(sig), _d(costs), _f(noise) if noise is not None else None, C.byref(traj), threads)
        assert rc == 0
        return costs, (noise.reshape(K, T, 4) if noise is not None else None), traj.value

    def update(self, f64, params, costs, noise, nominal, lam, fill):
        K = costs.size
        T = nominal.size // 4  # noise layout [k][t][4]
        costs = np.ascontiguousarray(costs, dtype=np.float64)
        noise = np.ascontiguousarray(noise, dtype=np.float32).reshape(-1)
        nominal = np.ascontiguousarray(nominal, dtype=np.float32).reshape(-1)
        fill = np.ascontiguousarray(fill, dtype=np.float32)
        upd = np.empty(T * 4)
        act = np.empty(4)
        sh = np.empty(T * 4)
        w = np.empty(K)
        mn, nu = C.c_double(), C.c_double()
        rc = self.L.or_mppi_update(int(f64), C.byref(params), _d(costs), _f(noise), _f(nominal), K, T, lam, _f(fill),
                                   _d(upd), _d(act), _d(sh), C.byref(mn), C.byref(nu), _d(w))
        assert rc == 0
        return dict(updated=upd.reshape(T, 4), action=act, shifted=sh.reshape(T, 4), min_cost=mn.value, nu=nu.value,
                    weights=w)

    def cost_grad(self, f64, params, cost, track, ego, opponents, U, ck=10, grad=True, threads=8):
        """J, dJ/dU and the discrete-decision signature of each sequence U[n][T][4]."""
        U = np.ascontiguousarray(U, dtype=np.float32)
        n, T = U.shape[0], U.shape[1]
        J = np.empty(n)
        G = np.empty((n, T, 4)) if grad else None
        sig = np.empty(n, dtype=np.uint64)
        opp = (self.abi.Opponent * max(1, len(opponents)))(*opponents)
        rc = self.L.or_cost_grad(int(f64), C.byref(params), C.byref(cost), C.byref(track), C.byref(ego), opp,
                                 len(opponents), _f(U), n, T, ck, _d(J), _d(G) if grad else None,
                                 sig.ctypes.data_as(C.POINTER(C.c_uint64)), threads)
        assert rc == 0
        return J, G, sig

    def world_run(self, sc, drones, nominal, steps, lambdas, n_steps, threads=1, refine=(0, 0.05, 10)):
        """Closed-loop FP32 run; drones/nominal/steps/lambdas are updated in place.  refine =
        (gradient steps, step size, checkpoint interval) of the adjoint refinement (0: plain MPPI)."""
        abi = self.abi
        n_t, n_i, n_d = len(sc.tracks), len(sc.instances), len(sc.drone_world)
        tracks = (abi.Track * n_t)(*sc.tracks)
        inst = (abi.Instance * n_i)(*sc.instances)
        world = (C.c_int32 * n_d)(*sc.drone_world)
        st = (C.c_uint32 * n_i)(*steps)
        lam = np.ascontiguousarray(lambdas, dtype=np.float32)
        acts = np.empty(n_i * 4, dtype=np.float32)
        traj = np.empty(n_i, dtype=np.float32)
        rc = self.L.or_world_run(C.byref(sc.params), C.byref(sc.cost), tracks, inst, n_i, world, drones, n_d,
                                 _f(nominal), st, _f(lam), n_steps, threads, _f(acts), _f(traj), int(refine[0]),
                                 float(refine[1]), int(refine[2]))
        assert rc == 0
        steps[:] = list(st)
        lambdas[:] = lam
        return acts.reshape(n_i, 4), traj

    # ------------------------------------------------------------ PR-MPPI

    def pr_step(self, f64, params, cost, track, ego, opponents, ins, belief, step, lam, nom, rob, threads=8,
                want_noise=True, want_pairs=True):
        """One PR-MPPI step of one controller from the shifted plans nom, rob (T x 4)."""
        K, T = ins.n_samples, ins.n_timesteps
        M = ins.n_models if ins.sampled else ins.n_branches
        nom = np.ascontiguousarray(nom, dtype=np.float32).reshape(-1)
        rob = np.ascontiguousarray(rob, dtype=np.float32).reshape(-1)
        bel = np.zeros(self.abi.MAX_BRANCHES, dtype=np.float32)
        bel[:len(belief)] = belief
        lam = np.ascontiguousarray(lam, dtype=np.float32)
        noise = np.empty(K * T * 4, dtype=np.float32) if want_noise else None
        pairs = np.empty(2 * M * K * 2) if want_pairs else None
        agg = np.empty(K * 4)
        models = (C.c_int * M)()
        cand = np.empty(3 * T * 4)
        res = np.empty(6)
        opp = (self.abi.Opponent * max(1, len(opponents)))(*opponents)
        rc = self.L.or_pr_step(int(f64), C.byref(params), C.byref(cost), C.byref(track), C.byref(ego), opp, len(opponents),
                               C.byref(ins), _f(bel), step, _f(lam), _f(nom), _f(rob), threads,
                               _f(noise) if noise is not None else None, _d(pairs) if pairs is not None else None, _d(agg),
                               models, _d(cand), _d(res))
        assert rc == 0
        return dict(noise=noise.reshape(K, T, 4) if noise is not None else None,
                    pairs=pairs.reshape(2, M, K, 2) if pairs is not None else None, agg=agg.reshape(K, 4),
                    models=np.array(models[:], dtype=np.int32), cand=cand.reshape(3, T, 4), min=res[:3], nu=res[3:6])

    def pr_choose(self, f64, params, cost, track, ego, opponents, ins, belief, models, cand, lam, nu):
        """Candidate costs and choice for candidates cand[3][T][4] (cand1, cand2, new robust plan)."""
        M = ins.n_models if ins.sampled else ins.n_branches
        bel = np.zeros(self.abi.MAX_BRANCHES, dtype=np.float32)
        bel[:len(belief)] = belief
        mdl = (C.c_int * M)(*[int(x) for x in models])
        cand = np.ascontiguousarray(cand, dtype=np.float64).reshape(-1)
        lam = np.ascontiguousarray(lam, dtype=np.float32)
        nu = np.ascontiguousarray(nu, dtype=np.float64)
        cc = np.empty(2 * M)
        cs = np.empty(2 * M)
        res = np.empty(11)
        opp = (self.abi.Opponent * max(1, len(opponents)))(*opponents)
        rc = self.L.or_pr_choose(int(f64), C.byref(params), C.byref(cost), C.byref(track), C.byref(ego), opp,
                                 len(opponents), C.byref(ins), _f(bel), mdl, _d(cand), _f(lam), _d(nu), _d(cc), _d(cs),
                                 _d(res))
        assert rc == 0
        return dict(cand_cost=res[0:2], use_nom=int(res[2]), reset_nom=int(res[3]), action=res[4:8], lam=res[8:11],
                    cc=cc.reshape(2, M), cs=cs.reshape(2, M))

    def world_opponents(self, sc, drones, d):
        """Opponents of drone d: the other drones of its world in increasing index."""
        out = []
        for e in range(len(sc.drone_world)):
            if e != d and sc.drone_world[e] == sc.drone_world[d]:
                o = self.abi.Opponent()
                for c in range(3):
                    o.p[c] = drones[e].p[c]
                    o.v[c] = drones[e].v[c]
                out.append(o)
        return out

    def pr_world_step(self, sc, drones, works, steps, lambdas, threads=8):
        """One closed-loop FP32 PR-MPPI step of every controller of sc; drones, works (shifted
        (nom, rob) plans), steps and lambdas are updated in place.  Returns the actions and the
        per-instance step and choice records."""
        recs = []
        for i, ins in enumerate(sc.instances):
            belief = np.array(ins.belief[:ins.n_branches], dtype=np.float32)
            opps = self.world_opponents(sc, drones, ins.drone)
            r = self.pr_step(0, sc.params, sc.cost, sc.tracks[ins.track], drones[ins.drone], opps, ins, belief, steps[i],
                             lambdas[i], works[i][0], works[i][1], threads, want_noise=False, want_pairs=False)
            c = self.pr_choose(0, sc.params, sc.cost, sc.tracks[ins.track], drones[ins.drone], opps, ins, belief,
                               r["models"], r["cand"], lambdas[i], r["nu"])
            recs.append((r, c))
        acts = np.zeros((len(sc.instances), 4), dtype=np.float32)
        for i, ins in enumerate(sc.instances):
            r, c = recs[i]
            pick = 0 if c["cand_cost"][0] <= c["cand_cost"][1] else 1
            chosen = r["cand"][pick].astype(np.float32)
            robp = r["cand"][2].astype(np.float32)
            fill = np.array(ins.shift_fill[:], dtype=np.float32)
            works[i] = (np.vstack([chosen[1:], fill[None]]), np.vstack([robp[1:], fill[None]]))
            lambdas[i] = np.array(c["lam"], dtype=np.float32)
            steps[i] += 1
            acts[i] = np.array(c["action"], dtype=np.float32)
        controlled = {ins.drone: i for i, ins in enumerate(sc.instances)}
        for d in range(len(sc.drone_world)):
            s = drones[d]
            if d in controlled:
                ins = sc.instances[controlled[d]]
                x = np.array(list(s.p) + list(s.q) + list(s.v) + list(s.w), dtype=np.float64)
                x, g, l = self.quad_step(0, sc.params, x, acts[controlled[d]].astype(np.float64), sc.tracks[ins.track],
                                         sc.cost.gate_traversal_margin, s.gate, s.laps)
                for c in range(3):
                    s.p[c], s.v[c], s.w[c] = x[c], x[7 + c], x[10 + c]
                for c in range(4):
                    s.q[c] = x[3 + c]
                s.gate, s.laps = g, l
            else:
                raise NotImplementedError("pr_world_step: every drone must have a controller")
        return acts, recs

    # ------------------------------------------------------------ StratRace2D

    def sr_env_step(self, f64, env, theta, ego, state, seed, k, t, step, lose_on_opp_win=True):
        """One environment step (in place on a copy); returns (state, terminal code)."""
        st = self.abi.SrState.from_buffer_copy(state)
        u = np.ascontiguousarray(ego, dtype=np.float32)
        r = self.L.or_sr_env_step(int(f64), C.byref(env), theta, _f(u), C.byref(st), seed, k, t, step,
                                  1 if lose_on_opp_win else 0)
        return st, r

    def sr_project(self, f64, env, x, y, prev_s):
        lat = C.c_double()
        s = self.L.or_sr_project(int(f64), C.byref(env), x, y, prev_s, C.byref(lat))
        return s, lat.value

    def sr_opp_action(self, f64, env, theta, state):
        o = np.empty(2)
        self.L.or_sr_opp_action(int(f64), C.byref(env), theta, C.byref(state), _d(o))
        return o

    def sr_state_cost(self, f64, env, cost, state):
        return self.L.or_sr_state_cost(int(f64), C.byref(env), C.byref(cost), C.byref(state))

    def sr_rollout(self, f64, env, cost, state, theta, nominal, K, T, seed, step, sigma, threads=8, want_noise=True):
        nominal = np.ascontiguousarray(nominal, dtype=np.float32).reshape(-1)
        costs = np.empty(K)
        noise = np.empty(K * T * 2, dtype=np.float32) if want_noise else None
        rc = self.L.or_sr_rollout(int(f64), C.byref(env), C.byref(cost), C.byref(state), theta, _f(nominal), K, T, seed, step,
                                  sigma, threads, _d(costs), _f(noise) if noise is not None else None)
        assert rc == 0
        return costs, (noise.reshape(K, T, 2) if noise is not None else None)

    def sr_update(self, f64, costs, noise, nominal, lam, max_accel, fill=(0.0, 0.0)):
        K = costs.size
        T = nominal.size // 2
        costs = np.ascontiguousarray(costs, dtype=np.float64)
        noise = np.ascontiguousarray(noise, dtype=np.float32).reshape(-1)
        nominal = np.ascontiguousarray(nominal, dtype=np.float32).reshape(-1)
        fill = np.ascontiguousarray(fill, dtype=np.float32)
        upd, act, sh = np.empty(T * 2), np.empty(2), np.empty(T * 2)
        mn, nu = C.c_double(), C.c_double()
        rc = self.L.or_sr_update(int(f64), _d(costs), _f(noise), _f(nominal), K, T, lam, max_accel, _f(fill), _d(upd), _d(act),
                                 _d(sh), C.byref(mn), C.byref(nu))
        assert rc == 0
        return dict(updated=upd.reshape(T, 2), action=act, shifted=sh.reshape(T, 2), min_cost=mn.value, nu=nu.value)

    def prsr_step(self, f64, env, cost, state, ins, belief, step, lam, nom, rob, threads=8, want_noise=True, want_pairs=True):
        """One PR-MPPI step of a StratRace2D controller from the shifted plans nom, rob (T x 2)."""
        K, T = ins.n_samples, ins.n_timesteps
        M = ins.n_draws if ins.sampled else env.n_models
        nom = np.ascontiguousarray(nom, dtype=np.float32).reshape(-1)
        rob = np.ascontiguousarray(rob, dtype=np.float32).reshape(-1)
        bel = np.zeros(self.abi.SR_MAX_MODELS, dtype=np.float32)
        bel[:len(belief)] = belief
        lam = np.ascontiguousarray(lam, dtype=np.float32)
        noise = np.empty(K * T * 2, dtype=np.float32) if want_noise else None
        pairs = np.empty(2 * M * K * 2) if want_pairs else None
        agg = np.empty(K * 4)
        models = (C.c_int * M)()
        cand = np.empty(3 * T * 2)
        res = np.empty(6)
        rc = self.L.or_prsr_step(int(f64), C.byref(env), C.byref(cost), C.byref(state), C.byref(ins), _f(bel), step, _f(lam),
                                 _f(nom), _f(rob), threads, _f(noise) if noise is not None else None,
                                 _d(pairs) if pairs is not None else None, _d(agg), models, _d(cand), _d(res))
        assert rc == 0
        return dict(noise=noise.reshape(K, T, 2) if noise is not None else None,
                    pairs=pairs.reshape(2, M, K, 2) if pairs is not None else None, agg=agg.reshape(K, 4),
                    models=np.array(models[:], dtype=np.int32), cand=cand.reshape(3, T, 2), min=res[:3], nu=res[3:6])

    def prsr_choose(self, f64, env, cost, state, ins, belief, models, cand, step, lam, nu):
        M = ins.n_draws if ins.sampled else env.n_models
        bel = np.zeros(self.abi.SR_MAX_MODELS, dtype=np.float32)
        bel[:len(belief)] = belief
        mdl = (C.c_int * M)(*[int(x) for x in models])
        cand = np.ascontiguousarray(cand, dtype=np.float64).reshape(-1)
        lam = np.ascontiguousarray(lam, dtype=np.float32)
        nu = np.ascontiguousarray(nu, dtype=np.float64)
        cc, cs, res = np.empty(2 * M), np.empty(2 * M), np.empty(9)
        rc = self.L.or_prsr_choose(int(f64), C.byref(env), C.byref(cost), C.byref(state), C.byref(ins), _f(bel), mdl, _d(cand), step,
                                   _f(lam), _d(nu), _d(cc), _d(cs), _d(res))
        assert rc == 0
        return dict(cand_cost=res[0:2], use_nom=int(res[2]), reset_nom=int(res[3]), action=res[4:6], lam=res[6:9],
                    cc=cc.reshape(2, M), cs=cs.reshape(2, M))

    # ---------------------------------------------------------- branching spline MPPI

    def bm_basis(self, f64, knots, T):
        """Natural cubic spline basis [T][M] of the integer knots."""
        tau = (C.c_int * len(knots))(*[int(x) for x in knots])
        B = np.empty(T * len(knots))
        assert self.L.or_bm_basis(int(f64), tau, len(knots), T, _d(B)) == 0
        return B.reshape(T, len(knots))

    def sr_update_belief(self, f64, env, belief, obs, noms):
        """Belief after one observed opponent action, given each model's nominal action noms [n_models][2]."""
        b = np.array(belief, dtype=np.float64)
        o = np.ascontiguousarray(obs, dtype=np.float64)
        n = np.ascontiguousarray(noms, dtype=np.float64).reshape(-1)
        assert self.L.or_sr_update_belief(int(f64), C.byref(env), _d(b), _d(o), _d(n)) == 0
        return b

    def sr_env_step_belief(self, f64, env, theta, ego, state, belief, seed, k, t, step, lose_on_opp_win=True, stream=4):
        """One environment step with belief update; returns (state, belief, terminal code)."""
        st = self.abi.SrState.from_buffer_copy(state)
        u = np.ascontiguousarray(ego, dtype=np.float32)
        b = np.zeros(self.abi.SR_MAX_MODELS, dtype=np.float32)
        b[:len(belief)] = belief
        r = self.L.or_sr_env_step_belief(int(f64), C.byref(env), theta, _f(u), C.byref(st), _f(b), seed, k, t, step,
                                         1 if lose_on_opp_win else 0, stream)
        return st, b[:env.n_models].copy(), r

    def bm_step(self, f64, env, cost, state, ins, belief, step, lam, knots, threads=8, want_noise=True, prev_dense=None):
        """One branching spline MPPI step of a StratRace2D or HiddenObs controller from its belief and knots
        [n_models + 1][M][2] (and, with verification, the previous step's executed plans [n_models + 1][T][2])."""
        ho = isinstance(env, self.abi.HoEnv)
        drc = isinstance(env, self.abi.DraceEnv)
        K, T, M, NT = ins.n_samples, ins.n_timesteps, ins.n_knots, 2 if ho else env.n_models
        NB = NT + 1
        kn = np.ascontiguousarray(knots, dtype=np.float32).reshape(-1)
        assert kn.size == NB * M * 2
        prev = None if prev_dense is None else np.ascontiguousarray(prev_dense, dtype=np.float32).reshape(-1)
        dnew = np.empty(NB * T * 2)
        fail = np.zeros(2 * T, dtype=np.int32)
        bel = np.zeros(self.abi.SR_MAX_MODELS, dtype=np.float32)
        bel[:NT] = belief[:NT]
        noise = np.empty(NB * M * K * 2, dtype=np.float32) if want_noise else None
        ct = np.empty(NT * K)
        used = np.empty(NT * K, dtype=np.int32)
        bt = np.empty(NT * K, dtype=np.int32)
        lens = np.empty(NT * K, dtype=np.int32)
        costs = np.empty(NB * K)
        mins = np.empty(NB * M)
        nu = np.empty(NB)
        kout = np.empty(NB * M * 2)
        dense = np.empty(NB * T * 2)
        res = np.empty(4)
        ip = C.POINTER(C.c_int)
        rc = (self.L.or_bmho_step if ho else self.L.or_bmdrace_step if drc else self.L.or_bm_step)(int(f64), C.byref(env), C.byref(cost), C.byref(state), C.byref(ins), _f(bel), step, lam, _f(kn),
                               _f(prev) if prev is not None else None, threads,
                               _f(noise) if noise is not None else None, _d(ct), used.ctypes.data_as(ip), bt.ctypes.data_as(ip),
                               _d(costs), _d(mins), _d(nu), _d(kout), _d(dense), _d(dnew), fail.ctypes.data_as(ip),
                               lens.ctypes.data_as(ip), _d(res))
        assert rc == 0
        return dict(noise=noise.reshape(NB, M, K, 2) if noise is not None else None, ctrue=ct.reshape(NT, K),
                    used=used.reshape(NT, K), btime=bt.reshape(NT, K), len=lens.reshape(NT, K), costs=costs.reshape(NB, K), mins=mins.reshape(NB, M),
                    nu=nu, knots=kout.reshape(NB, M, 2), dense=dense.reshape(NB, T, 2), action=res[:2], lambda_=float(res[2]),
                    branch=int(res[3]), dense_new=dnew.reshape(NB, T, 2), fail_new=fail[:T].copy(), fail_old=fail[T:].copy())

    # ---------------------------------------------------------- session display predictions

    def sr_predict(self, env, state, theta, plans, T, branching, belief=None, min_conf=0.0):
        """SimulationEngine::sendState's noise-free prediction: (full_pos [T][4], actions [T][2],
        (initial branch, branching time, final branch, stop reason, stop time))."""
        pl = np.ascontiguousarray(plans, dtype=np.float32).reshape(-1)
        b = np.zeros(self.abi.SR_MAX_MODELS, dtype=np.float32)
        if belief is not None:
            b[:len(belief)] = belief
        pos = np.empty(T * 4, dtype=np.float32)
        act = np.empty(T * 2, dtype=np.float32)
        info = (C.c_int * 5)()
        rc = self.L.or_sr_predict(C.byref(env), C.byref(state), theta, _f(pl), T, 1 if branching else 0, _f(b), min_conf, _f(pos),
                                  _f(act), info)
        assert rc == 0
        return pos.reshape(T, 4), act.reshape(T, 2), tuple(info[:])

    def sr_opp_target(self, f64, env, theta, state):
        return self.L.or_sr_opp_target(int(f64), C.byref(env), theta, C.byref(state))

    def sr_opp_target_goal(self, f64, env, theta, state):
        """(target lateral offset, speed goal) of computeOppTargetLatDist."""
        goal = C.c_double()
        lat = self.L.or_sr_opp_target_goal(int(f64), C.byref(env), theta, C.byref(state), C.byref(goal))
        return lat, goal.value

    # ---------------------------------------------------------- HiddenObs

    HO_PRIMS = {"sq_rect": 0, "sq_half_plane": 1, "annulus": 2, "seg_rect": 3, "seg_annulus": 4}

    def ho_prim(self, f64, name, *args):
        """A HiddenObs geometry primitive (sq_rect, sq_half_plane, annulus, seg_rect, seg_annulus)."""
        a = np.zeros(8)
        a[:len(args)] = args
        return self.L.or_ho_prim(int(f64), self.HO_PRIMS[name], _d(a))

    def ho_eval(self, f64, what, env, state, theta=0, cost=None):
        """bdist, can_see, advance, state_cost or final_cost of a HiddenObs state."""
        w = {"bdist": 0, "can_see": 1, "advance": 2, "state_cost": 3, "final_cost": 4}[what]
        return self.L.or_ho_eval(int(f64), w, C.byref(env), C.byref(cost) if cost is not None else None, C.byref(state), theta)

    def ho_env_step(self, f64, env, theta, ego, state, belief=None, seed=0, k=0, t=0, step=0, stream=4):
        """One HiddenObs environment step; returns (state, belief or None, terminal code)."""
        st = self.abi.HoState.from_buffer_copy(state)
        u = np.ascontiguousarray(ego, dtype=np.float32)
        b = None if belief is None else np.ascontiguousarray(belief, dtype=np.float32).copy()
        r = self.L.or_ho_env_step(int(f64), C.byref(env), theta, _f(u), C.byref(st), _f(b) if b is not None else None, seed, k, t, step,
                                  stream)
        return st, b, r

    def count_ho(self, env, cost, state, ins, knots, belief, n):
        """Counted FP32 work of n branching HiddenObs rollouts: (ops, rollout-steps)."""
        kn = np.ascontiguousarray(knots, dtype=np.float32).reshape(-1)
        b = np.zeros(self.abi.SR_MAX_MODELS, dtype=np.float32)
        b[:2] = belief[:2]
        steps = C.c_longlong(0)
        ops = self.L.or_count_ho_ops(C.byref(env), C.byref(cost), C.byref(state), _f(kn), n, C.byref(ins), _f(b), C.byref(steps))
        return ops, steps.value

    # ---------------------------------------------------------- DroneRace

    def drace_project(self, f64, line, x, y, prev_s=-1.0):
        """Progress in [0, 1) of (x, y) on one raceline [512][2] (full scan when prev_s < 0)."""
        tp = np.ascontiguousarray(line, dtype=np.float32).reshape(-1)
        return self.L.or_drace_project(int(f64), _f(tp), x, y, prev_s)

    def drace_pid(self, f64, env, state, agent, model):
        out = np.empty(2)
        assert self.L.or_drace_pid(int(f64), C.byref(env), C.byref(state), agent, model, _d(out)) == 0
        return out

    DRACE_EVAL = {"bdist": 0, "bdist_lose": 1, "winner": 2, "opp_winner": 3, "advance": 4, "opp_advance": 5, "state_cost": 6,
                  "final_cost": 7}

    def drace_eval(self, f64, what, env, state, cost=None):
        return self.L.or_drace_eval(int(f64), self.DRACE_EVAL[what], C.byref(env), C.byref(cost) if cost is not None else None,
                                    C.byref(state))

    def drace_env_step(self, f64, env, theta, ego, state, belief=None, seed=0, k=0, t=0, step=0, stream=4, lose_on_opp_win=True):
        """One DroneRace environment step; returns (state, belief or None, terminal code)."""
        st = self.abi.DraceState.from_buffer_copy(state)
        u = np.ascontiguousarray(ego, dtype=np.float32)
        b = None if belief is None else np.ascontiguousarray(belief, dtype=np.float32).copy()
        r = self.L.or_drace_env_step(int(f64), C.byref(env), theta, _f(u), C.byref(st), _f(b) if b is not None else None, seed, k, t,
                                     step, stream, 1 if lose_on_opp_win else 0)
        return st, b, r

    def count_drace(self, env, cost, state, ins, knots, belief, n):
        """Counted FP32 work of n branching DroneRace rollouts: (ops, rollout-steps)."""
        kn = np.ascontiguousarray(knots, dtype=np.float32).reshape(-1)
        b = np.zeros(self.abi.SR_MAX_MODELS, dtype=np.float32)
        b[:env.n_models] = belief[:env.n_models]
        steps = C.c_longlong(0)
        ops = self.L.or_count_drace_ops(C.byref(env), C.byref(cost), C.byref(state), _f(kn), n, C.byref(ins), _f(b), C.byref(steps))
        return ops, steps.value
