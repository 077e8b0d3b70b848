"""Python handle over the C ABI (include/dr_api.h).

Mirrors the reference controller interface (include/controllers.h:12-111):
MPPI.step(state, opponents) plays Controller::getControl(state, outAction);
Engine drives a whole batch of controllers and their closed-loop worlds.
"""
import ctypes as C

import numpy as np

from . import abi
from .abi import check


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _autotune(fn, h, trials):
    """Runs one engine's launch-shape search; returns (chosen launch fields, ms per step)."""
    cfg = abi.LaunchCfg()
    ms = C.c_float()
    check(fn(h, int(trials), C.byref(cfg), C.byref(ms)))
    return {f: getattr(cfg, f) for f, _ in abi.LaunchCfg._fields_}, ms.value


class Engine:
    def __init__(self, scenario):
        lib = abi.load()
        self.lib = lib
        self.sc = scenario
        n_t = len(scenario.tracks)
        self._tracks = (abi.Track * n_t)(*scenario.tracks)
        n_i = len(scenario.instances)
        self._inst = (abi.Instance * n_i)(*scenario.instances)
        n_d = len(scenario.drone_world)
        self._world = (C.c_int32 * n_d)(*scenario.drone_world)
        self.n_inst, self.n_drones = n_i, n_d
        h = C.c_void_p()
        check(lib.dr_engine_create(C.byref(scenario.params), C.byref(scenario.cost), self._tracks, n_t, self._inst, n_i,
                                   self._world, n_d, C.byref(h)))
        self.h = h
        self.set_drones(scenario.drones)
        for i, ins in enumerate(scenario.instances):
            nom = np.tile(np.array(ins.shift_fill[:], dtype=np.float32), (ins.n_timesteps, 1))
            self.set_nominal(i, nom)

    def close(self):
        if self.h:
            self.lib.dr_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_drones(self, states):
        arr = (abi.DroneState * self.n_drones)(*states)
        check(self.lib.dr_engine_set_drones(self.h, arr, self.n_drones))

    def get_drones(self):
        arr = (abi.DroneState * self.n_drones)()
        check(self.lib.dr_engine_get_drones(self.h, arr, self.n_drones))
        return arr

    def set_nominal(self, i, nom):
        a = np.ascontiguousarray(nom, dtype=np.float32).reshape(-1)
        check(self.lib.dr_engine_set_nominal(self.h, i, _fptr(a), a.size))

    def get_nominal(self, i):
        T = self.sc.instances[i].n_timesteps
        a = np.empty(T * 4, dtype=np.float32)
        check(self.lib.dr_engine_get_nominal(self.h, i, _fptr(a), a.size))
        return a.reshape(T, 4)

    def set_step(self, i, step):
        check(self.lib.dr_engine_set_step(self.h, i, step))

    def step(self, n_steps=1, advance=False):
        a = np.empty(self.n_inst * 4, dtype=np.float32)
        check(self.lib.dr_engine_step(self.h, n_steps, 1 if advance else 0, _fptr(a)))
        return a.reshape(self.n_inst, 4)

    def step_async(self, n_steps=1, advance=False):
        check(self.lib.dr_engine_step_async(self.h, n_steps, 1 if advance else 0))

    def synchronize(self):
        check(self.lib.dr_engine_synchronize(self.h))

    def costs(self, i):
        K = self.sc.instances[i].n_samples
        a = np.empty(K, dtype=np.float32)
        check(self.lib.dr_engine_get_costs(self.h, i, _fptr(a), K))
        return a

    def noise(self, i):
        ins = self.sc.instances[i]
        a = np.empty(ins.n_timesteps * ins.n_samples * 4, dtype=np.float32)
        check(self.lib.dr_engine_get_noise(self.h, i, _fptr(a), a.size))
        return a.reshape(ins.n_samples, ins.n_timesteps, 4)  # [k][t][d]

    def actions(self):
        a = np.empty(self.n_inst * 4, dtype=np.float32)
        check(self.lib.dr_engine_get_actions(self.h, _fptr(a), a.size))
        return a.reshape(self.n_inst, 4)

    def diag(self):
        d = (abi.Diag * self.n_inst)()
        check(self.lib.dr_engine_get_diag(self.h, d, self.n_inst))
        return [dict(min_cost=x.min_cost, nu=x.nu, lam=x.lambda_, traj_cost=x.traj_cost, step=x.step, exec_cost=x.exec_cost,
                     exec_progress=x.exec_progress, exec_collisions=x.exec_collisions) for x in d]

    def set_launch(self, **kw):
        cfg = self.get_launch()
        for k, v in kw.items():
            setattr(cfg, k, v)
        check(self.lib.dr_engine_set_launch(self.h, C.byref(cfg)))

    def get_launch(self):
        cfg = abi.LaunchCfg()
        check(self.lib.dr_engine_get_launch(self.h, C.byref(cfg)))
        return cfg

    def autotune(self, trials=5):
        return _autotune(self.lib.dr_engine_autotune, self.h, trials)

    def set_eval(self, on):
        check(self.lib.dr_engine_set_eval(self.h, 1 if on else 0))

    def set_exec_cost(self, on):
        """Track the executed closed loop's running cost, collisions and progress (diag exec_*); on clears them."""
        check(self.lib.dr_engine_set_exec_cost(self.h, 1 if on else 0))

    def horizon(self, i):
        """T of instance i."""
        if not 0 <= i < self.n_inst:
            raise IndexError(f"instance {i} out of range [0, {self.n_inst})")
        return self.sc.instances[i].n_timesteps

    def cost_grad(self, i, U, ckpt_every=10, grad=True):
        """J and dJ/dU of control sequences U[n][T][4] of instance i (noise-free adjoint rollout)."""
        U = np.ascontiguousarray(U, dtype=np.float32)
        T = self.horizon(i)
        if U.ndim != 3 or U.shape[0] < 1 or U.shape[1:] != (T, 4):
            raise ValueError(f"U must have shape (n, {T}, 4) for instance {i}, got {U.shape}")
        n = U.shape[0]
        J = np.empty(n, dtype=np.float32)
        G = np.empty_like(U) if grad else None
        check(self.lib.dr_engine_cost_grad(self.h, i, _fptr(U), n, U.size, _fptr(J), _fptr(G) if grad else None, ckpt_every))
        return J, G

    def cost_grad_device(self, i, u_ptr, n_seq, n_floats, j_ptr, g_ptr, ckpt_every, stream):
        check(self.lib.dr_engine_cost_grad_device(self.h, i, u_ptr, n_seq, n_floats, j_ptr, g_ptr, ckpt_every, stream))

    def set_refine(self, n_grad_steps, step=0.05, ckpt_every=10):
        check(self.lib.dr_engine_set_refine(self.h, n_grad_steps, step, ckpt_every))

    def set_timing(self, on):
        check(self.lib.dr_engine_set_timing(self.h, 1 if on else 0))

    def last_timing(self):
        """Device ms of the last timed run, summed over its steps, per phase."""
        a = np.zeros(5, dtype=np.float32)
        check(self.lib.dr_engine_last_timing(self.h, _fptr(a), 5))
        return dict(zip(("rollout_ms", "reduce_ms", "update_ms", "refine_ms", "total_ms"), [float(x) for x in a]))


class MPPI:
    """Single controller with the reference's getControl shape (controllers.h:23-111)."""

    def __init__(self, params, cost, track, n_samples, n_timesteps, seed, lam=1.0, sigma=(2.0, 0.5, 0.5, 0.5),
                 fill=(0.0, 0.0, 0.0, 0.0), adapt=0, max_opponents=0):
        lib = abi.load()
        self.lib = lib
        self.T = n_timesteps
        self.max_opp = max_opponents
        cfg = abi.MPPICfg(n_samples=n_samples, n_timesteps=n_timesteps, lambda_=lam, sigma=(C.c_float * 4)(*sigma),
                          shift_fill=(C.c_float * 4)(*fill), adapt_lambda=adapt, max_opponents=max_opponents)
        self._track = track
        h = C.c_void_p()
        check(lib.dr_mppi_create(C.byref(params), C.byref(cost), C.byref(track), C.byref(cfg), seed, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.dr_mppi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, state, opponents=()):
        opp = (abi.Opponent * max(1, len(opponents)))(*opponents)
        out = np.empty(4, dtype=np.float32)
        check(self.lib.dr_mppi_step(self.h, C.byref(state), opp, len(opponents), _fptr(out)))
        return out

    def get_nominal(self):
        a = np.empty(self.T * 4, dtype=np.float32)
        check(self.lib.dr_mppi_get_nominal(self.h, _fptr(a), a.size))
        return a.reshape(self.T, 4)

    def set_nominal(self, nom):
        a = np.ascontiguousarray(nom, dtype=np.float32).reshape(-1)
        check(self.lib.dr_mppi_set_nominal(self.h, _fptr(a), a.size))

    def diag(self):
        d = abi.Diag()
        check(self.lib.dr_mppi_get_diag(self.h, C.byref(d)))
        return dict(min_cost=d.min_cost, nu=d.nu, lam=d.lambda_, traj_cost=d.traj_cost, step=d.step)


class PREngine:
    """A batch of PR-MPPI controllers (PRMPPIController, include/controllers.h:114-179).

    get_nominals() returns (h_nom_nominal, h_rob_nominal); diag() carries
    use_nom_plan / reset_nom / the three minima, nu and temperatures."""

    def __init__(self, scenario):
        lib = abi.load()
        self.lib = lib
        self.sc = scenario
        n_t = len(scenario.tracks)
        self._tracks = (abi.Track * n_t)(*scenario.tracks)
        n_i = len(scenario.instances)
        self._inst = (abi.PrInstance * n_i)(*scenario.instances)
        n_d = len(scenario.drone_world)
        self._world = (C.c_int32 * n_d)(*scenario.drone_world)
        self.n_inst, self.n_drones = n_i, n_d
        h = C.c_void_p()
        check(lib.dr_pr_create(C.byref(scenario.params), C.byref(scenario.cost), self._tracks, n_t, self._inst, n_i,
                               self._world, n_d, C.byref(h)))
        self.h = h
        self.set_drones(scenario.drones)
        for i, ins in enumerate(scenario.instances):
            nom = np.tile(np.array(ins.shift_fill[:], dtype=np.float32), (ins.n_timesteps, 1))
            self.set_nominals(i, nom, nom)

    def close(self):
        if self.h:
            self.lib.dr_pr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n_models(self, i):
        ins = self.sc.instances[i]
        return ins.n_models if ins.sampled else ins.n_branches

    def set_drones(self, states):
        arr = (abi.DroneState * self.n_drones)(*states)
        check(self.lib.dr_pr_set_drones(self.h, arr, self.n_drones))

    def get_drones(self):
        arr = (abi.DroneState * self.n_drones)()
        check(self.lib.dr_pr_get_drones(self.h, arr, self.n_drones))
        return arr

    def set_nominals(self, i, nom, rob):
        a = np.ascontiguousarray(nom, dtype=np.float32).reshape(-1)
        b = np.ascontiguousarray(rob, dtype=np.float32).reshape(-1)
        check(self.lib.dr_pr_set_nominals(self.h, i, _fptr(a), _fptr(b), a.size))

    def get_nominals(self, i):
        T = self.sc.instances[i].n_timesteps
        a = np.empty(T * 4, dtype=np.float32)
        b = np.empty(T * 4, dtype=np.float32)
        check(self.lib.dr_pr_get_nominals(self.h, i, _fptr(a), _fptr(b), a.size))
        return a.reshape(T, 4), b.reshape(T, 4)

    def set_belief(self, i, belief):
        a = np.ascontiguousarray(belief, dtype=np.float32)
        check(self.lib.dr_pr_set_belief(self.h, i, _fptr(a), a.size))

    def belief(self, i, n_branches):
        a = np.empty(n_branches, dtype=np.float32)
        check(self.lib.dr_pr_get_belief(self.h, i, _fptr(a), a.size))
        return a

    def step(self, n_steps=1, advance=False):
        a = np.empty(self.n_inst * 4, dtype=np.float32)
        check(self.lib.dr_pr_step(self.h, n_steps, 1 if advance else 0, _fptr(a)))
        return a.reshape(self.n_inst, 4)

    def step_async(self, n_steps=1, advance=False):
        check(self.lib.dr_pr_step_async(self.h, n_steps, 1 if advance else 0))

    def synchronize(self):
        check(self.lib.dr_pr_synchronize(self.h))

    def diag(self):
        arr = (abi.PrDiag * self.n_inst)()
        check(self.lib.dr_pr_get_diag(self.h, arr, self.n_inst))
        return arr

    def noise(self, i):
        ins = self.sc.instances[i]
        a = np.empty(ins.n_samples * ins.n_timesteps * 4, dtype=np.float32)
        check(self.lib.dr_pr_get_noise(self.h, i, _fptr(a), a.size))
        return a.reshape(ins.n_samples, ins.n_timesteps, 4)

    def pair_costs(self, i):
        """[plan][m][k][cost, max safety] of the last step."""
        ins = self.sc.instances[i]
        M = self.n_models(i)
        a = np.empty(2 * M * ins.n_samples * 2, dtype=np.float32)
        check(self.lib.dr_pr_get_pair_costs(self.h, i, _fptr(a), a.size))
        return a.reshape(2, M, ins.n_samples, 2)

    def aggregates(self, i):
        ins = self.sc.instances[i]
        a = np.empty(ins.n_samples * 4, dtype=np.float32)
        check(self.lib.dr_pr_get_aggregates(self.h, i, _fptr(a), a.size))
        return a.reshape(ins.n_samples, 4)

    def candidates(self, i):
        T = self.sc.instances[i].n_timesteps
        a = np.empty(T * 4, dtype=np.float32)
        b = np.empty(T * 4, dtype=np.float32)
        check(self.lib.dr_pr_get_candidates(self.h, i, _fptr(a), _fptr(b), a.size))
        return a.reshape(T, 4), b.reshape(T, 4)

    def models(self, i):
        M = self.n_models(i)
        a = (C.c_int32 * M)()
        check(self.lib.dr_pr_get_models(self.h, i, a, M))
        return np.array(a[:], dtype=np.int32)

    def set_launch(self, **kw):
        cfg = self.get_launch()
        for k, v in kw.items():
            setattr(cfg, k, int(v))
        check(self.lib.dr_pr_set_launch(self.h, C.byref(cfg)))

    def get_launch(self):
        cfg = abi.LaunchCfg()
        check(self.lib.dr_pr_get_launch(self.h, C.byref(cfg)))
        return cfg

    def autotune(self, trials=5):
        """Times the rollout/reduction launch shapes on this workload and keeps the fastest; state is restored."""
        return _autotune(self.lib.dr_pr_autotune, self.h, trials)

    def set_timing(self, on=True):
        check(self.lib.dr_pr_set_timing(self.h, 1 if on else 0))

    def last_timing(self):
        """Device ms summed over the last timed run: rollout, aggregation, weighted averages, update, step."""
        a = np.empty(5, dtype=np.float32)
        check(self.lib.dr_pr_last_timing(self.h, _fptr(a), 5))
        return a


class SREngine:
    """A batch of plain-MPPI controllers on the 2-D race environment, one world each."""

    def __init__(self, scenario):
        lib = abi.load()
        self.lib = lib
        self.sc = scenario
        n = len(scenario.instances)
        self._inst = (abi.SrInstance * n)(*scenario.instances)
        self.n_inst = n
        h = C.c_void_p()
        check(lib.dr_sr_create(C.byref(scenario.env), C.byref(scenario.cost), self._inst, n, C.byref(h)))
        self.h = h
        self.set_states(scenario.states)
        for i, ins in enumerate(scenario.instances):
            self.set_nominal(i, np.zeros((ins.n_timesteps, 2), dtype=np.float32))

    def close(self):
        if self.h:
            self.lib.dr_sr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_states(self, states):
        arr = (abi.SrState * self.n_inst)(*states)
        check(self.lib.dr_sr_set_states(self.h, arr, self.n_inst))

    def get_states(self):
        arr = (abi.SrState * self.n_inst)()
        check(self.lib.dr_sr_get_states(self.h, arr, self.n_inst))
        return arr

    def set_nominal(self, i, nom):
        a = np.ascontiguousarray(nom, dtype=np.float32).reshape(-1)
        check(self.lib.dr_sr_set_nominal(self.h, i, _fptr(a), a.size))

    def get_nominal(self, i):
        T = self.sc.instances[i].n_timesteps
        a = np.empty(T * 2, dtype=np.float32)
        check(self.lib.dr_sr_get_nominal(self.h, i, _fptr(a), a.size))
        return a.reshape(T, 2)

    def step(self, n_steps=1, advance=False):
        a = np.empty(self.n_inst * 2, dtype=np.float32)
        check(self.lib.dr_sr_step(self.h, n_steps, 1 if advance else 0, _fptr(a)))
        return a.reshape(self.n_inst, 2)

    def events(self):
        a = (C.c_int32 * self.n_inst)()
        check(self.lib.dr_sr_get_events(self.h, a, self.n_inst))
        return np.array(a[:], dtype=np.int32)

    def costs(self, i):
        K = self.sc.instances[i].n_samples
        a = np.empty(K, dtype=np.float32)
        check(self.lib.dr_sr_get_costs(self.h, i, _fptr(a), K))
        return a

    def noise(self, i):
        ins = self.sc.instances[i]
        a = np.empty(ins.n_samples * ins.n_timesteps * 2, dtype=np.float32)
        check(self.lib.dr_sr_get_noise(self.h, i, _fptr(a), a.size))
        return a.reshape(ins.n_samples, ins.n_timesteps, 2)

    def diag(self):
        arr = (abi.Diag * self.n_inst)()
        check(self.lib.dr_sr_get_diag(self.h, arr, self.n_inst))
        return [dict(min_cost=d.min_cost, nu=d.nu, lambda_=d.lambda_, step=d.step) for d in arr]

    def set_launch(self, **kw):
        cfg = abi.LaunchCfg()
        check(self.lib.dr_sr_get_launch(self.h, C.byref(cfg)))
        for k, v in kw.items():
            setattr(cfg, k, int(v))
        check(self.lib.dr_sr_set_launch(self.h, C.byref(cfg)))

    def get_launch(self):
        cfg = abi.LaunchCfg()
        check(self.lib.dr_sr_get_launch(self.h, C.byref(cfg)))
        return cfg

    def autotune(self, trials=5):
        """Times the rollout/reduction launch shapes on this workload and keeps the fastest; state is restored."""
        return _autotune(self.lib.dr_sr_autotune, self.h, trials)

    def set_timing(self, on=True):
        check(self.lib.dr_sr_set_timing(self.h, 1 if on else 0))

    def last_timing(self):
        """Device ms summed over the last timed run: rollout, weighted average, update + world, step."""
        a = np.empty(4, dtype=np.float32)
        check(self.lib.dr_sr_last_timing(self.h, _fptr(a), 4))
        return a


class PRSREngine(PREngine):
    """PR-MPPI controllers on StratRace2D worlds (the reference's PR-MPPI on its own environment):
    plans are T x 2, one world state (abi.SrState) per instance."""

    D = 2

    def __init__(self, scenario):
        lib = abi.load()
        self.lib = lib
        self.sc = scenario
        n = len(scenario.instances)
        self._inst = (abi.PrsrInstance * n)(*scenario.instances)
        self.n_inst = n
        h = C.c_void_p()
        check(lib.dr_prsr_create(C.byref(scenario.env), C.byref(scenario.cost), self._inst, n, C.byref(h)))
        self.h = h
        self.set_states(scenario.states)

    def n_models(self, i):
        ins = self.sc.instances[i]
        return ins.n_draws if ins.sampled else self.sc.env.n_models

    def set_states(self, states):
        arr = (abi.SrState * self.n_inst)(*states)
        check(self.lib.dr_pr_set_sr_states(self.h, arr, self.n_inst))

    def get_states(self):
        arr = (abi.SrState * self.n_inst)()
        check(self.lib.dr_pr_get_sr_states(self.h, arr, self.n_inst))
        return arr

    def events(self):
        a = (C.c_int32 * self.n_inst)()
        check(self.lib.dr_pr_get_events(self.h, a, self.n_inst))
        return np.array(a[:], dtype=np.int32)

    def get_nominals(self, i):
        T = self.sc.instances[i].n_timesteps
        a = np.empty(T * 2, dtype=np.float32)
        b = np.empty(T * 2, dtype=np.float32)
        check(self.lib.dr_pr_get_nominals(self.h, i, _fptr(a), _fptr(b), a.size))
        return a.reshape(T, 2), b.reshape(T, 2)

    def step(self, n_steps=1, advance=False):
        a = np.empty(self.n_inst * 2, dtype=np.float32)
        check(self.lib.dr_pr_step(self.h, n_steps, 1 if advance else 0, _fptr(a)))
        return a.reshape(self.n_inst, 2)

    def noise(self, i):
        ins = self.sc.instances[i]
        a = np.empty(ins.n_samples * ins.n_timesteps * 2, dtype=np.float32)
        check(self.lib.dr_pr_get_noise(self.h, i, _fptr(a), a.size))
        return a.reshape(ins.n_samples, ins.n_timesteps, 2)

    def candidates(self, i):
        T = self.sc.instances[i].n_timesteps
        a = np.empty(T * 2, dtype=np.float32)
        b = np.empty(T * 2, dtype=np.float32)
        check(self.lib.dr_pr_get_candidates(self.h, i, _fptr(a), _fptr(b), a.size))
        return a.reshape(T, 2), b.reshape(T, 2)


class BMEngine:
    """Branching spline MPPI controllers on StratRace2D worlds (the reference's default
    MPPIController with splines, without plan verification), one world each."""

    _create, _set, _get, _state = "dr_bm_create", "dr_bm_set_states", "dr_bm_get_states", abi.SrState

    def __init__(self, scenario):
        lib = abi.load()
        self.lib = lib
        self.sc = scenario
        n = len(scenario.instances)
        self._inst = (abi.BmInstance * n)(*scenario.instances)
        self.n_inst = n
        self.n_models = self._n_models(scenario.env)
        self.recording = False
        h = C.c_void_p()
        check(getattr(lib, self._create)(C.byref(scenario.env), C.byref(scenario.cost), self._inst, n, C.byref(h)))
        self.h = h
        self.set_states(scenario.states)

    @staticmethod
    def _n_models(env):
        return env.n_models

    def close(self):
        if self.h:
            self.lib.dr_bm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _dims(self, i):
        ins = self.sc.instances[i]
        return ins.n_samples, ins.n_timesteps, ins.n_knots, self.n_models

    def set_states(self, states):
        arr = (self._state * self.n_inst)(*states)
        check(getattr(self.lib, self._set)(self.h, arr, self.n_inst))

    def get_states(self):
        arr = (self._state * self.n_inst)()
        check(getattr(self.lib, self._get)(self.h, arr, self.n_inst))
        return arr

    def set_belief(self, i, belief):
        b = np.ascontiguousarray(belief, dtype=np.float32)
        check(self.lib.dr_bm_set_belief(self.h, i, _fptr(b), b.size))

    def belief(self, i):
        b = np.empty(self.n_models, dtype=np.float32)
        check(self.lib.dr_bm_get_belief(self.h, i, _fptr(b), b.size))
        return b

    def set_knots(self, i, knots):
        a = np.ascontiguousarray(knots, dtype=np.float32).reshape(-1)
        check(self.lib.dr_bm_set_knots(self.h, i, _fptr(a), a.size))

    def knots(self, i):
        K, T, M, NT = self._dims(i)
        a = np.empty((NT + 1) * M * 2, dtype=np.float32)
        check(self.lib.dr_bm_get_knots(self.h, i, _fptr(a), a.size))
        return a.reshape(NT + 1, M, 2)

    def nominal(self, i):
        K, T, M, NT = self._dims(i)
        a = np.empty((NT + 1) * T * 2, dtype=np.float32)
        check(self.lib.dr_bm_get_nominal(self.h, i, _fptr(a), a.size))
        return a.reshape(NT + 1, T, 2)

    def basis(self, i):
        K, T, M, NT = self._dims(i)
        a = np.empty(T * M, dtype=np.float32)
        check(self.lib.dr_bm_get_basis(self.h, i, _fptr(a), a.size))
        return a.reshape(T, M)

    def step(self, n_steps=1, advance=False):
        a = np.empty(self.n_inst * 2, dtype=np.float32)
        check(self.lib.dr_bm_step(self.h, n_steps, 1 if advance else 0, _fptr(a)))
        return a.reshape(self.n_inst, 2)

    def events(self):
        a = (C.c_int32 * self.n_inst)()
        check(self.lib.dr_bm_get_events(self.h, a, self.n_inst))
        return np.array(a[:], dtype=np.int32)

    def noise(self, i):
        K, T, M, NT = self._dims(i)
        a = np.empty((NT + 1) * M * K * 2, dtype=np.float32)
        check(self.lib.dr_bm_get_noise(self.h, i, _fptr(a), a.size))
        return a.reshape(NT + 1, M, K, 2)

    def costs_true(self, i):
        K, T, M, NT = self._dims(i)
        a = np.empty(NT * K, dtype=np.float32)
        check(self.lib.dr_bm_get_costs_true(self.h, i, _fptr(a), a.size))
        return a.reshape(NT, K)

    def branches(self, i):
        K, T, M, NT = self._dims(i)
        u = np.empty(NT * K, dtype=np.int32)
        t = np.empty(NT * K, dtype=np.int32)
        ip = C.POINTER(C.c_int32)
        check(self.lib.dr_bm_get_branches(self.h, i, u.ctypes.data_as(ip), t.ctypes.data_as(ip), u.size))
        return u.reshape(NT, K), t.reshape(NT, K)

    def set_record_lengths(self, on=True):
        """Record the rollout lengths of the following steps (a slower rollout instantiation)."""
        check(self.lib.dr_bm_set_record_lengths(self.h, 1 if on else 0))
        self.recording = bool(on)

    def set_launch(self, rollout_grid=0):
        """rollout_grid 0: one workgroup per rollout tile; > 0: that many persistent workgroups."""
        cfg = self.get_launch()
        cfg.rollout_grid = int(rollout_grid)
        check(self.lib.dr_bm_set_launch(self.h, C.byref(cfg)))

    def get_launch(self):
        cfg = abi.LaunchCfg()
        check(self.lib.dr_bm_get_launch(self.h, C.byref(cfg)))
        return cfg

    def autotune(self, trials=5):
        """Times the rollout grid shapes on this workload and keeps the fastest; state is restored."""
        return _autotune(self.lib.dr_bm_autotune, self.h, trials)

    def rollout_lengths(self, i):
        """Environment steps each rollout of the last step ran [n_models][K] (recording on)."""
        K, T, M, NT = self._dims(i)
        a = np.empty(NT * K, dtype=np.int32)
        check(self.lib.dr_bm_get_rollout_lengths(self.h, i, a.ctypes.data_as(C.POINTER(C.c_int32)), a.size))
        return a.reshape(NT, K)

    def costs(self, i):
        K, T, M, NT = self._dims(i)
        a = np.empty((NT + 1) * K, dtype=np.float32)
        check(self.lib.dr_bm_get_costs(self.h, i, _fptr(a), a.size))
        return a.reshape(NT + 1, K)

    def min_costs(self, i):
        K, T, M, NT = self._dims(i)
        a = np.empty((NT + 1) * M, dtype=np.float32)
        check(self.lib.dr_bm_get_min_costs(self.h, i, _fptr(a), a.size))
        return a.reshape(NT + 1, M)

    def nu(self, i):
        a = np.empty(self.n_models + 1, dtype=np.float32)
        check(self.lib.dr_bm_get_nu(self.h, i, _fptr(a), a.size))
        return a

    def diag(self):
        arr = (abi.BmDiag * self.n_inst)()
        check(self.lib.dr_bm_get_diag(self.h, arr, self.n_inst))
        return arr

    def set_timing(self, on=True):
        check(self.lib.dr_bm_set_timing(self.h, 1 if on else 0))

    def fail_counts(self, i):
        """Verification failures per horizon step of the last step: (new plan, previous plan)."""
        T = self.sc.instances[i].n_timesteps
        a = np.empty(T, dtype=np.int32)
        b = np.empty(T, dtype=np.int32)
        ip = C.POINTER(C.c_int32)
        check(self.lib.dr_bm_get_fail_counts(self.h, i, a.ctypes.data_as(ip), b.ctypes.data_as(ip), T))
        return a, b

    def last_timing(self):
        """Device ms summed over the last timed run: rollout, aggregation + minima, weighted averages,
        update + verification, plan choice + world, step (host Clopper-Pearson time included)."""
        a = np.empty(6, dtype=np.float32)
        check(self.lib.dr_bm_last_timing(self.h, _fptr(a), 6))
        return a


class HoBMEngine(BMEngine):
    """Branching spline MPPI controllers on HiddenObs worlds (two models: which hidden
    obstacle exists); the same engine and calls, with HiddenObs states."""

    _create, _set, _get, _state = "dr_bmho_create", "dr_bmho_set_states", "dr_bmho_get_states", abi.HoState

    @staticmethod
    def _n_models(env):
        return 2


class DraceBMEngine(BMEngine):
    """Branching spline MPPI controllers on DroneRace worlds (model theta: the opponent's PD
    configuration); the same engine and calls, with DroneRace states."""

    _create, _set, _get, _state = "dr_bmdrace_create", "dr_bmdrace_set_states", "dr_bmdrace_get_states", abi.DraceState
