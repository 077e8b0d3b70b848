"""ctypes mirror of include/dr_api.h and loader for the HIP engine library.

The library is built in-tree (drone-race_amd/libdr_amd.so) by build.py.  No
Python fallback exists: if the library is missing, load() raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdr_amd.so")

ACTION_DIM = 4
MAX_GATES = 16
MAX_T = 256
MAX_WORLD_DRONES = 8

MAX_BRANCHES = 8
COSTGRAD_NOALLOC_MAX_FLOATS = 1 << 21  # dr_api.h DR_COSTGRAD_NOALLOC_MAX_FLOATS
MAX_MODELS = 16
SR_TRACK_SAMPLES = 512
SR_MAX_MODELS = 4

DR_OK, DR_EINVAL, DR_EHIP, DR_ENOMEM, DR_ESTATE = 0, 1, 2, 3, 4


class QuadParams(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "mass", "gravity", "drag", "rate_gain", "thrust_min", "thrust_max", "rate_max", "drone_radius", "dt")]


class Gate(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("normal", C.c_float * 3), ("radius", C.c_float), ("_pad", C.c_float)]


class Track(C.Structure):
    _fields_ = [("n_gates", C.c_int32), ("n_win_laps", C.c_int32), ("ground_z", C.c_float), ("_pad", C.c_float),
                ("gates", Gate * MAX_GATES)]


class CostCfg(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "gate_traversal_margin", "coll_dist_factor", "boundary_cost", "boundary_threshold_factor", "outside_cost",
        "win_cost", "opp_win_cost", "final_adv_weight", "final_opp_adv_weight", "opp_prox_radius", "opp_prox_cost",
        "_pad")]


class DroneState(C.Structure):
    _fields_ = [("p", C.c_float * 3), ("q", C.c_float * 4), ("v", C.c_float * 3), ("w", C.c_float * 3),
                ("gate", C.c_int32), ("laps", C.c_int32), ("_pad", C.c_int32)]


class Opponent(C.Structure):
    _fields_ = [("p", C.c_float * 3), ("v", C.c_float * 3)]


class Instance(C.Structure):
    _fields_ = [("n_samples", C.c_int32), ("n_timesteps", C.c_int32), ("track", C.c_int32), ("drone", C.c_int32),
                ("seed", C.c_uint64), ("lambda_", C.c_float), ("sigma", C.c_float * 4), ("shift_fill", C.c_float * 4),
                ("adapt_lambda", C.c_int32)]


class Diag(C.Structure):
    _fields_ = [("min_cost", C.c_float), ("nu", C.c_float), ("lambda_", C.c_float), ("traj_cost", C.c_float),
                ("step", C.c_uint32), ("exec_collisions", C.c_uint32), ("exec_cost", C.c_float), ("exec_progress", C.c_float)]


class LaunchCfg(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "rollout_block", "rollout_grid", "reduce_chunk", "noise_batch", "steps_per_graph", "use_graph", "reduce_tgroup",
        "rollout_split")]


class MPPICfg(C.Structure):
    _fields_ = [("n_samples", C.c_int32), ("n_timesteps", C.c_int32), ("lambda_", C.c_float),
                ("sigma", C.c_float * 4), ("shift_fill", C.c_float * 4), ("adapt_lambda", C.c_int32),
                ("max_opponents", C.c_int32)]


class PrInstance(C.Structure):
    _fields_ = [("n_samples", C.c_int32), ("n_timesteps", C.c_int32), ("track", C.c_int32), ("drone", C.c_int32),
                ("seed", C.c_uint64), ("lambda_", C.c_float), ("sigma", C.c_float * 4), ("shift_fill", C.c_float * 4),
                ("adapt_lambda", C.c_int32), ("n_branches", C.c_int32), ("sampled", C.c_int32), ("n_models", C.c_int32),
                ("safety_weight", C.c_float), ("belief", C.c_float * MAX_BRANCHES),
                ("vel_shift", ((C.c_float * 3) * (MAX_WORLD_DRONES - 1)) * MAX_BRANCHES)]


class PrDiag(C.Structure):
    _fields_ = [("min_costs", C.c_float * 3), ("nu", C.c_float * 3), ("lambda_", C.c_float * 3), ("cand_costs", C.c_float * 2),
                ("use_nom_plan", C.c_int32), ("reset_nom", C.c_int32), ("step", C.c_uint32), ("_pad", C.c_int32)]


class SrOppConfig(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("s1", "s2", "s3", "speed_scale", "action_noise")]


class SrEnv(C.Structure):
    _fields_ = [("dt", C.c_float), ("drone_radius", C.c_float), ("pos_noise", C.c_float), ("speed_noise", C.c_float),
                ("action_noise", C.c_float), ("max_speed", C.c_float * 2), ("max_accel", C.c_float * 2),
                ("track_width", C.c_float), ("n_win_laps", C.c_int32), ("kp", C.c_float), ("kv", C.c_float),
                ("max_opp_lat_dist_fact", C.c_float), ("track_length", C.c_float), ("n_models", C.c_int32),
                ("opp", SrOppConfig * SR_MAX_MODELS), ("p_grid", (C.c_float * 2) * SR_TRACK_SAMPLES),
                ("t_grid", (C.c_float * 2) * SR_TRACK_SAMPLES), ("kappa_grid", C.c_float * SR_TRACK_SAMPLES)]


class SrState(C.Structure):
    _fields_ = [("pos", C.c_float * 4), ("vel", C.c_float * 4), ("s", C.c_float * 2), ("lat_dist", C.c_float * 2),
                ("laps", C.c_int32 * 2)]


class SrCost(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "coll_dist_factor", "boundary_cost", "boundary_threshold_factor", "outside_cost", "win_cost", "opp_win_cost",
        "final_adv_weight", "final_opp_adv_weight")]


class SrInstance(C.Structure):
    _fields_ = [("n_samples", C.c_int32), ("n_timesteps", C.c_int32), ("true_theta", C.c_int32), ("adapt_lambda", C.c_int32),
                ("seed", C.c_uint64), ("lambda_", C.c_float), ("sigma", C.c_float), ("shift_fill", C.c_float * 2),
                ("_pad", C.c_int32 * 2)]


class PrsrInstance(C.Structure):
    _fields_ = [("n_samples", C.c_int32), ("n_timesteps", C.c_int32), ("sampled", C.c_int32), ("n_draws", C.c_int32),
                ("seed", C.c_uint64), ("lambda_", C.c_float), ("sigma", C.c_float), ("safety_weight", C.c_float),
                ("adapt_lambda", C.c_int32), ("belief", C.c_float * SR_MAX_MODELS), ("shift_fill", C.c_float * 2),
                ("true_theta", C.c_int32), ("_pad", C.c_int32)]


MAX_KNOTS = 60


class BmInstance(C.Structure):
    _fields_ = [("n_samples", C.c_int32), ("n_timesteps", C.c_int32), ("n_knots", C.c_int32), ("knots", C.c_int32 * MAX_KNOTS),
                ("seed", C.c_uint64), ("lambda_", C.c_float), ("sigma", C.c_float), ("min_confidence", C.c_float),
                ("true_theta", C.c_int32), ("adapt_lambda", C.c_int32), ("init_belief", C.c_float * SR_MAX_MODELS),
                ("n_verif_samples", C.c_int32), ("verif_horizon", C.c_int32), ("beta", C.c_float), ("max_verif_eps", C.c_float)]


class BmDiag(C.Structure):
    _fields_ = [("lambda_", C.c_float), ("nu", C.c_float), ("min_cost", C.c_float), ("step", C.c_uint32),
                ("branch", C.c_int32), ("use_new_plan", C.c_int32), ("fail_count", C.c_uint32), ("_pad", C.c_int32),
                ("epsilon", C.c_double), ("epsilon_partial", C.c_double), ("certified_loss", C.c_double)]


class HoEnv(C.Structure):
    _fields_ = [("arena_min", C.c_float * 2), ("arena_max", C.c_float * 2), ("dt", C.c_float), ("drone_radius", C.c_float),
                ("pos_noise", C.c_float), ("speed_noise", C.c_float), ("action_noise", C.c_float), ("max_speed", C.c_float),
                ("max_accel", C.c_float), ("n_win_laps", C.c_int32), ("gate_centers", (C.c_float * 2) * 3),
                ("gate_vectors", (C.c_float * 2) * 3), ("gate_radius", C.c_float * 3), ("rect_obstacles", (C.c_float * 4) * 3),
                ("hidden_centers", (C.c_float * 2) * 2), ("hidden_radius", C.c_float * 2), ("dbl_hp_limits", C.c_float * 4),
                ("dbl_hp_lambdas", C.c_float * 4), ("annulus_center", C.c_float * 2), ("annulus_radius", C.c_float * 2)]


class HoState(C.Structure):
    _fields_ = [("pos", C.c_float * 2), ("vel", C.c_float * 2), ("laps", C.c_int32), ("gates", C.c_int32)]


DRACE_SAMPLES, DRACE_MAX_GATES, DRACE_MAX_OBSTACLES, DRACE_MAX_RACELINES = 512, 8, 4, 4


class DracePid(C.Structure):
    _fields_ = [("kp", C.c_float), ("kd", C.c_float), ("repulsion_factor", C.c_float), ("repulsion_power", C.c_float),
                ("repulsion_dist_fact", C.c_float), ("raceline_index", C.c_int32), ("action_noise", C.c_float)]


class DraceEnv(C.Structure):
    _fields_ = [("arena_min", C.c_float * 2), ("arena_max", C.c_float * 2), ("dt", C.c_float), ("drone_radius", C.c_float),
                ("pos_noise", C.c_float), ("speed_noise", C.c_float), ("action_noise", C.c_float), ("max_speed", C.c_float * 2),
                ("max_accel", C.c_float * 2), ("n_win_laps", C.c_int32), ("target_distance", C.c_float), ("n_gates", C.c_int32),
                ("gate_centers", (C.c_float * 2) * DRACE_MAX_GATES), ("gate_vectors", (C.c_float * 2) * DRACE_MAX_GATES),
                ("gate_radius", C.c_float * DRACE_MAX_GATES), ("n_obstacles", C.c_int32),
                ("obstacles", (C.c_float * 4) * DRACE_MAX_OBSTACLES), ("n_round_obstacles", C.c_int32),
                ("round_centers", (C.c_float * 2) * DRACE_MAX_OBSTACLES), ("round_radius", C.c_float * DRACE_MAX_OBSTACLES),
                ("n_racelines", C.c_int32), ("n_models", C.c_int32), ("opp_pid", DracePid * SR_MAX_MODELS), ("i_mppi", C.c_int32),
                ("track_points", ((C.c_float * 2) * DRACE_SAMPLES) * DRACE_MAX_RACELINES)]


class DraceState(C.Structure):
    _fields_ = [("pos", C.c_float * 4), ("vel", C.c_float * 4), ("s", (C.c_float * DRACE_MAX_RACELINES) * 2), ("laps", C.c_int32 * 2),
                ("gates", C.c_int32 * 2)]


MSG_HEADER, MSG_STATE, MSG_EVENT, MSG_DONE = 0, 1, 2, 3
EVT_OUTSIDE, EVT_WINNER, EVT_OPP_WINNER, EVT_TRUNCATED = 0, 1, 2, 3
CONT_MPPI, CONT_PRMPPI = 0, 1
ENV_STRATRACE = 2


class WireEnvInfo(C.Structure):
    _fields_ = [("seed", C.c_int32), ("true_theta", C.c_int32), ("send_states", C.c_int32), ("n_models", C.c_int32),
                ("init_belief", C.c_float * SR_MAX_MODELS), ("arena_min", C.c_float * 2), ("arena_max", C.c_float * 2)]


class WireController(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_samples", C.c_int32), ("n_timesteps", C.c_int32), ("n_knots", C.c_int32),
                ("knots", C.c_int32 * MAX_KNOTS), ("use_splines", C.c_int32), ("lambda_", C.c_float), ("sigma", C.c_float),
                ("gate_traversal_margin", C.c_float), ("cost", SrCost), ("min_confidence", C.c_float), ("n_verif_samples", C.c_int32),
                ("beta", C.c_float), ("verif_horizon", C.c_int32), ("max_verif_eps", C.c_float), ("safety_weight", C.c_float),
                ("min_safe_dist", C.c_float), ("delta", C.c_float), ("P", C.c_int32)]


# Every symbol include/dr_api.h declares, with its ctypes signature.
_P = C.c_void_p
_I = C.c_int
_F = C.POINTER(C.c_float)
SIGNATURES = {
    "dr_last_error": (C.c_char_p, []),
    "dr_version": (_I, []),
    "dr_engine_create": (_I, [C.POINTER(QuadParams), C.POINTER(CostCfg), C.POINTER(Track), _I, C.POINTER(Instance), _I,
                              C.POINTER(C.c_int32), _I, C.POINTER(_P)]),
    "dr_engine_destroy": (None, [_P]),
    "dr_engine_set_drones": (_I, [_P, C.POINTER(DroneState), _I]),
    "dr_engine_get_drones": (_I, [_P, C.POINTER(DroneState), _I]),
    "dr_engine_set_nominal": (_I, [_P, _I, _F, _I]),
    "dr_engine_get_nominal": (_I, [_P, _I, _F, _I]),
    "dr_engine_set_step": (_I, [_P, _I, C.c_uint32]),
    "dr_engine_step": (_I, [_P, _I, _I, _F]),
    "dr_engine_step_async": (_I, [_P, _I, _I]),
    "dr_engine_synchronize": (_I, [_P]),
    "dr_engine_get_costs": (_I, [_P, _I, _F, _I]),
    "dr_engine_get_noise": (_I, [_P, _I, _F, C.c_size_t]),
    "dr_engine_get_actions": (_I, [_P, _F, _I]),
    "dr_engine_get_diag": (_I, [_P, C.POINTER(Diag), _I]),
    "dr_engine_set_launch": (_I, [_P, C.POINTER(LaunchCfg)]),
    "dr_engine_get_launch": (_I, [_P, C.POINTER(LaunchCfg)]),
    "dr_engine_autotune": (_I, [_P, _I, C.POINTER(LaunchCfg), _F]),
    "dr_engine_set_timing": (_I, [_P, _I]),
    "dr_engine_set_eval": (_I, [_P, _I]),
    "dr_engine_set_exec_cost": (_I, [_P, _I]),
    "dr_engine_cost_grad": (_I, [_P, _I, _F, _I, C.c_size_t, _F, _F, _I]),
    "dr_engine_cost_grad_device": (_I, [_P, _I, _P, _I, C.c_size_t, _P, _P, _I, _P]),
    "dr_engine_cost_grad_noalloc": (_I, [_P, _I, _F, _I, C.c_size_t, _F, _F, _I]),
    "dr_engine_set_refine": (_I, [_P, _I, C.c_float, _I]),
    "dr_engine_last_timing": (_I, [_P, _F, _I]),
    "dr_mppi_create": (_I, [C.POINTER(QuadParams), C.POINTER(CostCfg), C.POINTER(Track), C.POINTER(MPPICfg), C.c_uint64,
                            C.POINTER(_P)]),
    "dr_mppi_destroy": (None, [_P]),
    "dr_mppi_step": (_I, [_P, C.POINTER(DroneState), C.POINTER(Opponent), _I, _F]),
    "dr_mppi_get_nominal": (_I, [_P, _F, _I]),
    "dr_mppi_set_nominal": (_I, [_P, _F, _I]),
    "dr_mppi_get_diag": (_I, [_P, C.POINTER(Diag)]),
    "dr_mppi_engine": (_P, [_P]),
    "dr_pr_create": (_I, [C.POINTER(QuadParams), C.POINTER(CostCfg), C.POINTER(Track), _I, C.POINTER(PrInstance), _I,
                          C.POINTER(C.c_int32), _I, C.POINTER(_P)]),
    "dr_pr_destroy": (None, [_P]),
    "dr_pr_set_drones": (_I, [_P, C.POINTER(DroneState), _I]),
    "dr_pr_get_drones": (_I, [_P, C.POINTER(DroneState), _I]),
    "dr_pr_set_nominals": (_I, [_P, _I, _F, _F, _I]),
    "dr_pr_get_nominals": (_I, [_P, _I, _F, _F, _I]),
    "dr_pr_set_belief": (_I, [_P, _I, _F, _I]),
    "dr_pr_get_belief": (_I, [_P, _I, _F, _I]),
    "dr_pr_step": (_I, [_P, _I, _I, _F]),
    "dr_pr_get_diag": (_I, [_P, C.POINTER(PrDiag), _I]),
    "dr_pr_get_noise": (_I, [_P, _I, _F, C.c_size_t]),
    "dr_pr_get_aggregates": (_I, [_P, _I, _F, _I]),
    "dr_pr_get_candidates": (_I, [_P, _I, _F, _F, _I]),
    "dr_pr_get_models": (_I, [_P, _I, C.POINTER(C.c_int32), _I]),
    "dr_pr_get_pair_costs": (_I, [_P, _I, _F, _I]),
    "dr_pr_step_async": (_I, [_P, _I, _I]),
    "dr_pr_synchronize": (_I, [_P]),
    "dr_pr_set_launch": (_I, [_P, C.POINTER(LaunchCfg)]),
    "dr_pr_get_launch": (_I, [_P, C.POINTER(LaunchCfg)]),
    "dr_pr_autotune": (_I, [_P, _I, C.POINTER(LaunchCfg), _F]),
    "dr_pr_set_timing": (_I, [_P, _I]),
    "dr_pr_last_timing": (_I, [_P, _F, _I]),
    "dr_sr_create": (_I, [C.POINTER(SrEnv), C.POINTER(SrCost), C.POINTER(SrInstance), _I, C.POINTER(_P)]),
    "dr_sr_destroy": (None, [_P]),
    "dr_sr_set_states": (_I, [_P, C.POINTER(SrState), _I]),
    "dr_sr_get_states": (_I, [_P, C.POINTER(SrState), _I]),
    "dr_sr_set_nominal": (_I, [_P, _I, _F, _I]),
    "dr_sr_get_nominal": (_I, [_P, _I, _F, _I]),
    "dr_sr_step": (_I, [_P, _I, _I, _F]),
    "dr_sr_get_events": (_I, [_P, C.POINTER(C.c_int32), _I]),
    "dr_sr_get_costs": (_I, [_P, _I, _F, _I]),
    "dr_sr_get_noise": (_I, [_P, _I, _F, C.c_size_t]),
    "dr_sr_get_diag": (_I, [_P, C.POINTER(Diag), _I]),
    "dr_sr_set_launch": (_I, [_P, C.POINTER(LaunchCfg)]),
    "dr_sr_get_launch": (_I, [_P, C.POINTER(LaunchCfg)]),
    "dr_sr_autotune": (_I, [_P, _I, C.POINTER(LaunchCfg), _F]),
    "dr_sr_set_timing": (_I, [_P, _I]),
    "dr_sr_last_timing": (_I, [_P, _F, _I]),
    "dr_prsr_create": (_I, [C.POINTER(SrEnv), C.POINTER(SrCost), C.POINTER(PrsrInstance), _I, C.POINTER(_P)]),
    "dr_pr_set_sr_states": (_I, [_P, C.POINTER(SrState), _I]),
    "dr_pr_get_sr_states": (_I, [_P, C.POINTER(SrState), _I]),
    "dr_pr_get_events": (_I, [_P, C.POINTER(C.c_int32), _I]),
    "dr_wire_decode_env": (_I, [C.c_void_p, C.c_size_t, C.POINTER(SrEnv), C.POINTER(WireEnvInfo)]),
    "dr_wire_decode_controller": (_I, [C.c_void_p, C.c_size_t, C.POINTER(WireController)]),
    "dr_wire_decode_state": (_I, [C.c_void_p, C.c_size_t, C.POINTER(SrState)]),
    "dr_session_create": (_I, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int32, C.POINTER(_P)]),
    "dr_session_destroy": (None, [_P]),
    "dr_session_next": (_I, [_P, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t), C.POINTER(C.c_int32)]),
    "dr_bm_create": (_I, [C.POINTER(SrEnv), C.POINTER(SrCost), C.POINTER(BmInstance), _I, C.POINTER(_P)]),
    "dr_bm_destroy": (None, [_P]),
    "dr_bm_set_states": (_I, [_P, C.POINTER(SrState), _I]),
    "dr_bm_get_states": (_I, [_P, C.POINTER(SrState), _I]),
    "dr_bm_set_belief": (_I, [_P, _I, _F, _I]),
    "dr_bm_get_belief": (_I, [_P, _I, _F, _I]),
    "dr_bm_set_knots": (_I, [_P, _I, _F, _I]),
    "dr_bm_get_knots": (_I, [_P, _I, _F, _I]),
    "dr_bm_get_nominal": (_I, [_P, _I, _F, _I]),
    "dr_bm_get_basis": (_I, [_P, _I, _F, _I]),
    "dr_bm_step": (_I, [_P, _I, _I, _F]),
    "dr_bm_get_events": (_I, [_P, C.POINTER(C.c_int32), _I]),
    "dr_bm_get_noise": (_I, [_P, _I, _F, C.c_size_t]),
    "dr_bm_set_launch": (_I, [_P, C.POINTER(LaunchCfg)]),
    "dr_bm_get_launch": (_I, [_P, C.POINTER(LaunchCfg)]),
    "dr_bm_autotune": (_I, [_P, _I, C.POINTER(LaunchCfg), _F]),
    "dr_bm_get_costs_true": (_I, [_P, _I, _F, _I]),
    "dr_bm_get_branches": (_I, [_P, _I, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _I]),
    "dr_bm_get_costs": (_I, [_P, _I, _F, _I]),
    "dr_bm_get_min_costs": (_I, [_P, _I, _F, _I]),
    "dr_bm_get_nu": (_I, [_P, _I, _F, _I]),
    "dr_bm_set_record_lengths": (_I, [_P, _I]),
    "dr_bm_get_rollout_lengths": (_I, [_P, _I, C.POINTER(C.c_int32), _I]),
    "dr_bm_get_diag": (_I, [_P, C.POINTER(BmDiag), _I]),
    "dr_bm_get_fail_counts": (_I, [_P, _I, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _I]),
    "dr_bm_set_timing": (_I, [_P, _I]),
    "dr_bm_last_timing": (_I, [_P, _F, _I]),
    "dr_bmho_create": (_I, [C.POINTER(HoEnv), C.POINTER(SrCost), C.POINTER(BmInstance), _I, C.POINTER(_P)]),
    "dr_bmho_set_states": (_I, [_P, C.POINTER(HoState), _I]),
    "dr_bmho_get_states": (_I, [_P, C.POINTER(HoState), _I]),
    "dr_bmdrace_create": (_I, [C.POINTER(DraceEnv), C.POINTER(SrCost), C.POINTER(BmInstance), _I, C.POINTER(_P)]),
    "dr_bmdrace_set_states": (_I, [_P, C.POINTER(DraceState), _I]),
    "dr_bmdrace_get_states": (_I, [_P, C.POINTER(DraceState), _I]),
}

_lib = None


class EngineError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the HIP engine library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EngineError(f"HIP engine library not built: {path} is missing (run __graft_entry__.build())")
    # PyTorch ships its own libamdhip64.so.7 (same soname as ROCm's).  Loading
    # torch first makes the engine bind to that runtime, so that torch and the
    # engine share one HIP runtime in the process (tensors, streams, autograd).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != DR_OK:
        msg = _lib.dr_last_error().decode() if _lib is not None else ""
        raise EngineError(f"dr status {rc}: {msg}")
