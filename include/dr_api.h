/*
 * drone-race MI355X control engine: C ABI.
 *
 * Plain pointers and sizes only; every entry point returns a dr_status and
 * never throws.  dr_last_error() returns a thread-local description of the
 * last failure.  Handles are not thread-safe; distinct handles are.  Each
 * handle owns its device memory (a private copy of the tracks included) and
 * one HIP stream; callers own every host buffer passed in.
 *
 * Reference interface being replaced (pablo-gr-rey/drone-race):
 *   Controller::getControl(const SimState&, float* outAction)
 *       include/controllers.h:12-21  -> dr_mppi_step / dr_engine_step
 *   MPPIController(const EnvironmentConfig&, const MPPIConfig&, int seed,
 *       std::optional<std::vector<float>> nominal)
 *       include/controllers.h:23-111 -> dr_mppi_create (+ dr_mppi_set_nominal)
 *   MPPIController::~MPPIController  src/mppi_controller.cu:51-54
 *                                     -> dr_mppi_destroy
 *   MPPIController::h_nominal         include/controllers.h:37
 *                                     -> dr_mppi_get_nominal
 *   MPPIConfig                        include/config.h:116-154 -> dr_mppi_cfg
 *   EnvironmentConfig                 include/environments/env_stratrace_defs.h:53-86
 *                                     -> dr_quad_params + dr_track (+ dr_cost_cfg)
 *   SimState                          include/environments/env_stratrace_defs.h:88-96
 *                                     -> dr_drone_state (quadrotor race), dr_sr_state (StratRace2D)
 *   PRMPPIController(EnvironmentConfig, PRMPPIConfig, seed, nominal),
 *       getControl, h_nom_nominal, h_rob_nominal, useNomPlan, resetNom
 *       include/controllers.h:114-179 -> dr_pr_create / dr_pr_step /
 *                                     dr_pr_get_nominals / dr_pr_get_diag
 *   PRMPPIConfig                      include/config.h:157-192 -> dr_pr_instance
 *   EnvironmentConfig, SimState of the StratRace2D build
 *       include/environments/env_stratrace_defs.h:42-96 -> dr_sr_env, dr_sr_state
 */
#ifndef DR_API_H
#define DR_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DR_ABI_VERSION 1

#define DR_ACTION_DIM 4
#define DR_STATE_DIM 13
#define DR_MAX_GATES 16
#define DR_MAX_T 256
#define DR_MAX_WORLD_DRONES 8

typedef enum dr_status
{
    DR_OK = 0,
    DR_EINVAL = 1,   /* bad argument or size */
    DR_EHIP = 2,     /* HIP runtime error */
    DR_ENOMEM = 3,   /* device or host allocation failed */
    DR_ESTATE = 4,   /* call not valid in the handle's current state */
} dr_status;

/* Quadrotor model.  SI units; dt is the Euler step and the control period. */
typedef struct dr_quad_params
{
    float mass;         /* kg */
    float gravity;      /* m/s^2 */
    float drag;         /* 1/s, linear drag */
    float rate_gain;    /* 1/s, first-order body-rate tracking gain */
    float thrust_min;   /* N */
    float thrust_max;   /* N */
    float rate_max;     /* rad/s */
    float drone_radius; /* m */
    float dt;           /* s */
} dr_quad_params;

typedef struct dr_gate
{
    float center[3];
    float normal[3]; /* unit */
    float radius;
    float _pad;
} dr_gate;

typedef struct dr_track
{
    int32_t n_gates;    /* 2..DR_MAX_GATES */
    int32_t n_win_laps; /* laps counted at each passage of gate 0 */
    float ground_z;     /* m */
    float _pad;
    dr_gate gates[DR_MAX_GATES];
} dr_track;

/* Cost weights.  Field names follow MPPIConfig (include/config.h:116-154). */
typedef struct dr_cost_cfg
{
    float gate_traversal_margin;     /* gate counts if crossed within radius*margin */
    float coll_dist_factor;          /* collision if boundary distance < radius*factor */
    float boundary_cost;             /* near-boundary penalty weight */
    float boundary_threshold_factor; /* penalty starts at radius*factor */
    float outside_cost;              /* collision cost (terminal) */
    float win_cost;                  /* bonus when laps reach n_win_laps (terminal) */
    float opp_win_cost;              /* reserved: opponents' laps are not simulated */
    float final_adv_weight;          /* terminal -weight * race progress */
    float final_opp_adv_weight;      /* reserved: opponents' progress is not simulated */
    float opp_prox_radius;           /* opponent-proximity radius, m */
    float opp_prox_cost;             /* weight of (1 - d/r)^2 inside the radius */
    float _pad;
} dr_cost_cfg;

/* Drone state: position, unit quaternion (w, x, y, z), velocity, body
 * rates, index of the last gate passed, laps. */
typedef struct dr_drone_state
{
    float p[3];
    float q[4];
    float v[3];
    float w[3];
    int32_t gate;
    int32_t laps;
    int32_t _pad;
} dr_drone_state;

/* Constant-velocity opponent prediction seen by one controller. */
typedef struct dr_opponent
{
    float p[3];
    float v[3];
} dr_opponent;

/* One controller (one drone) in a batch. */
typedef struct dr_instance
{
    int32_t n_samples;    /* K >= 1 */
    int32_t n_timesteps;  /* T in 1..DR_MAX_T */
    int32_t track;        /* index into the engine's track table */
    int32_t drone;        /* ego drone slot; opponents = other drones of its world */
    uint64_t seed;        /* Philox key */
    float lambda;         /* temperature: w = exp(-(c - min)/lambda) */
    float sigma[DR_ACTION_DIM];
    float shift_fill[DR_ACTION_DIM]; /* written at t = T-1 by the shift (reference: 0) */
    int32_t adapt_lambda; /* 1: the reference's nu-driven temperature update */
} dr_instance;

typedef struct dr_diag
{
    float min_cost;   /* min over samples of the last rollout */
    float nu;         /* sum of weights of the last update */
    float lambda;     /* temperature used by the next step */
    float traj_cost;  /* noise-free cost of the nominal the last rollout started from */
    uint32_t step;    /* control steps taken */
    uint32_t exec_collisions; /* closed-loop steps whose executed state was a collision (dr_engine_set_exec_cost) */
    float exec_cost;          /* running cost of the executed states, summed over closed-loop steps (dr_engine_set_exec_cost) */
    float exec_progress;      /* race progress of the last executed state (dr_engine_set_exec_cost) */
} dr_diag;

/* Launch shape of the control step; chosen by dr_engine_autotune or set by
 * hand.  Zero fields select the defaults. */
typedef struct dr_launch_cfg
{
    int32_t rollout_block;    /* samples per rollout workgroup: 64, 128, 256 */
    int32_t rollout_grid;     /* 0: one workgroup per tile; >0 persistent grid */
    int32_t reduce_chunk;     /* samples per weighted-average workgroup: 64..4096, multiple of 64 */
    int32_t noise_batch;      /* rollout steps whose noise is generated together: 1, 2, 4 */
    int32_t steps_per_graph;  /* control steps captured per hipGraph */
    int32_t use_graph;        /* 0: plain stream launches */
    int32_t reduce_tgroup;    /* timesteps per weighted-average workgroup: 1..256 */
    int32_t rollout_split;    /* 0: one lane per sample; 1: noise wave + dynamics wave; 2: four-stage wave pipeline;
                                 3: eight-stage wave pipeline (noise_batch 1 or 2); 4: with the adjoint refinement on,
                                 an eight-wave pipeline that reads noise rows made ahead by the previous step's
                                 refinement launch on the CUs it leaves idle, chunks of 4 * noise_batch steps (as 3
                                 without the refinement) */
} dr_launch_cfg;

typedef struct dr_engine dr_engine;

const char* dr_last_error(void);
int dr_version(void);

/* Batch engine.  drone_world[d] groups drones into worlds; an instance's
 * opponents are the other drones of its ego drone's world, predicted at
 * constant velocity from their state at the start of each control step. */
int dr_engine_create(const dr_quad_params* params, const dr_cost_cfg* cost, const dr_track* tracks, int n_tracks,
                     const dr_instance* instances, int n_instances, const int32_t* drone_world, int n_drones,
                     dr_engine** out);
void dr_engine_destroy(dr_engine* e);

int dr_engine_set_drones(dr_engine* e, const dr_drone_state* states, int n_drones);
int dr_engine_get_drones(dr_engine* e, dr_drone_state* states, int n_drones);
int dr_engine_set_nominal(dr_engine* e, int instance, const float* nominal, int n_floats);
int dr_engine_get_nominal(dr_engine* e, int instance, float* nominal, int n_floats);
int dr_engine_set_step(dr_engine* e, int instance, uint32_t step);

/* Run n_steps control steps (rollout, min, weighted average, update, clamp,
 * shift).  advance != 0 integrates every controlled drone with its action
 * after each step (closed loop); actions (n_instances*4, may be NULL)
 * receives the last step's actions.  Synchronous. */
int dr_engine_step(dr_engine* e, int n_steps, int advance, float* actions);
/* Same work, enqueued only; dr_engine_synchronize waits for it. */
int dr_engine_step_async(dr_engine* e, int n_steps, int advance);
int dr_engine_synchronize(dr_engine* e);

/* Results of the last control step, for tests and telemetry. */
int dr_engine_get_costs(dr_engine* e, int instance, float* costs, int n);
/* Noise rows, layout [k][t][4]: those of the last control step, or with
 * rollout_split 4 and the refinement on those made ahead for the next one. */
int dr_engine_get_noise(dr_engine* e, int instance, float* noise, size_t n);
int dr_engine_get_actions(dr_engine* e, float* actions, int n_floats);
int dr_engine_get_diag(dr_engine* e, dr_diag* diag, int n_instances);

int dr_engine_set_launch(dr_engine* e, const dr_launch_cfg* cfg);
int dr_engine_get_launch(dr_engine* e, dr_launch_cfg* cfg);
/* Time candidate launch shapes on the engine's own workload (state is
 * saved and restored) and keep the fastest.  trials = replays per shape. */
int dr_engine_autotune(dr_engine* e, int trials, dr_launch_cfg* best, float* best_ms);
/* on != 0: each step also rolls the instance's nominal out without noise
 * (dr_diag.traj_cost), on a second stream concurrently with the sampled
 * rollout.  Off by default; traj_cost is not updated while off. */
int dr_engine_set_eval(dr_engine* e, int on);
/* on != 0: every later closed-loop step (advance != 0) also evaluates, for
 * each controlled drone, the running cost of the state it was just moved to
 * (opponents where the other drones of its world now are, undiscounted) and
 * adds it to dr_diag.exec_cost, counts the states in collision in
 * exec_collisions and writes the state's race progress to exec_progress: the
 * quality of the executed trajectory.  Turning it on clears the three.  Off
 * by default (one more small launch per step).  No reference counterpart:
 * the reference prints the executed state instead (engine.cpp:336-338). */
int dr_engine_set_exec_cost(dr_engine* e, int on);
/* on != 0: later steps run as plain stream launches with HIP events
 * between the kernels (no graph), so that dr_engine_last_timing can report
 * per-kernel device time. */
int dr_engine_set_timing(dr_engine* e, int on);
/* Device time of the last timed dr_engine_step / step_async batch, per
 * kernel kind, from HIP events on the engine stream (ms, summed over steps),
 * the first n (1..5) of: rollout, weighted average, update (with the
 * closed-loop advance when it ran), adjoint refinement (0 without it; with
 * it, the MPPI update runs inside the refinement kernel and is counted
 * there), whole step. */
int dr_engine_last_timing(dr_engine* e, float* ms, int n);

/* Adjoint of the noise-free rollout cost.  For n_seq control sequences
 * U[n_seq][T][4] of `instance` (from its current drone state, opponents and
 * track) writes J[n_seq] and dJdU[n_seq][T][4] (may be NULL), checkpointing
 * every ckpt_every steps.  n_floats is the length of U (and of dJdU): it must
 * equal n_seq * T * 4 for the instance's T, else DR_EINVAL.  Host pointers;
 * synchronous.  No reference counterpart: the reference has no
 * differentiable rollout. */
int dr_engine_cost_grad(dr_engine* e, int instance, const float* U, int n_seq, size_t n_floats, float* J, float* dJdU,
                        int ckpt_every);
/* Same with device pointers, enqueued on `stream` (a hipStream_t; NULL = the
 * engine's stream).  Used by the PyTorch autograd wrapper. */
int dr_engine_cost_grad_device(dr_engine* e, int instance, const float* U, int n_seq, size_t n_floats, float* J, float* dJdU,
                               int ckpt_every, void* stream);
/* No-allocation variant of dr_engine_cost_grad, for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid allocation): the same adjoint kernel, with U, J and
 * dJdU staged through a fixed buffer of DR_COSTGRAD_NOALLOC_MAX_FLOATS floats
 * (8 MiB) that the library holds on each GPU for the life of the process;
 * the call makes no device allocation.  n_seq * T * 8 + n_seq must not exceed
 * that size (DR_EINVAL otherwise).  Every call on a GPU shares that GPU's
 * buffer: calls that overlap on one GPU (other threads, other handles) race.
 * No counterpart in the reference. */
#define DR_COSTGRAD_NOALLOC_MAX_FLOATS (1 << 21)
int dr_engine_cost_grad_noalloc(dr_engine* e, int instance, const float* U, int n_seq, size_t n_floats, float* J, float* dJdU,
                                int ckpt_every);
/* Adjoint refinement inside every control step: after the MPPI update,
 * n_grad_steps projected gradient steps U <- clamp(U - step * dJ/dU) on each
 * instance's nominal (ckpt_every in 1..256 is accepted for the interface of
 * the checkpointed adjoint; the refinement keeps every state of the
 * horizon on chip, so it changes neither the result nor the work), the J of
 * every iterate evaluated and the lowest kept (the MPPI update itself on ties),
 * then action and shift.  n_grad_steps = 0 (default) is plain MPPI.  No
 * counterpart in the reference (its controllers have no gradient path). */
int dr_engine_set_refine(dr_engine* e, int n_grad_steps, float step, int ckpt_every);

/* ---------------------------------------------------------------------
 * PR-MPPI ("Parameter-Robust MPPI"), restating PRMPPIController
 * (include/controllers.h:114-179, src/prmppi_controller.cu:142-349) for the
 * quadrotor race.  Opponent-behaviour models: branch b shifts every
 * opponent's constant velocity by vel_shift[b][j] and has belief[b].
 * sampled = 0: the models are the B branches, costs averaged with the
 * belief; sampled = 1: P = n_models models are drawn from the belief each
 * step (prmppi_kernels.cu:23-51) and averaged uniformly (the reference). */
#define DR_MAX_BRANCHES 8
#define DR_MAX_MODELS 16

typedef struct dr_pr_instance
{
    int32_t n_samples;
    int32_t n_timesteps;
    int32_t track;
    int32_t drone;
    uint64_t seed;
    float lambda;                     /* initial temperature of the three weightings */
    float sigma[DR_ACTION_DIM];
    float shift_fill[DR_ACTION_DIM];
    int32_t adapt_lambda;             /* 1: prmppi_controller.cu:328-346 */
    int32_t n_branches;               /* B in 1..DR_MAX_BRANCHES */
    int32_t sampled;                  /* 0: belief-weighted branches; 1: sampled models */
    int32_t n_models;                 /* P when sampled, 1..DR_MAX_MODELS */
    float safety_weight;              /* added to a sample's cost if any model is unsafe */
    float belief[DR_MAX_BRANCHES];
    float vel_shift[DR_MAX_BRANCHES][DR_MAX_WORLD_DRONES - 1][3];
} dr_pr_instance;

typedef struct dr_pr_diag
{
    float min_costs[3];  /* nominal/full, robust/full, robust/safety */
    float nu[3];
    float lambda[3];
    float cand_costs[2]; /* model-averaged full cost of the two candidates */
    int32_t use_nom_plan;
    int32_t reset_nom;
    uint32_t step;
    int32_t _pad;
} dr_pr_diag;

typedef struct dr_prengine dr_prengine;

int dr_pr_create(const dr_quad_params* params, const dr_cost_cfg* cost, const dr_track* tracks, int n_tracks,
                 const dr_pr_instance* instances, int n_instances, const int32_t* drone_world, int n_drones, dr_prengine** out);
void dr_pr_destroy(dr_prengine* e);
int dr_pr_set_drones(dr_prengine* e, const dr_drone_state* states, int n_drones);
int dr_pr_get_drones(dr_prengine* e, dr_drone_state* states, int n_drones);
/* Both plans (T x 4 each, nom may equal rob).  As the reference does at the
 * start of every getControl, the plans are shifted by one step before they
 * are used (prmppi_controller.cu:163-175). */
int dr_pr_set_nominals(dr_prengine* e, int instance, const float* nom, const float* rob, int n_floats);
/* The plans after the last step (unshifted, as h_nom_nominal / h_rob_nominal). */
int dr_pr_get_nominals(dr_prengine* e, int instance, float* nom, float* rob, int n_floats);
int dr_pr_set_belief(dr_prengine* e, int instance, const float* belief, int n_branches);
/* The belief as the engine holds it (on StratRace2D the closed-loop world step updates it, as
 * SimulationEngine::dynStep updates h_belief, engine.cpp:286-312). */
int dr_pr_get_belief(dr_prengine* e, int instance, float* belief, int n_branches);
int dr_pr_step(dr_prengine* e, int n_steps, int advance, float* actions);
int dr_pr_get_diag(dr_prengine* e, dr_pr_diag* diag, int n_instances);
/* Results of the last step for tests: noise [k][t][4]; per-sample
 * aggregates [k] x (nominal full cost, nominal max safety, robust full
 * cost, robust max safety); candidates cand1, cand2 (T x 4); branch of each
 * model this step (n_models ints). */
int dr_pr_get_noise(dr_prengine* e, int instance, float* noise, size_t n);
int dr_pr_get_aggregates(dr_prengine* e, int instance, float* agg, int n_floats);
int dr_pr_get_candidates(dr_prengine* e, int instance, float* cand1, float* cand2, int n_floats);
int dr_pr_get_models(dr_prengine* e, int instance, int32_t* branch_of_model, int n_models);
/* Cost and max safety of every rollout of the last step, [plan][m][k][2]. */
int dr_pr_get_pair_costs(dr_prengine* e, int instance, float* out, int n_floats);
int dr_pr_step_async(dr_prengine* e, int n_steps, int advance);
int dr_pr_synchronize(dr_prengine* e);
/* rollout_block: waves*64 per rollout workgroup (0: every rollout of a
 * sample group, at most 16 waves); reduce_chunk <= 2048; noise_batch and
 * rollout_split are unused. */
int dr_pr_set_launch(dr_prengine* e, const dr_launch_cfg* cfg);
int dr_pr_get_launch(dr_prengine* e, dr_launch_cfg* cfg);
/* Times the launch shapes (rollout waves per workgroup x reduce_chunk x
 * reduce_tgroup) on the engine's own workload, trials steps each, and keeps
 * the fastest; the plans and state the timing steps change are restored.
 * No reference counterpart (its block size is fixed, prmppi_controller.cu:185). */
int dr_pr_autotune(dr_prengine* e, int trials, dr_launch_cfg* best, float* best_ms);
int dr_pr_set_timing(dr_prengine* e, int enable);
/* Device ms summed over the last timed run: rollout, aggregation, weighted
 * averages, update (+ closed-loop advance), whole step. */
int dr_pr_last_timing(dr_prengine* e, float* ms, int n);

/* ---------------------------------------------------------------------
 * StratRace2D: the reference's compiled environment (2-D race on a closed
 * track, one opponent driven by the reference's opponent model), restating
 * include/environments/env_stratrace_defs.h and env_stratrace.h:33-894 and
 * the running/terminal costs of include/costs.cuh:7-39, with plain (dense,
 * single committed opponent model) MPPI on top (src/kernels.cu:34-171,
 * 307-456, src/mppi_controller.cu:250-466).  Control = 2-D acceleration. */
#define DR_SR_TRACK_SAMPLES 512
#define DR_SR_MAX_MODELS 4

typedef struct dr_sr_opp_config  /* OppConfig, env_stratrace_defs.h:42-50, one opponent */
{
    float s1;           /* overtaking lateral distance */
    float s2;           /* longitudinal influence exp(-s2 ds^2) */
    float s3;           /* blocking strength */
    float speed_scale;
    float action_noise;
} dr_sr_opp_config;

typedef struct dr_sr_env         /* EnvironmentConfig, env_stratrace_defs.h:53-86 */
{
    float dt;
    float drone_radius;
    float pos_noise;              /* posNoiseLevel */
    float speed_noise;            /* speedNoiseLevel */
    float action_noise;           /* actionNoiseLevel */
    float max_speed[2];           /* ego, opponent */
    float max_accel[2];
    float track_width;
    int32_t n_win_laps;
    float kp;
    float kv;
    float max_opp_lat_dist_fact;
    float track_length;
    int32_t n_models;             /* N_TRUE_MODELS in 1..DR_SR_MAX_MODELS */
    dr_sr_opp_config opp[DR_SR_MAX_MODELS];
    float p_grid[DR_SR_TRACK_SAMPLES][2];   /* centre line at s = L i / 512 */
    float t_grid[DR_SR_TRACK_SAMPLES][2];   /* unit tangent */
    float kappa_grid[DR_SR_TRACK_SAMPLES];  /* curvature */
} dr_sr_env;

typedef struct dr_sr_state       /* SimState, env_stratrace_defs.h:88-96 (ego first) */
{
    float pos[4];
    float vel[4];
    float s[2];
    float lat_dist[2];
    int32_t laps[2];
} dr_sr_state;

typedef struct dr_sr_cost        /* cost fields of MPPIConfig, config.h:116-154 */
{
    float coll_dist_factor;
    float boundary_cost;
    float boundary_threshold_factor;
    float outside_cost;
    float win_cost;
    float opp_win_cost;
    float final_adv_weight;
    float final_opp_adv_weight;
} dr_sr_cost;

typedef struct dr_sr_instance
{
    int32_t n_samples;
    int32_t n_timesteps;
    int32_t true_theta;   /* opponent model the rollouts and the world follow */
    int32_t adapt_lambda; /* 1: mppi_controller.cu:413-424 */
    uint64_t seed;
    float lambda;         /* temperature ("invTemperature") */
    float sigma;          /* samplingNoise, both control dimensions */
    float shift_fill[2];  /* freed last step after the shift (the reference: 0) */
    int32_t _pad[2];
} dr_sr_instance;

/* PR-MPPI on StratRace2D (PRMPPIController as the reference runs it on its
 * environment): models are the environment's true models theta; sampled =
 * 1 draws P = n_draws models from the belief each step and averages them
 * uniformly (prmppi_kernels.cu:23-51, :160), sampled = 0 weights all
 * n_models with the belief.  Running cost PRMPPIstateCost, safety
 * PRMPPIsafetyCost, terminal finalCost (costs.cuh:44-86) with the fields of
 * dr_sr_cost.  Plans are T x 2. */
typedef struct dr_prsr_instance
{
    int32_t n_samples;
    int32_t n_timesteps;
    int32_t sampled;
    int32_t n_draws;       /* P when sampled, 1..DR_MAX_MODELS */
    uint64_t seed;
    float lambda;          /* initial temperature of the three weightings */
    float sigma;           /* samplingNoise */
    float safety_weight;
    int32_t adapt_lambda;
    float belief[DR_SR_MAX_MODELS];
    float shift_fill[2];
    int32_t true_theta;    /* opponent model of the instance's world (closed loop) */
    int32_t _pad;
} dr_prsr_instance;

typedef struct dr_sr_engine dr_sr_engine;

int dr_sr_create(const dr_sr_env* env, const dr_sr_cost* cost, const dr_sr_instance* instances, int n_instances, dr_sr_engine** out);
void dr_sr_destroy(dr_sr_engine* e);
/* One world state per instance. */
int dr_sr_set_states(dr_sr_engine* e, const dr_sr_state* states, int n);
int dr_sr_get_states(dr_sr_engine* e, dr_sr_state* states, int n);
int dr_sr_set_nominal(dr_sr_engine* e, int instance, const float* nominal, int n_floats); /* T x 2 */
int dr_sr_get_nominal(dr_sr_engine* e, int instance, float* nominal, int n_floats);
/* n_steps control steps of every instance; advance != 0 steps each world
 * with its action (SimulationEngine::dynStep, engine.cpp:286-312, with the
 * environment's noise); a world stops at its first event. */
int dr_sr_step(dr_sr_engine* e, int n_steps, int advance, float* actions /* n x 2, may be NULL */);
/* Last world event per instance: 0 none, 1 ego outside, 2 ego wins, 3 opponent wins. */
int dr_sr_get_events(dr_sr_engine* e, int32_t* events, int n);
int dr_sr_get_costs(dr_sr_engine* e, int instance, float* costs, int n);
/* Control noise of the last step, [k][t][2]. */
int dr_sr_get_noise(dr_sr_engine* e, int instance, float* noise, size_t n);
int dr_sr_get_diag(dr_sr_engine* e, dr_diag* diag, int n);
/* rollout_block in {64, 128, 256}, reduce_chunk, reduce_tgroup, use_graph,
 * steps_per_graph are used. */
int dr_sr_set_launch(dr_sr_engine* e, const dr_launch_cfg* cfg);
int dr_sr_get_launch(dr_sr_engine* e, dr_launch_cfg* cfg);
/* As dr_pr_autotune over rollout_block x reduce_chunk x reduce_tgroup.  No
 * reference counterpart (mppi_controller.cu:277 fixes the block). */
int dr_sr_autotune(dr_sr_engine* e, int trials, dr_launch_cfg* best, float* best_ms);
int dr_sr_set_timing(dr_sr_engine* e, int enable);
/* Device ms summed over the last timed run: rollout, weighted average, update + world step, whole step. */
int dr_sr_last_timing(dr_sr_engine* e, float* ms, int n);

/* A PR-MPPI engine over StratRace2D worlds (one per instance); the other
 * dr_pr_* calls apply, with T x 2 plans and [k][t][2] noise. */
int dr_prsr_create(const dr_sr_env* env, const dr_sr_cost* cost, const dr_prsr_instance* instances, int n_instances, dr_prengine** out);
int dr_pr_set_sr_states(dr_prengine* e, const dr_sr_state* states, int n);
int dr_pr_get_sr_states(dr_prengine* e, dr_sr_state* states, int n);
/* Last world event per instance (0 none, 1 ego outside, 2 ego wins, 3 opponent wins). */
int dr_pr_get_events(dr_prengine* e, int32_t* events, int n);

/* ---------------------------------------------------------------------
 * Branching MPPI on StratRace2D: the reference's default MPPIController
 * (USE_SPLINES, src/kernels.cu:34-486, src/mppi_controller.cu:56-141,
 * 250-466) without plan verification.  Plans: n_models + 1 branch plans
 * (0 = nominal, theta + 1 = specialised to model theta), each M spline
 * knots x 2; rollouts follow every true model compatible with the
 * committed branch, with belief updates and branching in the rollout;
 * masked minima and belief-weighted averages per (plan, knot); natural
 * cubic spline interpolation, L2 clamp of the knots, spline shift of the
 * chosen branch. */
#define DR_MAX_KNOTS 60

typedef struct dr_bm_instance
{
    int32_t n_samples;
    int32_t n_timesteps;
    int32_t n_knots;                 /* M in 2..DR_MAX_KNOTS */
    int32_t knots[DR_MAX_KNOTS];     /* tau_m: increasing, tau_0 >= 0, tau_{M-1} <= T - 1 */
    uint64_t seed;
    float lambda;                    /* invTemperature */
    float sigma;                     /* samplingNoise */
    float min_confidence;            /* belief above which a branch commits */
    int32_t true_theta;              /* opponent model of the instance's world */
    int32_t adapt_lambda;
    float init_belief[DR_SR_MAX_MODELS];
    /* plan verification (verifyNominalFailureKernel + Clopper-Pearson, kernels.cu:488-551,
     * mppi_controller.cu:352-377, 468-514); n_verif_samples = 0 runs none and always uses the
     * new plan (the reference would keep the previous plan, its certified loss being 1) */
    int32_t n_verif_samples;         /* nVerifSamples, 0..2^22 */
    int32_t verif_horizon;           /* verifHorizon, 1..T */
    float beta;                      /* confidence parameter of the bounds */
    float max_verif_eps;             /* maxVerifEps: the new plan is used iff its certified loss is below */
} dr_bm_instance;

typedef struct dr_bm_engine dr_bm_engine;

int dr_bm_create(const dr_sr_env* env, const dr_sr_cost* cost, const dr_bm_instance* instances, int n_instances, dr_bm_engine** out);
void dr_bm_destroy(dr_bm_engine* e);
int dr_bm_set_states(dr_bm_engine* e, const dr_sr_state* states, int n);
int dr_bm_get_states(dr_bm_engine* e, dr_sr_state* states, int n);
/* The controller's belief (h_belief), n_models values per instance. */
int dr_bm_set_belief(dr_bm_engine* e, int instance, const float* belief, int n_models);
int dr_bm_get_belief(dr_bm_engine* e, int instance, float* belief, int n_models);
/* Spline knots of every branch plan, [n_models + 1][M][2]. */
int dr_bm_set_knots(dr_bm_engine* e, int instance, const float* knots, int n_floats);
int dr_bm_get_knots(dr_bm_engine* e, int instance, float* knots, int n_floats);
/* Dense plans after the last step ([n_models + 1][T][2]; the chosen branch shifted). */
int dr_bm_get_nominal(dr_bm_engine* e, int instance, float* nominal, int n_floats);
/* The natural cubic spline basis B ([T][M]), mppi_controller.cu:67-141. */
int dr_bm_get_basis(dr_bm_engine* e, int instance, float* B, int n_floats);
int dr_bm_step(dr_bm_engine* e, int n_steps, int advance, float* actions /* n x 2, may be NULL */);
int dr_bm_get_events(dr_bm_engine* e, int32_t* events, int n);
/* Results of the last step: noise [n_models + 1][M][K][2]; per-model
 * costs [n_models][K] (+inf where the model was skipped); branch used and
 * branching time per model [n_models][K]; aggregated costs [n_models + 1][K];
 * masked minima [n_models + 1][M]; nu [n_models + 1]. */
int dr_bm_get_noise(dr_bm_engine* e, int instance, float* noise, size_t n);
/* Launch shape of the branching rollout: rollout_grid = 0 runs one workgroup
 * per rollout tile, > 0 a persistent grid of that many workgroups looping over
 * the tiles; the other fields are ignored (get_launch reports the fixed
 * ones).  dr_bm_autotune times rollout_grid in {0, 1, 2, 4, 8 per CU} on the
 * engine's own workload, trials steps each, keeps the fastest and restores the
 * knots, plans and state the timing steps change.  No reference counterpart. */
int dr_bm_set_launch(dr_bm_engine* e, const dr_launch_cfg* cfg);
int dr_bm_get_launch(dr_bm_engine* e, dr_launch_cfg* cfg);
int dr_bm_autotune(dr_bm_engine* e, int trials, dr_launch_cfg* best, float* best_ms);
int dr_bm_get_costs_true(dr_bm_engine* e, int instance, float* costs, int n);
int dr_bm_get_branches(dr_bm_engine* e, int instance, int32_t* used, int32_t* time, int n);
int dr_bm_get_costs(dr_bm_engine* e, int instance, float* costs, int n);
int dr_bm_get_min_costs(dr_bm_engine* e, int instance, float* mins, int n);
int dr_bm_get_nu(dr_bm_engine* e, int instance, float* nu, int n);
/* Environment steps each rollout of the last step ran, per model [n_models][K] (T unless a terminal
 * event ended it; 0 for a model skipped by the commitment), recorded while dr_bm_set_record_lengths
 * is on (off by default: recording runs a slower rollout instantiation).  No counterpart in the
 * reference: it measures the work of a rollout launch. */
int dr_bm_set_record_lengths(dr_bm_engine* e, int enable);
int dr_bm_get_rollout_lengths(dr_bm_engine* e, int instance, int32_t* lengths, int n);
typedef struct dr_bm_diag
{
    float lambda;           /* temperature of the next step */
    float nu;               /* nu of the chosen branch in the last step */
    float min_cost;         /* masked minimum of the chosen branch at knot 0 */
    uint32_t step;          /* control steps taken */
    int32_t branch;         /* branch of the last action: 0 nominal, theta + 1 committed */
    int32_t use_new_plan;   /* 1: the new plan was executed, 0: the previous one */
    uint32_t fail_count;    /* verification failures of the executed plan */
    int32_t _pad;
    double epsilon;         /* epsilonPartial + verifHorizon * maxVerifEps */
    double epsilon_partial; /* Clopper-Pearson upper bound of the failure rate */
    double certified_loss;  /* max over t of upper(new) - lower(old); -1 before any test */
} dr_bm_diag;
/* Per-instance diagnostics (what getControl prints, mppi_controller.cu:394-424). */
int dr_bm_get_diag(dr_bm_engine* e, dr_bm_diag* diag, int n);
/* Verification failures per horizon step of the last step, new and previous plan ([T] each). */
int dr_bm_get_fail_counts(dr_bm_engine* e, int instance, int32_t* fail_new, int32_t* fail_old, int n);
int dr_bm_set_timing(dr_bm_engine* e, int enable);
/* Device ms summed over the last timed run (6 values): rollout, aggregation + masked minima, weighted
 * averages, update + verification, plan choice + world, step (with verification the step includes the
 * host's Clopper-Pearson bounds). */
int dr_bm_last_timing(dr_bm_engine* e, float* ms, int n);

/* ---------------------------------------------------------------------
 * HiddenObs: the reference's second 2-D environment (include/environments/
 * env_hiddenobs_defs.h:48-105, env_hiddenobs.h): one drone, three gates,
 * three rectangles, two double half-planes, a quarter annulus and two hidden
 * circular obstacles of which model theta in {0, 1} exists; the belief
 * collapses on the true model once either hidden obstacle is in sight.
 * The branching spline MPPI engine runs on it (dr_bmho_create; every other
 * dr_bm_* call applies, with two models). */
typedef struct dr_ho_env
{
    float arena_min[2];
    float arena_max[2];
    float dt;
    float drone_radius;
    float pos_noise;
    float speed_noise;
    float action_noise;
    float max_speed;               /* L2 */
    float max_accel;               /* L2 */
    int32_t n_win_laps;
    float gate_centers[3][2];
    float gate_vectors[3][2];
    float gate_radius[3];
    float rect_obstacles[3][4];    /* xmin, ymin, xmax, ymax */
    float hidden_centers[2][2];
    float hidden_radius[2];
    float dbl_hp_limits[4];        /* minx1, maxx1, minx2, maxx2 */
    float dbl_hp_lambdas[4];       /* lambda1top, lambda1bot, lambda2top, lambda2bot */
    float annulus_center[2];
    float annulus_radius[2];       /* inner, outer */
} dr_ho_env;

typedef struct dr_ho_state
{
    float pos[2];
    float vel[2];
    int32_t laps;
    int32_t gates;
} dr_ho_state;

int dr_bmho_create(const dr_ho_env* env, const dr_sr_cost* cost, const dr_bm_instance* instances, int n_instances, dr_bm_engine** out);
int dr_bmho_set_states(dr_bm_engine* e, const dr_ho_state* states, int n);
int dr_bmho_get_states(dr_bm_engine* e, dr_ho_state* states, int n);

/* ---------------------------------------------------------------------
 * DroneRace: the reference's third 2-D environment (include/environments/
 * env_dronerace_defs.h:45-112, env_dronerace.h): two agents race through
 * gates; the opponent is a PD controller (PIDConfig) following one of
 * several racelines sampled at 512 points, model theta is its PD
 * configuration; the belief is updated from the opponent's applied action.
 * The reference fixes the counts at compile time (N_GATES, N_OBSTACLES,
 * N_ROUND_OBSTACLES, N_RACELINES, N_TRUE_MODELS); here they are fields with
 * the maxima below.  The branching spline MPPI engine runs on it
 * (dr_bmdrace_create; every other dr_bm_* call applies). */
#define DR_DRACE_SAMPLES 512
#define DR_DRACE_MAX_GATES 8
#define DR_DRACE_MAX_OBSTACLES 4
#define DR_DRACE_MAX_RACELINES 4

typedef struct dr_drace_pid           /* PIDConfig, env_dronerace_defs.h:45-58 */
{
    float kp;
    float kd;
    float repulsion_factor;
    float repulsion_power;
    float repulsion_dist_fact;
    int32_t raceline_index;
    float action_noise;
} dr_drace_pid;

typedef struct dr_drace_env           /* EnvironmentConfig, env_dronerace_defs.h:61-97 */
{
    float arena_min[2];
    float arena_max[2];
    float dt;
    float drone_radius;
    float pos_noise;
    float speed_noise;
    float action_noise;
    float max_speed[2];                /* per agent, L2 */
    float max_accel[2];                /* per agent, L2 */
    int32_t n_win_laps;
    float target_distance;             /* PD target at S + targetDistance */
    int32_t n_gates;                   /* 2..DR_DRACE_MAX_GATES */
    float gate_centers[DR_DRACE_MAX_GATES][2];
    float gate_vectors[DR_DRACE_MAX_GATES][2];
    float gate_radius[DR_DRACE_MAX_GATES];
    int32_t n_obstacles;               /* rectangles, 0..4 */
    float obstacles[DR_DRACE_MAX_OBSTACLES][4];  /* xmin, ymin, xmax, ymax */
    int32_t n_round_obstacles;         /* 0..4 */
    float round_centers[DR_DRACE_MAX_OBSTACLES][2];
    float round_radius[DR_DRACE_MAX_OBSTACLES];
    int32_t n_racelines;               /* 1..DR_DRACE_MAX_RACELINES */
    int32_t n_models;                  /* N_TRUE_MODELS, 1..DR_SR_MAX_MODELS */
    dr_drace_pid opp_pid[DR_SR_MAX_MODELS];
    int32_t i_mppi;                    /* the controlled agent (0 or 1) */
    float track_points[DR_DRACE_MAX_RACELINES][DR_DRACE_SAMPLES][2];
} dr_drace_env;

typedef struct dr_drace_state         /* SimState, env_dronerace_defs.h:99-107 */
{
    float pos[4];                      /* agent-major (x, y) */
    float vel[4];
    float s[2][DR_DRACE_MAX_RACELINES];/* progress on each raceline in [0, 1); < 0: not tracked */
    int32_t laps[2];
    int32_t gates[2];
} dr_drace_state;

int dr_bmdrace_create(const dr_drace_env* env, const dr_sr_cost* cost, const dr_bm_instance* instances, int n_instances,
                      dr_bm_engine** out);
int dr_bmdrace_set_states(dr_bm_engine* e, const dr_drace_state* states, int n);
int dr_bmdrace_get_states(dr_bm_engine* e, dr_drace_state* states, int n);

/* ---------------------------------------------------------------------
 * Wire protocol and simulation session of the reference's front-end for the
 * StratRace2D build: the byte codec (include/protocol.h:9-142,
 * src/protocol.cpp, python-src/protocol.py:13-116), the header decoders
 * (src/environments/env_stratrace.cpp:98-226, src/config.cpp:5-115) and
 * SimulationEngine (include/engine.h, src/engine.cpp:14-363).  Frames are
 * little-endian; an array is an int32 length followed by that many float32
 * values (integer arrays included). */
enum { DR_MSG_HEADER = 0, DR_MSG_STATE = 1, DR_MSG_EVENT = 2, DR_MSG_DONE = 3 };                /* MsgType, protocol.h:9-15 */
enum { DR_EVT_OUTSIDE = 0, DR_EVT_WINNER = 1, DR_EVT_OPP_WINNER = 2, DR_EVT_TRUNCATED = 3 };  /* EventType, protocol.h:17-23 */
enum { DR_CONT_MPPI = 0, DR_CONT_PRMPPI = 1 };                                               /* CONTROLLER_KIND, utils.py:31-34 */
#define DR_ENV_STRATRACE 2                                                                    /* ENV_KIND, utils.py:37-41 */

/* Environment-header fields that are not part of dr_sr_env. */
typedef struct dr_wire_env_info
{
    int32_t seed;          /* controller and world seed */
    int32_t true_theta;    /* opponent model of the world */
    int32_t send_states;   /* 0: no STATE frames */
    int32_t n_models;      /* modelSizes[0] */
    float init_belief[DR_SR_MAX_MODELS];
    float arena_min[2];
    float arena_max[2];
} dr_wire_env_info;

/* Controller header (MPPIConfig / PRMPPIConfig, config.h:116-192). */
typedef struct dr_wire_controller
{
    int32_t kind;                   /* DR_CONT_MPPI or DR_CONT_PRMPPI */
    int32_t n_samples;
    int32_t n_timesteps;
    int32_t n_knots;                /* MPPI */
    int32_t knots[DR_MAX_KNOTS];    /* MPPI */
    int32_t use_splines;            /* MPPI; 1 (the reference is built with USE_SPLINES) */
    float lambda;                   /* invTemperature */
    float sigma;                    /* samplingNoise */
    float gate_traversal_margin;    /* unused by StratRace2D */
    dr_sr_cost cost;                /* outside_cost is MPPI only */
    float min_confidence;           /* MPPI */
    int32_t n_verif_samples;        /* MPPI plan verification (not run) */
    float beta;
    int32_t verif_horizon;
    float max_verif_eps;
    float safety_weight;            /* PR-MPPI */
    float min_safe_dist;            /* PR-MPPI, unused by StratRace2D */
    float delta;                    /* PR-MPPI */
    int32_t P;                      /* PR-MPPI models drawn per step */
} dr_wire_controller;

/* Header decoders (unpackEnvConfig env_stratrace.cpp:127-226, loadControllerConfig config.cpp:86-115,
 * unpackSimState env_stratrace.cpp:98-118): DR_EINVAL with the reason in dr_last_error() on any
 * mismatch of message type, kind, sizes, array lengths or trailing bytes. */
int dr_wire_decode_env(const void* frame, size_t len, dr_sr_env* env, dr_wire_env_info* info);
int dr_wire_decode_controller(const void* frame, size_t len, dr_wire_controller* cont);
int dr_wire_decode_state(const void* frame, size_t len, dr_sr_state* state);

/* SimulationEngine(envConfig, contConfig, initState, seed, trueTheta) from the three header frames
 * (protocol.py:225-229); the MPPI header selects the branching spline engine (dr_bm_*), the PR-MPPI
 * header the StratRace2D PR-MPPI engine (dr_prsr_create).  max_steps <= 0 selects 500 (main.cpp:11). */
typedef struct dr_session dr_session;
int dr_session_create(const void* env_frame, size_t env_len, const void* cont_frame, size_t cont_len, const void* state_frame,
                      size_t state_len, int32_t max_steps, dr_session** out);
void dr_session_destroy(dr_session* s);
/* The next frame SimulationEngine::run sends (engine.cpp:314-352): the step-0 STATE, then per control
 * step getControl + dynStep + STATE, an EVENT on termination or at max_steps, then DONE.  *frame stays
 * valid until the next call on s; after DONE, *frame = NULL, *len = 0 and *msg_type = -1. */
int dr_session_next(dr_session* s, const uint8_t** frame, size_t* len, int32_t* msg_type);

/* Single-controller entry points with the reference's shape. */
typedef struct dr_mppi_cfg
{
    int32_t n_samples;
    int32_t n_timesteps;
    float lambda;
    float sigma[DR_ACTION_DIM];
    float shift_fill[DR_ACTION_DIM];
    int32_t adapt_lambda;
    int32_t max_opponents; /* <= DR_MAX_WORLD_DRONES - 1 */
} dr_mppi_cfg;

typedef struct dr_mppi dr_mppi;

int dr_mppi_create(const dr_quad_params* params, const dr_cost_cfg* cost, const dr_track* track, const dr_mppi_cfg* cfg,
                   uint64_t seed, dr_mppi** out);
void dr_mppi_destroy(dr_mppi* h);
/* One control step from `state` with the given opponents; writes the first
 * action of the updated nominal, then shifts the nominal. */
int dr_mppi_step(dr_mppi* h, const dr_drone_state* state, const dr_opponent* opponents, int n_opponents, float* out_action);
int dr_mppi_get_nominal(dr_mppi* h, float* nominal, int n_floats);
int dr_mppi_set_nominal(dr_mppi* h, const float* nominal, int n_floats);
int dr_mppi_get_diag(dr_mppi* h, dr_diag* diag);
dr_engine* dr_mppi_engine(dr_mppi* h);

#ifdef __cplusplus
}
#endif

#endif /* DR_API_H */
