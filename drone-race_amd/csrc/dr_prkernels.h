// Device-side data layout of the PR-MPPI engine.
//
// A PR-MPPI controller evaluates each of its K noise samples under M
// opponent models for two plans (nominal and robust), 2*M rollouts per
// sample sharing one noise sequence (common random numbers).  HBM layout of
// one batch (all instances):
//   work    float[sum_i 2*T_i*4]    plans as used this step (shifted), [plan][t][d]
//   plan    float[sum_i 2*T_i*4]    plans after the last step (unshifted), [plan][t][d]
//   cand    float[sum_i 2*T_i*4]    candidates 1 and 2 of the last step
//   noise   float4[sum_i K_i*T_i]   [k][t], written once per step
//   pair    float2[sum_i 2*M_i*K_i] (cost, max safety) per rollout, [plan][m][k]
//   agg     float4[sum_i K_i]       (full nom, max safety nom, full rob, max safety rob)
//   ablk    float4[n_agg_blocks]    minima of full nom, full rob, safety rob per block
//   part    double[sum_i chunks_i*3*T_i*4], den double[sum_i chunks_i*3]
//   models  int[sum_i M_i]          branch of each model this step
#pragma once

#include <stdint.h>

#include "dr_kernels.h"

namespace dr {

constexpr int PR_CH = 8;             // noise steps per LDS chunk of the rollout
constexpr int PR_MAX_WAVES = 16;     // waves (rollouts of one sample group) per rollout workgroup
constexpr int PR_AGG_BLOCK = 256;    // samples per aggregation workgroup
constexpr int PR_WAVG_BLOCK = 256;
constexpr int PR_WAVG_MAX_CHUNK = 2048;

struct PrInstDev
{
    int K, T, track, drone;
    uint32_t key0, key1;
    float sigma[4];
    float shift_fill[4];
    int adapt;
    int n_opp;
    int opp[MAX_OPP];
    int M, B, sampled;
    float safety_weight;
    float shift[DR_MAX_BRANCHES][MAX_OPP][3]; // velocity shift of branch b for opponent slot j
    long long noise_off;                      // float4 units
    long long pair_off;                       // float2 units
    int agg_off;                              // float4 units
    int plan_off;                             // float units, 2*T*4 per instance (work, plan, cand)
    int model_off;
    int ablk_first, ablk_count;
    long long part_off; // double units
    int den_off;
    int n_chunks;
    int rtile_first, rtile_count;
    int world_theta; // StratRace2D: opponent model of the instance's world
};

struct PrDyn
{
    uint32_t step;
    float lambda[3];   // nominal/full, robust/full, robust/safety
    float min_cost[3];
    float nu[3];
    float cand_cost[2];
    int use_nom, reset_nom;
    float belief[DR_MAX_BRANCHES];
};

struct PrTile
{
    int inst, k0, pair0, npairs;
};

struct PrATile
{
    int inst, k0;
};

} // namespace dr
