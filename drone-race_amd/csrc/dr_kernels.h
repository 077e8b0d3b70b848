// Device-side data layout shared by the kernels and the host engine.
//
// HBM layout of one batch (all instances of all worlds):
//   nominal  float[sum_i T_i*4]            per instance, [t][d]
//   costs    float[sum_i K_i]              per instance, [k]
//   noise    float4[sum_i K_i*T_i]         per instance, [k][t] (one float4 = the 4 control dims)
//   tmin     float[n_rollout_tiles]        per-rollout-workgroup minimum cost
//   part     double[sum_i chunks_i*T_i*4]  weighted-average partial numerators, [chunk][t][d]
//   den      double[sum_i chunks_i]        weighted-average partial denominators
//   drones   dr_drone_state[n_drones]      world states (ego and opponents)
// The noise is written once by the rollout (each lane its own sample row) and
// read once by the weighted average (lanes along t: each row is one coalesced
// read); no per-sample RNG state exists.
#pragma once

#include <stdint.h>

#include "dr_quad.h"

namespace dr {

constexpr int MAX_OPP = DR_MAX_WORLD_DRONES - 1;
constexpr int WAVE = 64;
constexpr int WAVG_BLOCK = 256;
constexpr int WAVG_MAX_CHUNK = 4096; // max samples per weighted-average workgroup

struct InstDev
{
    int K, T, track, drone;
    uint32_t key0, key1;
    float sigma[4];
    float shift_fill[4];
    int adapt;
    int n_opp;
    int opp[MAX_OPP];
    long long noise_off; // float4 units
    int cost_off;
    int nom_off;
    int rtile_first, rtile_count; // rollout tiles, contiguous in tile order
    long long part_off;           // double units
    int den_off;
    int n_chunks; // weighted-average sample chunks
};

struct InstDyn
{
    uint32_t step;
    float lambda;
    float min_cost;
    float nu;
    float traj_cost;
    float exec_cost;          // running cost of the executed closed-loop states, summed (k_exec)
    float exec_progress;      // race progress of the last executed state (k_exec)
    uint32_t exec_collisions; // executed states in collision (k_exec)
    uint32_t step_ahead;      // step + 1, written by k_wavg: the step whose noise k_refine's spare workgroups make
};

struct RTile
{
    int inst;
    int k0;
};

struct WTile
{
    int inst;
    int chunk;
    int t0, t1;
};

} // namespace dr
