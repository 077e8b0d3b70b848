// Device-side data of the branching spline MPPI engine on StratRace2D.
//
// Per instance (NT = the environment's models, NB = NT + 1 branch plans,
// M knots, K samples, T steps):
//   knots  float[NB][M][2]      spline knots of every branch plan
//   dense  float[NB][T][2]      interpolated plans (output)
//   B      float[T][M]          natural cubic spline basis
//   noise  float2[NB][M][K]     knot noise, sample-contiguous per (plan, knot)
//   ctrue  float[NT][K]         cost of each sample under each model (+inf: skipped)
//   used   int[NT][K]           final branch (0 nominal, theta + 1 specialised)
//   btime  int[NT][K]           branching time (-1: never)
//   costs  float[NB][K]         belief-weighted branch costs
//   mins   float[NB][M]         masked minima per (plan, knot)
//   pmin   float[NB][M][chunks] masked minima per sample chunk
//   dnew   float[NB][T][2]      interpolated plans of this step (before verification)
//   fail   int[2][T]            verification failures per step: new plan, previous plan
//   dense (above) holds the executed plans: the previous step's, then this step's choice
//   part   double[NB][M][chunks][3]  weighted-average partials (num x, num y, den)
#pragma once

#include <stdint.h>

#include "dr_kernels.h"

namespace dr {

constexpr int BM_BLOCK = 128;        // rollout lanes per workgroup
constexpr int BM_MAX_TM = 8192;      // T * M held in LDS
constexpr int BM_RED_BLOCK = 256;    // workgroup of the masked minima and weighted averages
constexpr int BM_CHUNK = 4096;       // samples per weighted-average workgroup
constexpr int BM_VBLOCK = 128;       // verification lanes per workgroup
constexpr int BM_ENV_SR = 0;         // StratRace2D (dr_sr_env, dr_sr_state)
constexpr int BM_ENV_HO = 1;         // HiddenObs (dr_ho_env, dr_ho_state)
constexpr int BM_ENV_DRACE = 2;      // DroneRace (dr_drace_env, dr_drace_state)

struct BmInstDev
{
    int K, T, M, NB, NT;
    int knots[DR_MAX_KNOTS];
    uint32_t key0, key1;
    float sigma, min_conf;
    int true_theta, adapt;
    long long noise_off; // float2 units
    int ct_off;          // NT * K
    int cost_off;        // NB * K
    int min_off;         // NB * M
    int knot_off;        // NB * M * 2
    int dense_off;       // NB * T * 2
    int B_off;           // T * M
    int n_chunks;
    long long part_off;  // double units, NB * M * n_chunks * 3
    int pmin_off;        // NB * M * n_chunks
    int nu_off;          // NB
    int n_verif;         // verification samples (0: none)
    int verif_h;         // verification horizon
    int fail_off;        // 2 * T
};

struct BmDyn
{
    uint32_t step;
    float lambda;
    float belief[DR_SR_MAX_MODELS];
    int branch;  // branch index of the last action
    int use_new; // the last step executed the new plan
    float _pad;
};

// Verification tile: BM_VBLOCK samples of one instance against one plan set
// (0 = this step's plans, 1 = the previous step's).
struct BmVTile
{
    int inst, set, s0, _pad;
};

// Rollout tile: BM_BLOCK consecutive samples of one instance under one true model.
struct BmTile
{
    int inst, k0, theta, _pad;
};

struct BmRTile
{
    int inst, b, m, chunk;
};

} // namespace dr
