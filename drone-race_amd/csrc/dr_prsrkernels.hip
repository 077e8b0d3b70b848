// PR-MPPI kernels on the StratRace2D environment (the reference's PR-MPPI
// as it runs on its compiled environment).  One step of a batch:
//   k_prsr_rollout -> k_pr_agg -> k_pr_wavg -> k_prsr_update -> k_prsr_advance
// (aggregation and weighted averages are the environment-independent
// kernels of dr_prkernels.hip; plans and noise rows are 4 wide with the two
// StratRace2D control dimensions in lanes x, y).
//
// Reference kernels these replace (src/prmppi_kernels.cu):
//   PRMPPIgenerateNoiseKernel :8-20, PRMPPIsampleThetaValues :23-51,
//   PRMPPIfullRolloutKernel :55-130                        -> k_prsr_rollout
//   PRMPPIcomputeCandidateCostKernel :285-347, PRMPPIcomputeSafeCostKernel
//   :350-393, host choice prmppi_controller.cu:227-346     -> k_prsr_update
// The safety of the chosen plan is read from the candidate rollouts (same
// environment noise); the reference rolls it out again with fresh draws.
#include <hip/hip_runtime.h>

#include "dr_prcommon.h"
#include "dr_sr.h"

namespace dr {

using namespace sr;

// Environment-noise counter word of rollout (model m, sample k) and of
// candidate lane l (above any sample index: K * M < 2^28).
__device__ __forceinline__ uint32_t sample_stream(int m, int K, int k)
{
    return (uint32_t)m * (uint32_t)K + (uint32_t)k;
}
__device__ __forceinline__ uint32_t cand_stream(int lane)
{
    return 0xF0000000u + (uint32_t)lane;
}

__device__ __forceinline__ void stage_grids(const dr_sr_env& E, float* s_p, float* s_t, float* s_k)
{
    for (int i = threadIdx.x; i < NS * 2; i += blockDim.x)
    {
        s_p[i] = (&E.p_grid[0][0])[i];
        s_t[i] = (&E.t_grid[0][0])[i];
    }
    for (int i = threadIdx.x; i < NS; i += blockDim.x)
        s_k[i] = E.kappa_grid[i];
}

// W waves per 64 samples of one instance (wave w: plan (pair0 + w) / M,
// model (pair0 + w) % M); the control noise of a chunk of PR_CH steps is
// made once into an LDS ring and read by every rollout of the sample
// (common random numbers, prmppi_kernels.cu:107); each (model, sample) has
// its own environment-noise stream, shared by its two plans.
__global__ __launch_bounds__(1024) void k_prsr_rollout(const PrInstDev* __restrict__ insts, const PrDyn* __restrict__ dyn,
                                                       const PrTile* __restrict__ tiles, int n_tiles, const dr_sr_env* __restrict__ env,
                                                       dr_sr_cost KC, const dr_sr_state* __restrict__ states, const float* __restrict__ work,
                                                       float2* __restrict__ pair, float4* __restrict__ noise)
{
    __shared__ float4 s_eps[2][PR_CH][WAVE];
    __shared__ float s_nom[2 * DR_MAX_T * 4];
    __shared__ float s_p[NS * 2];
    __shared__ float s_t[NS * 2];
    __shared__ float s_k[NS];

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1), wave = tid / WAVE;
    const int W = (int)blockDim.x / WAVE;
    const dr_sr_env& E = *env;
    stage_grids(E, s_p, s_t, s_k);
    const Tables tb{s_p, s_t, s_k, E.track_length};

    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    {
        const PrTile pt = tiles[tile];
        const PrInstDev& I = insts[pt.inst];
        const PrDyn& Y = dyn[pt.inst];
        const int K = I.K, T = I.T, M = I.M;
        const uint32_t step = Y.step;
        const uint32_t key0 = I.key0, key1 = I.key1;
        const float sigma = I.sigma[0];

        __syncthreads(); // grids staged; plans of the previous tile no longer read
        for (int i = tid; i < 2 * T * 4; i += blockDim.x)
            s_nom[i] = work[I.plan_off + i];
        __syncthreads();

        const bool rolls = wave < pt.npairs;
        const int pidx = pt.pair0 + wave;
        const int m = rolls ? pidx % M : 0;
        const int theta = rolls ? model_branch(I, Y, step, m) : 0;
        const int k = pt.k0 + lane;
        const bool live = k < K;
        const bool write_noise = pt.pair0 == 0;
        const int n_chunks = (T + PR_CH - 1) / PR_CH;
        float4* nrow = noise + I.noise_off + (long long)k * T;
        const float* nomp = s_nom + (rolls ? (pidx / M) : 0) * T * 4;

        St s;
        float acc = 0.0f, decay = 1.0f, smax = -__builtin_inff();
        bool stop = false;
        if (rolls && live)
            s = state_from_abi(states[I.drone]);

        for (int i = 0; i <= n_chunks; ++i)
        {
            if (i < n_chunks && live)
            {
                float4* ring = &s_eps[i & 1][0][0];
                for (int u = wave; u < PR_CH; u += W)
                {
                    const int t = i * PR_CH + u;
                    if (t < T)
                    {
                        const U4 w = philox4x32_10(U4{(uint32_t)k, (uint32_t)t, step, STREAM_CONTROL}, key0, key1);
                        float z0, z1;
                        box_muller(w.x, w.y, &z0, &z1);
                        const float4 e = make_float4(sigma * z0, sigma * z1, 0.0f, 0.0f);
                        ring[u * WAVE + lane] = e;
                        if (write_noise)
                            nrow[t] = e;
                    }
                }
            }
            if (i > 0 && rolls && live)
            {
                const int j = i - 1;
                const float4* ring = &s_eps[j & 1][0][0];
                for (int u = 0; u < PR_CH; ++u)
                {
                    const int t = j * PR_CH + u;
                    if (t < T && !stop)
                    {
                        const float4 e = ring[u * WAVE + lane];
                        const float uin[2] = {nomp[t * 4 + 0] + e.x, nomp[t * 4 + 1] + e.y};
                        NormalSeq rng = normal_seq(key0, key1, sample_stream(m, K, k), (uint32_t)t, step);
                        const int term = env_step(tb, E, theta, uin, s, rng, true);
                        acc = fma_(pr_state_cost(E, KC, s), decay, acc);
                        smax = fmaxf(smax, pr_safety(E, KC, s));
                        decay = decay * DECAY;
                        stop = term != SR_NONE;
                    }
                }
            }
            __syncthreads();
        }
        if (rolls && live)
        {
            acc = acc + final_cost(E, KC, s, tb.L);
            pair[I.pair_off + (long long)pidx * K + k] = make_float2(acc, smax);
        }
    }
}

// Candidates, their rollouts under every model (lane which*M + m, with its
// own environment-noise stream), the choice and the next plans.
__global__ __launch_bounds__(256) void k_prsr_update(const PrInstDev* __restrict__ insts, PrDyn* __restrict__ dyn,
                                                     const double* __restrict__ part, const double* __restrict__ den,
                                                     float* __restrict__ work, float* __restrict__ plan, float* __restrict__ cand,
                                                     int* __restrict__ models, float* __restrict__ actions, const dr_sr_env* __restrict__ env,
                                                     dr_sr_cost KC, const dr_sr_state* __restrict__ states)
{
    __shared__ PrUpdShared sh;
    __shared__ float s_p[NS * 2];
    __shared__ float s_t[NS * 2];
    __shared__ float s_k[NS];

    const int inst = blockIdx.x;
    const PrInstDev& I = insts[inst];
    PrDyn& Y = dyn[inst];
    const int T = I.T, M = I.M, tid = threadIdx.x;
    const uint32_t step = Y.step;
    const dr_sr_env& E = *env;

    pr_begin(I, Y, step, den, sh, tid);
    stage_grids(E, s_p, s_t, s_k);
    __syncthreads();
    pr_candidates(I, part, work, sh, tid);
    __syncthreads();

    if (tid < 2 * M)
    {
        const Tables tb{s_p, s_t, s_k, E.track_length};
        const int which = tid / M, m = tid - which * M;
        const int theta = sh.br[m];
        const float* U = sh.c[which];
        St s = state_from_abi(states[I.drone]);
        float acc = 0.0f, decay = 1.0f, smax = -__builtin_inff();
        for (int t = 0; t < T; ++t)
        {
            const float uin[2] = {U[t * 4] + 0.0f, U[t * 4 + 1] + 0.0f};
            NormalSeq rng = normal_seq(I.key0, I.key1, cand_stream(tid), (uint32_t)t, step);
            const int term = env_step(tb, E, theta, uin, s, rng, true);
            acc = fma_(pr_state_cost(E, KC, s), decay, acc);
            smax = fmaxf(smax, pr_safety(E, KC, s));
            decay = decay * DECAY;
            if (term != SR_NONE)
                break;
        }
        acc = acc + final_cost(E, KC, s, tb.L);
        sh.cc[tid] = smax > 0.0f ? acc + I.safety_weight : acc;
        sh.cs[tid] = smax;
    }
    __syncthreads();
    pr_finish(I, Y, step, sh, work, plan, cand, models, actions, inst, tid);
}

// Closed loop of each instance's world (SimulationEngine::dynStep with
// loseOnOppWin = false, environment noise at counter (2^32 - 1, 4 n + b, 0,
// STREAM_ENV) for control step n); a world stops at its first event.
__global__ __launch_bounds__(64) void k_prsr_advance(const PrInstDev* __restrict__ insts, PrDyn* __restrict__ dyn,
                                                     const dr_sr_env* __restrict__ env, dr_sr_state* __restrict__ states,
                                                     const float* __restrict__ actions, int* __restrict__ events, int n_inst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst || events[i] != 0)
        return;
    const dr_sr_env& E = *env;
    const Tables tb{&E.p_grid[0][0], &E.t_grid[0][0], E.kappa_grid, E.track_length};
    const PrInstDev& I = insts[i];
    St s = state_from_abi(states[I.drone]);
    NormalSeq rng = normal_seq(I.key0, I.key1, 0xFFFFFFFFu, dyn[i].step - 1u, 0u);
    const float u[2] = {actions[i * 4], actions[i * 4 + 1]};
    // SimulationEngine::dynStep (engine.cpp:286-312): the controller's belief
    // is updated from the world opponent's applied action
    float belief[DR_SR_MAX_MODELS];
    for (int th = 0; th < DR_SR_MAX_MODELS; ++th)
        belief[th] = dyn[i].belief[th];
    const int term = env_step_t<true>(tb, E, I.world_theta, u, s, rng, false, belief);
    state_to_abi(s, states[I.drone]);
    for (int th = 0; th < DR_SR_MAX_MODELS; ++th)
        if (th < E.n_models)
            dyn[i].belief[th] = belief[th];
    events[i] = term;
}

// ---------------------------------------------------------------------------
// launch wrappers (host)

hipError_t launch_prsr_rollout(int waves, int grid, hipStream_t st, const PrInstDev* insts, const PrDyn* dyn, const PrTile* tiles, int n_tiles,
                               const dr_sr_env* env, const dr_sr_cost& KC, const dr_sr_state* states, const float* work, float2* pair,
                               float4* noise)
{
    if (waves < 1 || waves > PR_MAX_WAVES)
        return hipErrorInvalidValue;
    const int g = grid > 0 ? (grid < n_tiles ? grid : n_tiles) : n_tiles;
    k_prsr_rollout<<<g, waves * WAVE, 0, st>>>(insts, dyn, tiles, n_tiles, env, KC, states, work, pair, noise);
    return hipGetLastError();
}

hipError_t launch_prsr_update(int n_inst, hipStream_t st, const PrInstDev* insts, PrDyn* dyn, const double* part, const double* den,
                              float* work, float* plan, float* cand, int* models, float* actions, const dr_sr_env* env, const dr_sr_cost& KC,
                              const dr_sr_state* states)
{
    k_prsr_update<<<n_inst, 256, 0, st>>>(insts, dyn, part, den, work, plan, cand, models, actions, env, KC, states);
    return hipGetLastError();
}

hipError_t launch_prsr_advance(int n_inst, hipStream_t st, const PrInstDev* insts, PrDyn* dyn, const dr_sr_env* env, dr_sr_state* states,
                               const float* actions, int* events)
{
    k_prsr_advance<<<(n_inst + 63) / 64, 64, 0, st>>>(insts, dyn, env, states, actions, events, n_inst);
    return hipGetLastError();
}

} // namespace dr
