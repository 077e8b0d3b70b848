// Plain-MPPI kernels of the StratRace2D environment for gfx950 (wave64).
// One control step of a batch:
//   k_sr_rollout -> k_wavg (shared with the quadrotor engine) -> k_sr_update -> k_sr_advance
//
// Reference kernels these replace (src/kernels.cu, dense, one committed model):
//   generateNoiseKernel :18-31 + fullRolloutKernel :34-171  -> k_sr_rollout
//   weightedAverageKernelUnified :307-408                  -> k_wavg + k_sr_update
//   clampNominalKernel :434-456, host shift mppi_controller.cu:430-436 -> k_sr_update
//   SimulationEngine::dynStep (engine.cpp:286-312)         -> k_sr_advance
#include <hip/hip_runtime.h>

#include "dr_kernels.h"
#include "dr_sr.h"
#include "dr_srkernels.h"
#include "dr_wave.h"

namespace dr {

using namespace sr;

// One lane = one sample.  Per step: Philox control noise (first Box-Muller
// pair), float4 store (eps0, eps1, 0, 0) into the sample's noise row (the
// layout the shared weighted-average kernel reads), action = nominal + eps,
// environment step with its own noise stream, running cost with discount;
// terminal cost at the stop state.  The track grids (10 KB) and the nominal
// sit in LDS.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sr_rollout(const InstDev* __restrict__ insts, const InstDyn* __restrict__ dyn,
                                                      const SrInstDev* __restrict__ sinst, const RTile* __restrict__ tiles, int n_tiles,
                                                      const dr_sr_env* __restrict__ env, dr_sr_cost KC, const dr_sr_state* __restrict__ states,
                                                      const float* __restrict__ nominal, float* __restrict__ costs, float4* __restrict__ noise,
                                                      float* __restrict__ tmin)
{
    __shared__ float s_p[NS * 2];
    __shared__ float s_t[NS * 2];
    __shared__ float s_k[NS];
    __shared__ float s_nom[DR_MAX_T * 2];
    __shared__ float s_min[BLOCK / WAVE];

    const int tid = threadIdx.x;
    const dr_sr_env& E = *env;
    for (int i = tid; i < NS * 2; i += BLOCK)
    {
        s_p[i] = (&E.p_grid[0][0])[i];
        s_t[i] = (&E.t_grid[0][0])[i];
    }
    for (int i = tid; i < NS; i += BLOCK)
        s_k[i] = E.kappa_grid[i];
    const Tables tb{s_p, s_t, s_k, E.track_length};

    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    {
        const RTile rt = tiles[tile];
        const InstDev& I = insts[rt.inst];
        const SrInstDev& SI = sinst[rt.inst];
        const int K = I.K, T = I.T;
        const uint32_t step = dyn[rt.inst].step;
        const uint32_t key0 = I.key0, key1 = I.key1;

        __syncthreads(); // grids staged; nominal of the previous tile no longer read
        for (int i = tid; i < T * 2; i += BLOCK)
            s_nom[i] = nominal[I.nom_off + i];
        __syncthreads();

        const int k = rt.k0 + tid;
        float cost = __builtin_inff();
        if (k < K)
        {
            St s = state_from_abi(states[rt.inst]);
            float acc = 0.0f, decay = 1.0f;
            bool stop = false;
            float4* nrow = noise + I.noise_off + (long long)k * T;
            for (int t = 0; t < T; ++t)
            {
                const U4 w = philox4x32_10(U4{(uint32_t)k, (uint32_t)t, step, STREAM_CONTROL}, key0, key1);
                float z0, z1;
                box_muller(w.x, w.y, &z0, &z1);
                const float e0 = SI.sigma * z0, e1 = SI.sigma * z1;
                nrow[t] = make_float4(e0, e1, 0.0f, 0.0f);
                if (!stop)
                {
                    const float u[2] = {s_nom[t * 2] + e0, s_nom[t * 2 + 1] + e1};
                    NormalSeq rng = normal_seq(key0, key1, (uint32_t)k, (uint32_t)t, step);
                    const int term = env_step(tb, E, SI.theta, u, s, rng, true);
                    acc = fma_(state_cost(E, KC, s), decay, acc);
                    decay = decay * DECAY;
                    stop = term != SR_NONE;
                }
            }
            acc = acc + final_cost(E, KC, s, tb.L);
            costs[I.cost_off + k] = acc;
            cost = acc;
        }
        const float wm = wave_min(cost);
        if ((tid & (WAVE - 1)) == 0)
            s_min[tid / WAVE] = wm;
        __syncthreads();
        if (tid == 0)
        {
            float m = s_min[0];
            for (int w = 1; w < BLOCK / WAVE; ++w)
                m = fminf(m, s_min[w]);
            tmin[tile] = m;
        }
    }
}

// Per instance: nominal += sum(w eps)/sum(w) (when sum(w) > 1e-30,
// kernels.cu:402-403), L2 clamp of every step to the ego's max acceleration
// (kernels.cu:434-456), action = first step, shift with the fill value
// (mppi_controller.cu:430-436), nu and the optional temperature update
// (mppi_controller.cu:413-424).
__global__ __launch_bounds__(256) void k_sr_update(const InstDev* __restrict__ insts, InstDyn* __restrict__ dyn,
                                                   const SrInstDev* __restrict__ sinst, const double* __restrict__ part,
                                                   const double* __restrict__ den, float* __restrict__ nominal, float* __restrict__ actions,
                                                   float max_accel)
{
    __shared__ float s_u[DR_MAX_T * 2];
    const int inst = blockIdx.x;
    const InstDev& I = insts[inst];
    const SrInstDev& SI = sinst[inst];
    const int T = I.T, tid = threadIdx.x;

    const double dn = sum_in_order(den + I.den_off, I.n_chunks, 1);
    for (int i = tid; i < T * 2; i += blockDim.x)
    {
        const int t = i >> 1, d = i & 1;
        const double num = sum_in_order(part + I.part_off + t * 4 + d, I.n_chunks, (long long)T * 4);
        float u = nominal[I.nom_off + i];
        if (dn > 1e-30)
            u = u + (float)(num / dn);
        s_u[i] = u;
    }
    __syncthreads();
    for (int t = tid; t < T; t += blockDim.x)
    {
        const float a = s_u[t * 2], b = s_u[t * 2 + 1];
        const float sq = a * a + b * b;
        if (sq > max_accel * max_accel)
        {
            const float sc = max_accel / __builtin_sqrtf(sq);
            s_u[t * 2] = a * sc;
            s_u[t * 2 + 1] = b * sc;
        }
    }
    __syncthreads();
    if (tid < 2)
        actions[inst * 2 + tid] = s_u[tid];
    for (int i = tid; i < T * 2; i += blockDim.x)
    {
        const int t = i >> 1, d = i & 1;
        nominal[I.nom_off + i] = (t + 1 < T) ? s_u[i + 2] : SI.fill[d];
    }
    if (tid == 0)
    {
        InstDyn& Y = dyn[inst];
        const float nu = (float)dn;
        Y.nu = nu;
        if (SI.adapt)
        {
            if (nu > 0.001f * (float)I.K)
                Y.lambda = Y.lambda * 0.9f;
            else if (nu < 0.0005f * (float)I.K)
                Y.lambda = Y.lambda * 1.2f;
        }
        Y.step = Y.step + 1u;
    }
}

// Closed loop (SimulationEngine::dynStep with loseOnOppWin = false): each
// world takes its instance's action with the environment's noise (Philox
// counter (2^32 - 1, 4 n + block, 0, STREAM_ENV) at control step n); a world
// stops at its first event.
__global__ __launch_bounds__(64) void k_sr_advance(const InstDev* __restrict__ insts, const InstDyn* __restrict__ dyn,
                                                   const SrInstDev* __restrict__ sinst, const dr_sr_env* __restrict__ env,
                                                   dr_sr_state* __restrict__ states, const float* __restrict__ actions,
                                                   int* __restrict__ events, int n_inst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst || events[i] != 0)
        return;
    const dr_sr_env& E = *env;
    const Tables tb{&E.p_grid[0][0], &E.t_grid[0][0], E.kappa_grid, E.track_length};
    const InstDev& I = insts[i];
    const uint32_t n = dyn[i].step - 1u;
    St s = state_from_abi(states[i]);
    NormalSeq rng = normal_seq(I.key0, I.key1, 0xFFFFFFFFu, n, 0u);
    const float u[2] = {actions[i * 2], actions[i * 2 + 1]};
    const int term = env_step(tb, E, sinst[i].theta, u, s, rng, false);
    state_to_abi(s, states[i]);
    events[i] = term;
}

// ---------------------------------------------------------------------------
// launch wrappers (host)

hipError_t launch_sr_rollout(int block, int grid, hipStream_t st, const InstDev* insts, const InstDyn* dyn, const SrInstDev* sinst,
                             const RTile* tiles, int n_tiles, const dr_sr_env* env, const dr_sr_cost& KC, const dr_sr_state* states,
                             const float* nominal, float* costs, float4* noise, float* tmin)
{
    const int g = grid > 0 ? (grid < n_tiles ? grid : n_tiles) : n_tiles;
    switch (block)
    {
    case 64:
        k_sr_rollout<64><<<g, 64, 0, st>>>(insts, dyn, sinst, tiles, n_tiles, env, KC, states, nominal, costs, noise, tmin);
        break;
    case 128:
        k_sr_rollout<128><<<g, 128, 0, st>>>(insts, dyn, sinst, tiles, n_tiles, env, KC, states, nominal, costs, noise, tmin);
        break;
    case 256:
        k_sr_rollout<256><<<g, 256, 0, st>>>(insts, dyn, sinst, tiles, n_tiles, env, KC, states, nominal, costs, noise, tmin);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sr_update(int n_inst, hipStream_t st, const InstDev* insts, InstDyn* dyn, const SrInstDev* sinst, const double* part,
                            const double* den, float* nominal, float* actions, float max_accel)
{
    k_sr_update<<<n_inst, 256, 0, st>>>(insts, dyn, sinst, part, den, nominal, actions, max_accel);
    return hipGetLastError();
}

hipError_t launch_sr_advance(int n_inst, hipStream_t st, const InstDev* insts, const InstDyn* dyn, const SrInstDev* sinst,
                             const dr_sr_env* env, dr_sr_state* states, const float* actions, int* events)
{
    k_sr_advance<<<(n_inst + 63) / 64, 64, 0, st>>>(insts, dyn, sinst, env, states, actions, events, n_inst);
    return hipGetLastError();
}

} // namespace dr
