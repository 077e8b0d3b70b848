// Branching spline MPPI on StratRace2D for gfx950: the reference's default
// MPPI controller (USE_SPLINES, one model factor) without plan verification.
// One control step of a batch:
//   k_bm_rollout -> k_bm_aggregate -> k_bm_min -> k_bm_wavg -> k_bm_update
//   [-> k_bm_verify, Clopper-Pearson on the host] -> k_bm_certify -> k_bm_advance
//
// Reference kernels these replace (src/kernels.cu, src/mppi_controller.cu):
//   generateNoiseKernel :18-31 + fullRolloutKernel :34-171  -> k_bm_rollout
//   aggregateBranchCostsKernel :175-217                     -> k_bm_aggregate
//   computeMaskedMinCostsKernel :221-301                    -> k_bm_min
//   weightedAverageKernelUnified :307-408                   -> k_bm_wavg + k_bm_update
//   interpolateSplineKernel :411-431, clampNominalKernel :434-456,
//   shiftSplineKernel :461-486, host action / temperature / shift
//   mppi_controller.cu:385-457                              -> k_bm_update
//   verifyNominalFailureKernel :488-551                     -> k_bm_verify
//   plan choice, action, shift mppi_controller.cu:379-436   -> k_bm_certify
//   SimulationEngine::dynStep (engine.cpp:286-312)          -> k_bm_advance
//
// Arithmetic: sums written a += b * c in the reference are fused
// multiply-adds here (what nvcc emits for them); the weighted averages sum
// in FP64 with FP32 weights (the build's numerics contract).
#include <hip/hip_runtime.h>

#include "dr_bmkernels.h"
#include "dr_envpolicy.h"
#include "dr_wave.h"

namespace dr {

using namespace sr;

// Knot noise of sample k, plan b, knot m: sigma * (z0, z1) from the first
// Box-Muller pair of Philox (k, b * DR_MAX_KNOTS + m, step, STREAM_CONTROL).
__device__ __forceinline__ float2 knot_noise(uint32_t key0, uint32_t key1, uint32_t k, int b, int m, uint32_t step, float sigma)
{
    const U4 w = philox4x32_10(U4{k, (uint32_t)(b * DR_MAX_KNOTS + m), step, STREAM_CONTROL}, key0, key1);
    float z0, z1;
    box_muller(w.x, w.y, &z0, &z1);
    return make_float2(sigma * z0, sigma * z1);
}

// sampleContributed / splineSampleContributed (state.h:203-228), one model factor.
__device__ __forceinline__ bool sample_contributed(int b, int tLocal, int btime, int used, int T)
{
    const int origin = (b != 0 && btime > 0) ? btime : 0;
    const int tGlobal = tLocal + origin;
    if (tGlobal >= T)
        return false;
    if ((b != 0 && b != used) || (b == 0 && used != 0 && tGlobal >= btime))
        return false;
    return true;
}

__device__ __forceinline__ bool spline_contributed(int b, int m, int btime, int used, int T, int M, const int* knots)
{
    return (m > 0 && sample_contributed(b, knots[m - 1], btime, used, T)) || sample_contributed(b, knots[m], btime, used, T) ||
           (m < M - 1 && sample_contributed(b, knots[m + 1], btime, used, T));
}

// One lane = one (sample, true model) pair; the model's lanes also draw and
// store the knot noise of plans b = theta, theta + NT, ... (all plans when
// MAXM == 0), read by the weighted averages.  The lane rolls its model out if it is compatible
// with the committed branch (kernels.cu:78-161): the action at local time
// tau = t - t_origin is sum_m B[tau][m] (knot + noise) of the current
// branch, the environment step updates the belief, and a branch commits
// when the belief becomes confident (t_origin = t + 1).  With MAXM > 0
// (and M <= MAXM) the current branch's knot noise is held in registers;
// with MAXM == 0 it is read back from the noise rows in HBM; the
// environment policy picks one (P::REG_NOISE), by measurement.  With RLEN
// each lane also records how many environment steps it ran (a separate
// instantiation, so the plain loop carries nothing extra).
// The spline basis sits in dynamic LDS (T * M floats of the largest instance).
template <int MAXM, bool RLEN, class P>
__global__ __launch_bounds__(BM_BLOCK) void k_bm_rollout(const BmInstDev* __restrict__ insts, const BmDyn* __restrict__ dyn,
                                                         const BmTile* __restrict__ tiles, int n_tiles, const typename P::Env* __restrict__ env,
                                                         dr_sr_cost KC, const typename P::Abi* __restrict__ states,
                                                         const float* __restrict__ knots, const float* __restrict__ basis,
                                                         float2* __restrict__ noise, float* __restrict__ ctrue, int* __restrict__ used,
                                                         int* __restrict__ btime, int* __restrict__ rlen)
{
    __shared__ float s_env[P::LDS_FLOATS];
    extern __shared__ float s_B[];
    __shared__ float s_kn[(DR_SR_MAX_MODELS + 1) * DR_MAX_KNOTS * 2];

    const int tid = threadIdx.x;
    const typename P::Env& E = *env;
    const typename P::Ctx ctx = P::stage(s_env, E, tid, BM_BLOCK);

    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    {
        const BmTile bt = tiles[tile];
        const BmInstDev& I = insts[bt.inst];
        const BmDyn& Y = dyn[bt.inst];
        const int K = I.K, T = I.T, M = I.M, NB = I.NB, NT = I.NT;
        const uint32_t step = Y.step;
        const uint32_t key0 = I.key0, key1 = I.key1;

        __syncthreads(); // previous tile done with the LDS plans
        for (int i = tid; i < T * M; i += BM_BLOCK)
            s_B[i] = basis[I.B_off + i];
        for (int i = tid; i < NB * M * 2; i += BM_BLOCK)
            s_kn[i] = knots[I.knot_off + i];
        __syncthreads();

        const int k = bt.k0 + tid;
        const int th = bt.theta;
        if (k < K)
        {
            // MAXM == 0 lanes read their branch's noise back, so each writes every
            // plan's (the lanes of other models store the same values)
            const int b0 = MAXM > 0 ? th : 0, bstep = MAXM > 0 ? NT : 1;
            for (int b = b0; b < NB; b += bstep)
                for (int m = 0; m < M; ++m)
                    noise[I.noise_off + (long long)(b * M + m) * K + k] = knot_noise(key0, key1, (uint32_t)k, b, m, step, I.sigma);

            float belief0[DR_SR_MAX_MODELS];
            for (int q = 0; q < DR_SR_MAX_MODELS; ++q)
                belief0[q] = Y.belief[q];
            const int firstPred = find_confident(belief0, NT, I.min_conf);
            const int firstTime = firstPred == 0 ? -1 : 0;
            ctrue[I.ct_off + th * K + k] = __builtin_inff();
            used[I.ct_off + th * K + k] = firstPred;
            btime[I.ct_off + th * K + k] = firstTime;
            if constexpr (RLEN)
                rlen[I.ct_off + th * K + k] = 0;
            const typename P::State x0 = P::load(states[bt.inst]);
            if (firstPred == 0 || firstPred == th + 1)
            {
                typename P::State s = x0;
                float belief[DR_SR_MAX_MODELS];
                for (int q = 0; q < DR_SR_MAX_MODELS; ++q)
                    belief[q] = belief0[q];
                int pred = firstPred, time = firstTime, origin = 0;
                constexpr int RM = MAXM > 0 ? MAXM : 1;
                // knot + noise of the current branch, formed once per branch
                // (the same float sums the step would form)
                float2 e[RM];
                if constexpr (MAXM > 0)
                {
                    const float* kb = s_kn + pred * M * 2;
#pragma unroll
                    for (int m = 0; m < MAXM; ++m)
                        if (m < M)
                        {
                            const float2 z = knot_noise(key0, key1, (uint32_t)k, pred, m, step, I.sigma);
                            e[m] = make_float2(kb[m * 2] + z.x, kb[m * 2 + 1] + z.y);
                        }
                        else
                            e[m] = make_float2(0.0f, 0.0f);
                }
                float acc = 0.0f, decay = 1.0f;
                int t = 0;
                for (; t < T; ++t)
                {
                    const float* Brow = s_B + (t - origin) * M;
                    float u0 = 0.0f, u1 = 0.0f;
                    if constexpr (MAXM > 0)
                    {
#pragma unroll
                        for (int m = 0; m < MAXM; ++m)
                            if (m < M)
                            {
                                u0 = fma_(Brow[m], e[m].x, u0);
                                u1 = fma_(Brow[m], e[m].y, u1);
                            }
                    }
                    else
                    {
                        const float* kn = s_kn + pred * M * 2;
                        const float2* nrow = noise + I.noise_off + (long long)pred * M * K + k;
                        for (int m = 0; m < M; ++m)
                        {
                            const float2 em = nrow[(long long)m * K];
                            u0 = fma_(Brow[m], kn[m * 2] + em.x, u0);
                            u1 = fma_(Brow[m], kn[m * 2 + 1] + em.y, u1);
                        }
                    }
                    const float ego[2] = {u0, u1};
                    NormalSeq rng = normal_seq(key0, key1, (uint32_t)(th * K + k), (uint32_t)t, step);
                    // once committed, the rollout no longer reads its belief
                    const int term = P::step(ctx, E, th, ego, s, rng, true, belief, pred == 0);
                    const int np = pred == 0 ? find_confident(belief, NT, I.min_conf) : 0;
                    if (np != 0)
                    {
                        pred = np;
                        time = t + 1;
                        origin = t + 1;
                        if constexpr (MAXM > 0)
                        {
                            const float* kb = s_kn + pred * M * 2;
#pragma unroll
                            for (int m = 0; m < MAXM; ++m)
                                if (m < M)
                                {
                                    const float2 z = knot_noise(key0, key1, (uint32_t)k, pred, m, step, I.sigma);
                                    e[m] = make_float2(kb[m * 2] + z.x, kb[m * 2 + 1] + z.y);
                                }
                                else
                                    e[m] = make_float2(0.0f, 0.0f);
                        }
                    }
                    acc = fma_(P::stage_cost(ctx, E, KC, s, th), decay, acc);
                    decay = decay * DECAY;
                    if (term != SR_NONE)
                        break;
                }
                acc = acc + P::final_cost(ctx, E, KC, s);
                ctrue[I.ct_off + th * K + k] = acc;
                used[I.ct_off + th * K + k] = pred;
                btime[I.ct_off + th * K + k] = time;
                if constexpr (RLEN)
                    rlen[I.ct_off + th * K + k] = t < T ? t + 1 : T;
            }
        }
    }
}

// aggregateBranchCostsKernel: belief-weighted mean over the models
// compatible with each plan of the finite per-model costs (+inf if none).
__global__ __launch_bounds__(256) void k_bm_aggregate(const BmInstDev* __restrict__ insts, const BmDyn* __restrict__ dyn, int inst0,
                                                      const float* __restrict__ ctrue, float* __restrict__ costs)
{
    const int inst = inst0 + blockIdx.y;
    const BmInstDev& I = insts[inst];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I.NB * I.K)
        return;
    const int b = i / I.K, k = i - b * I.K;
    float num = 0.0f, den = 0.0f;
    for (int th = 0; th < I.NT; ++th)
    {
        if (b != 0 && b != th + 1)
            continue;
        const float c = ctrue[I.ct_off + th * I.K + k];
        if (!__builtin_isfinite(c))
            continue;
        const float w = dyn[inst].belief[th];
        num = fma_(w, c, num);
        den = den + w;
    }
    costs[I.cost_off + i] = den > 1e-30f ? num / den : __builtin_inff();
}

// Eligibility coefficient of sample k for (plan b, knot m): the belief mass
// of the models compatible with b whose rollout contributed to knot m.
__device__ __forceinline__ float knot_coeff(const BmInstDev& I, const float* belief, const int* __restrict__ used,
                                            const int* __restrict__ btime, int b, int m, int k)
{
    float coeff = 0.0f;
    for (int th = 0; th < I.NT; ++th)
    {
        if (belief[th] <= 0.0f)
            continue;
        if (b != 0 && b != th + 1)
            continue;
        const int u = used[I.ct_off + th * I.K + k], tm = btime[I.ct_off + th * I.K + k];
        if (spline_contributed(b, m, tm, u, I.T, I.M, I.knots))
            coeff = coeff + belief[th];
    }
    return coeff;
}

// computeMaskedMinCostsKernel: one workgroup per (instance, plan, knot,
// sample chunk); the chunk minima are combined by k_bm_wavg.
__global__ __launch_bounds__(BM_RED_BLOCK) void k_bm_min(const BmInstDev* __restrict__ insts, const BmDyn* __restrict__ dyn,
                                                         const BmRTile* __restrict__ rt, const float* __restrict__ costs,
                                                         const int* __restrict__ used, const int* __restrict__ btime, float* __restrict__ pmin)
{
    __shared__ float s_m[BM_RED_BLOCK / WAVE];
    const BmRTile r = rt[blockIdx.x];
    const BmInstDev& I = insts[r.inst];
    float belief[DR_SR_MAX_MODELS];
    for (int th = 0; th < DR_SR_MAX_MODELS; ++th)
        belief[th] = dyn[r.inst].belief[th];
    const int initPred = find_confident(belief, I.NT, I.min_conf);
    float v = __builtin_inff();
    if (initPred == 0 || initPred == r.b)
    {
        const int k0 = r.chunk * BM_CHUNK, k1 = min(I.K, k0 + BM_CHUNK);
        for (int k = k0 + threadIdx.x; k < k1; k += BM_RED_BLOCK)
            if (knot_coeff(I, belief, used, btime, r.b, r.m, k) > 0.01f)
                v = fminf(v, costs[I.cost_off + r.b * I.K + k]);
    }
    v = wave_min(v);
    if ((threadIdx.x & (WAVE - 1)) == 0)
        s_m[threadIdx.x / WAVE] = v;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        float m = s_m[0];
        for (int w = 1; w < BM_RED_BLOCK / WAVE; ++w)
            m = fminf(m, s_m[w]);
        pmin[I.pmin_off + (r.b * I.M + r.m) * I.n_chunks + r.chunk] = m;
    }
}

// weightedAverageKernelUnified: per (instance, plan, knot, sample chunk),
// the masked minimum over the chunk minima, then FP64 partial sums of w * coeff * eps (both dimensions) and w * coeff over
// the eligible samples, w = exp(-(cost - min) / lambda).
__global__ __launch_bounds__(BM_RED_BLOCK) void k_bm_wavg(const BmInstDev* __restrict__ insts, const BmDyn* __restrict__ dyn,
                                                          const BmRTile* __restrict__ rt, const float* __restrict__ costs,
                                                          const int* __restrict__ used, const int* __restrict__ btime,
                                                          const float* __restrict__ pmin, float* __restrict__ mins,
                                                          const float2* __restrict__ noise, double* __restrict__ part)
{
    __shared__ double s_a[3][BM_RED_BLOCK / WAVE];
    const BmRTile r = rt[blockIdx.x];
    const BmInstDev& I = insts[r.inst];
    const BmDyn& Y = dyn[r.inst];
    float belief[DR_SR_MAX_MODELS];
    for (int th = 0; th < DR_SR_MAX_MODELS; ++th)
        belief[th] = Y.belief[th];
    const float* pm = pmin + I.pmin_off + (r.b * I.M + r.m) * I.n_chunks;
    float minC = pm[0];
    for (int c = 1; c < I.n_chunks; ++c)
        minC = fminf(minC, pm[c]);
    if (r.chunk == 0 && threadIdx.x == 0)
        mins[I.min_off + r.b * I.M + r.m] = minC;
    const float lam = Y.lambda;
    double n0 = 0.0, n1 = 0.0, dn = 0.0;
    if (!__builtin_isinf(minC))
    {
        const int k0 = r.chunk * BM_CHUNK, k1 = min(I.K, k0 + BM_CHUNK);
        const float2* nrow = noise + I.noise_off + (long long)(r.b * I.M + r.m) * I.K;
        for (int k = k0 + threadIdx.x; k < k1; k += BM_RED_BLOCK)
        {
            const float coeff = knot_coeff(I, belief, used, btime, r.b, r.m, k);
            if (coeff > 0.01f)
            {
                const float w = expf_ieee(-(costs[I.cost_off + r.b * I.K + k] - minC) / lam);
                const double wc = (double)(w * coeff);
                const float2 e = nrow[k];
                n0 += wc * (double)e.x;
                n1 += wc * (double)e.y;
                dn += wc;
            }
        }
    }
    n0 = wave_sum(n0);
    n1 = wave_sum(n1);
    dn = wave_sum(dn);
    if ((threadIdx.x & (WAVE - 1)) == 0)
    {
        s_a[0][threadIdx.x / WAVE] = n0;
        s_a[1][threadIdx.x / WAVE] = n1;
        s_a[2][threadIdx.x / WAVE] = dn;
    }
    __syncthreads();
    if (threadIdx.x < 3)
    {
        double v = s_a[threadIdx.x][0];
        for (int w = 1; w < BM_RED_BLOCK / WAVE; ++w)
            v += s_a[threadIdx.x][w];
        part[I.part_off + ((long long)(r.b * I.M + r.m) * I.n_chunks + r.chunk) * 3 + threadIdx.x] = v;
    }
}

// Per instance: knots += num/den (den > 1e-30), this step's dense plans
// dnew = B * knots, L2 clamp of the knots to the ego's max acceleration,
// temperature update from the confident branch's nu, spline shift of that
// branch's knots (knot m <- spline at min(tau_m + 1, T - 1)); clears the
// verification counters.  The action and the executed plans are k_bm_certify's.
__global__ __launch_bounds__(256) void k_bm_update(const BmInstDev* __restrict__ insts, BmDyn* __restrict__ dyn,
                                                   const double* __restrict__ part, const float* __restrict__ basis, float* __restrict__ knots,
                                                   float* __restrict__ dnew, float* __restrict__ nu, int* __restrict__ fail, float max_accel)
{
    __shared__ float s_kn[(DR_SR_MAX_MODELS + 1) * DR_MAX_KNOTS * 2];
    __shared__ float s_sh[DR_MAX_KNOTS * 2];
    __shared__ int s_branch;
    const int inst = blockIdx.x;
    const BmInstDev& I = insts[inst];
    BmDyn& Y = dyn[inst];
    const int T = I.T, M = I.M, NB = I.NB, tid = threadIdx.x;
    const float* B = basis + I.B_off;

    for (int i = tid; i < NB * M * 2; i += blockDim.x)
    {
        const int bm = i >> 1, d = i & 1;
        const double* p = part + I.part_off + (long long)bm * I.n_chunks * 3;
        double num = 0.0, den = 0.0;
        for (int c = 0; c < I.n_chunks; ++c)
        {
            num += p[c * 3 + d];
            den += p[c * 3 + 2];
        }
        float v = knots[I.knot_off + i];
        if (den > 1e-30)
            v = v + (float)(num / den);
        s_kn[i] = v;
        if (d == 0 && bm % M == 0)
            nu[I.nu_off + bm / M] = (float)den;
    }
    if (tid == 0)
    {
        float belief[DR_SR_MAX_MODELS];
        for (int th = 0; th < DR_SR_MAX_MODELS; ++th)
            belief[th] = Y.belief[th];
        s_branch = find_confident(belief, I.NT, I.min_conf);
    }
    __syncthreads();
    // dense plans from the updated (unclamped) knots
    for (int i = tid; i < NB * T * 2; i += blockDim.x)
    {
        const int b = i / (T * 2), r = i - b * T * 2, t = r >> 1, d = r & 1;
        float sum = 0.0f;
        for (int m = 0; m < M; ++m)
            sum = fma_(B[t * M + m], s_kn[(b * M + m) * 2 + d], sum);
        dnew[I.dense_off + i] = sum;
    }
    if (I.n_verif > 0)
        for (int i = tid; i < 2 * T; i += blockDim.x)
            fail[I.fail_off + i] = 0;
    __syncthreads();
    // L2 clamp of every knot
    for (int bm = tid; bm < NB * M; bm += blockDim.x)
    {
        const float a = s_kn[bm * 2], c = s_kn[bm * 2 + 1];
        const float sq = a * a + c * c;
        if (sq > max_accel * max_accel)
        {
            const float sc = max_accel / __builtin_sqrtf(sq);
            s_kn[bm * 2] = a * sc;
            s_kn[bm * 2 + 1] = c * sc;
        }
    }
    __syncthreads();
    const int br = s_branch;
    if (tid == 0)
    {
        const float usedNu = nu[I.nu_off + br];
        if (I.adapt)
        {
            if (usedNu > 0.001f * (float)I.K)
                Y.lambda = Y.lambda * 0.9f;
            else if (usedNu < 0.0005f * (float)I.K)
                Y.lambda = Y.lambda * 1.2f;
        }
        Y.branch = br;
    }
    // spline shift of the chosen branch from the clamped knots
    for (int i = tid; i < M * 2; i += blockDim.x)
    {
        const int m = i >> 1, d = i & 1;
        int tEval = I.knots[m] + 1;
        if (tEval >= T)
            tEval = T - 1;
        float sum = 0.0f;
        for (int j = 0; j < M; ++j)
            sum = fma_(B[tEval * M + j], s_kn[(br * M + j) * 2 + d], sum);
        s_sh[i] = sum;
    }
    __syncthreads();
    for (int i = tid; i < NB * M * 2; i += blockDim.x)
    {
        const int b = i / (M * 2);
        knots[I.knot_off + i] = (b == br) ? s_sh[i - b * M * 2] : s_kn[i];
    }
}

// verifyNominalFailureKernel (kernels.cu:488-551): one lane per verification
// sample and plan set.  The opponent model is drawn from the controller's
// belief; the world runs the executed-plan rule of the controller (action of
// the current branch at local time, belief update and branching, noise on,
// the opponent's win not a loss) for verif_h steps; a loss at step t counts
// one failure at t.
template <class P>
__global__ __launch_bounds__(BM_VBLOCK) void k_bm_verify(const BmInstDev* __restrict__ insts, const BmDyn* __restrict__ dyn,
                                                         const BmVTile* __restrict__ tiles, int n_tiles,
                                                         const typename P::Env* __restrict__ env,
                                                         const typename P::Abi* __restrict__ states, const float* __restrict__ dnew,
                                                         const float* __restrict__ dense, int* __restrict__ fail)
{
    __shared__ float s_env[P::LDS_FLOATS];
    __shared__ float s_plan[(DR_SR_MAX_MODELS + 1) * DR_MAX_T * 2];
    __shared__ int s_fail[DR_MAX_T];

    const int tid = threadIdx.x;
    const typename P::Env& E = *env;
    const typename P::Ctx ctx = P::stage(s_env, E, tid, BM_VBLOCK);

    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    {
        const BmVTile vt = tiles[tile];
        const BmInstDev& I = insts[vt.inst];
        const BmDyn& Y = dyn[vt.inst];
        const int T = I.T, NB = I.NB, NT = I.NT, H = I.verif_h;
        const float* plans = (vt.set == 0 ? dnew : dense) + I.dense_off;
        __syncthreads(); // previous tile done with the LDS plans and counters
        for (int i = tid; i < NB * T * 2; i += BM_VBLOCK)
            s_plan[i] = plans[i];
        for (int i = tid; i < H; i += BM_VBLOCK)
            s_fail[i] = 0;
        __syncthreads();

        const int sidx = vt.s0 + tid;
        if (sidx < I.n_verif)
        {
            const uint32_t step = Y.step, word = (uint32_t)sidx | ((uint32_t)vt.set << 30);
            float belief[DR_SR_MAX_MODELS];
            for (int q = 0; q < DR_SR_MAX_MODELS; ++q)
                belief[q] = Y.belief[q];
            int pred = find_confident(belief, NT, I.min_conf);
            const U4 r = philox4x32_10(U4{word, 0xFFFFFFFFu, step, STREAM_VERIFY}, I.key0, I.key1);
            const int theta = model_from_belief(uniform_open0(r.x), belief, NT);
            typename P::State s = P::load(states[vt.inst]);
            int origin = 0;
            for (int t = 0; t < H; ++t)
            {
                const float* u = s_plan + (pred * T + (t - origin)) * 2;
                const float ego[2] = {u[0], u[1]};
                NormalSeq rng = normal_seq(I.key0, I.key1, word, (uint32_t)t, step, STREAM_VERIFY);
                const int term = P::step(ctx, E, theta, ego, s, rng, false, belief, pred == 0);
                if (pred == 0)
                {
                    const int np = find_confident(belief, NT, I.min_conf);
                    if (np != 0)
                    {
                        pred = np;
                        origin = t + 1;
                    }
                }
                if (term == SR_LOSE)
                {
                    atomicAdd(&s_fail[t], 1);
                    break;
                }
                if (term != SR_NONE)
                    break;
            }
        }
        __syncthreads();
        for (int i = tid; i < H; i += BM_VBLOCK)
            if (s_fail[i] != 0)
                atomicAdd(&fail[I.fail_off + vt.set * T + i], s_fail[i]);
    }
}

// The plan choice (mppi_controller.cu:379-436): the executed plans are this
// step's when verdict[i] is 1 (or no verification runs), else the previous
// step's; the action is the first step of the confident branch, whose row is
// then shifted left with a zero fill.
__global__ __launch_bounds__(256) void k_bm_certify(const BmInstDev* __restrict__ insts, BmDyn* __restrict__ dyn,
                                                    const int* __restrict__ verdict, const float* __restrict__ dnew,
                                                    float* __restrict__ dense, float* __restrict__ actions)
{
    const int inst = blockIdx.x;
    const BmInstDev& I = insts[inst];
    BmDyn& Y = dyn[inst];
    const int T = I.T, NB = I.NB, tid = threadIdx.x;
    const int use_new = I.n_verif > 0 ? verdict[inst] : 1;
    const int br = Y.branch;
    const float* src = (use_new ? dnew : dense) + I.dense_off;
    float* dst = dense + I.dense_off;
    const int n = NB * T * 2;
    float v[2 * (DR_SR_MAX_MODELS + 1) * DR_MAX_T / 256];
    int cnt = 0;
    for (int i = tid; i < n; i += 256, ++cnt)
    {
        const int b = i / (T * 2);
        v[cnt] = b != br ? src[i] : (i + 2 < (b + 1) * T * 2 ? src[i + 2] : 0.0f);
    }
    if (tid == 0)
    {
        actions[inst * 2] = src[br * T * 2];
        actions[inst * 2 + 1] = src[br * T * 2 + 1];
    }
    __syncthreads();
    cnt = 0;
    for (int i = tid; i < n; i += 256, ++cnt)
        dst[i] = v[cnt];
    if (tid == 0)
    {
        Y.use_new = use_new;
        Y.step = Y.step + 1u;
    }
}

// Closed loop (SimulationEngine::dynStep): the world's opponent follows its
// model; the controller's belief is updated from the opponent's action.
template <class P>
__global__ __launch_bounds__(64) void k_bm_advance(const BmInstDev* __restrict__ insts, BmDyn* __restrict__ dyn,
                                                   const typename P::Env* __restrict__ env, typename P::Abi* __restrict__ states,
                                                   const float* __restrict__ actions, int* __restrict__ events, int n_inst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst || events[i] != 0)
        return;
    const typename P::Env& E = *env;
    const typename P::Ctx ctx = P::direct(E);
    const BmInstDev& I = insts[i];
    BmDyn& Y = dyn[i];
    typename P::State s = P::load(states[i]);
    float belief[DR_SR_MAX_MODELS];
    for (int th = 0; th < DR_SR_MAX_MODELS; ++th)
        belief[th] = Y.belief[th];
    NormalSeq rng = normal_seq(I.key0, I.key1, 0xFFFFFFFFu, Y.step - 1u, 0u);
    const float u[2] = {actions[i * 2], actions[i * 2 + 1]};
    const int term = P::step(ctx, E, I.true_theta, u, s, rng, false, belief, true);
    P::store(s, states[i]);
    for (int th = 0; th < DR_SR_MAX_MODELS; ++th)
        Y.belief[th] = belief[th];
    events[i] = term;
}

// ---------------------------------------------------------------------------
// launch wrappers (host)

template <bool RLEN, class P>
static void rollout_t(int max_m, int g, size_t lds, hipStream_t st, const BmInstDev* insts, const BmDyn* dyn, const BmTile* tiles,
                      int n_tiles, const void* env, const dr_sr_cost& KC, const void* states, const float* knots, const float* basis,
                      float2* noise, float* ctrue, int* used, int* btime, int* rlen)
{
    const auto* E = static_cast<const typename P::Env*>(env);
    const auto* S = static_cast<const typename P::Abi*>(states);
    if constexpr (P::REG_NOISE)
    {
        if (max_m <= 20)
        {
            k_bm_rollout<20, RLEN, P><<<g, BM_BLOCK, lds, st>>>(insts, dyn, tiles, n_tiles, E, KC, S, knots, basis, noise, ctrue, used,
                                                                btime, rlen);
            return;
        }
    }
    k_bm_rollout<0, RLEN, P><<<g, BM_BLOCK, lds, st>>>(insts, dyn, tiles, n_tiles, E, KC, S, knots, basis, noise, ctrue, used, btime,
                                                       rlen);
}

template <class P>
static void rollout_p(int max_m, int g, size_t lds, hipStream_t st, const BmInstDev* insts, const BmDyn* dyn, const BmTile* tiles,
                      int n_tiles, const void* env, const dr_sr_cost& KC, const void* states, const float* knots, const float* basis,
                      float2* noise, float* ctrue, int* used, int* btime, int* rlen)
{
    if (rlen)
        rollout_t<true, P>(max_m, g, lds, st, insts, dyn, tiles, n_tiles, env, KC, states, knots, basis, noise, ctrue, used, btime, rlen);
    else
        rollout_t<false, P>(max_m, g, lds, st, insts, dyn, tiles, n_tiles, env, KC, states, knots, basis, noise, ctrue, used, btime, rlen);
}

// env_kind: BM_ENV_SR (env dr_sr_env, states dr_sr_state), BM_ENV_HO
// (dr_ho_env, dr_ho_state) or BM_ENV_DRACE (dr_drace_env, dr_drace_state);
// rlen may be NULL (no rollout lengths recorded)
hipError_t launch_bm_rollout(int env_kind, int max_m, int max_tm, int grid, hipStream_t st, const BmInstDev* insts, const BmDyn* dyn,
                             const BmTile* tiles, int n_tiles, const void* env, const dr_sr_cost& KC, const void* states,
                             const float* knots, const float* basis, float2* noise, float* ctrue, int* used, int* btime, int* rlen)
{
    const int g = grid > 0 ? (grid < n_tiles ? grid : n_tiles) : n_tiles;
    if (max_m > DR_MAX_KNOTS || max_tm > BM_MAX_TM || max_tm < 1)
        return hipErrorInvalidValue;
    const size_t lds = (size_t)max_tm * sizeof(float);
    if (env_kind == BM_ENV_SR)
        rollout_p<SrPolicy>(max_m, g, lds, st, insts, dyn, tiles, n_tiles, env, KC, states, knots, basis, noise, ctrue, used, btime, rlen);
    else if (env_kind == BM_ENV_HO)
        rollout_p<HoPolicy>(max_m, g, lds, st, insts, dyn, tiles, n_tiles, env, KC, states, knots, basis, noise, ctrue, used, btime, rlen);
    else if (env_kind == BM_ENV_DRACE)
        rollout_p<DracePolicy>(max_m, g, lds, st, insts, dyn, tiles, n_tiles, env, KC, states, knots, basis, noise, ctrue, used, btime,
                               rlen);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_bm_reduce(int n_inst, int max_nbk, int n_rtiles, hipStream_t st, const BmInstDev* insts, BmDyn* dyn, const BmRTile* rt,
                            const float* ctrue, float* costs, const int* used, const int* btime, float* pmin, float* mins,
                            const float2* noise, double* part, hipEvent_t mid)
{
    k_bm_aggregate<<<dim3((max_nbk + 255) / 256, n_inst), 256, 0, st>>>(insts, dyn, 0, ctrue, costs);
    k_bm_min<<<n_rtiles, BM_RED_BLOCK, 0, st>>>(insts, dyn, rt, costs, used, btime, pmin);
    if (mid)
    {
        hipError_t e = hipEventRecord(mid, st);
        if (e != hipSuccess)
            return e;
    }
    k_bm_wavg<<<n_rtiles, BM_RED_BLOCK, 0, st>>>(insts, dyn, rt, costs, used, btime, pmin, mins, noise, part);
    return hipGetLastError();
}

hipError_t launch_bm_update(int n_inst, hipStream_t st, const BmInstDev* insts, BmDyn* dyn, const double* part, const float* basis,
                            float* knots, float* dnew, float* nu, int* fail, float max_accel)
{
    k_bm_update<<<n_inst, 256, 0, st>>>(insts, dyn, part, basis, knots, dnew, nu, fail, max_accel);
    return hipGetLastError();
}

hipError_t launch_bm_verify(int env_kind, int n_tiles, hipStream_t st, const BmInstDev* insts, const BmDyn* dyn, const BmVTile* tiles,
                            const void* env, const void* states, const float* dnew, const float* dense, int* fail)
{
    if (n_tiles <= 0)
        return hipGetLastError();
    if (env_kind == BM_ENV_SR)
        k_bm_verify<SrPolicy><<<n_tiles, BM_VBLOCK, 0, st>>>(insts, dyn, tiles, n_tiles, static_cast<const dr_sr_env*>(env),
                                                             static_cast<const dr_sr_state*>(states), dnew, dense, fail);
    else if (env_kind == BM_ENV_HO)
        k_bm_verify<HoPolicy><<<n_tiles, BM_VBLOCK, 0, st>>>(insts, dyn, tiles, n_tiles, static_cast<const dr_ho_env*>(env),
                                                             static_cast<const dr_ho_state*>(states), dnew, dense, fail);
    else if (env_kind == BM_ENV_DRACE)
        k_bm_verify<DracePolicy><<<n_tiles, BM_VBLOCK, 0, st>>>(insts, dyn, tiles, n_tiles, static_cast<const dr_drace_env*>(env),
                                                                static_cast<const dr_drace_state*>(states), dnew, dense, fail);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_bm_certify(int n_inst, hipStream_t st, const BmInstDev* insts, BmDyn* dyn, const int* verdict, const float* dnew,
                             float* dense, float* actions)
{
    k_bm_certify<<<n_inst, 256, 0, st>>>(insts, dyn, verdict, dnew, dense, actions);
    return hipGetLastError();
}

hipError_t launch_bm_advance(int env_kind, int n_inst, hipStream_t st, const BmInstDev* insts, BmDyn* dyn, const void* env, void* states,
                             const float* actions, int* events)
{
    const int g = (n_inst + 63) / 64;
    if (env_kind == BM_ENV_SR)
        k_bm_advance<SrPolicy><<<g, 64, 0, st>>>(insts, dyn, static_cast<const dr_sr_env*>(env), static_cast<dr_sr_state*>(states), actions,
                                                 events, n_inst);
    else if (env_kind == BM_ENV_HO)
        k_bm_advance<HoPolicy><<<g, 64, 0, st>>>(insts, dyn, static_cast<const dr_ho_env*>(env), static_cast<dr_ho_state*>(states), actions,
                                                 events, n_inst);
    else if (env_kind == BM_ENV_DRACE)
        k_bm_advance<DracePolicy><<<g, 64, 0, st>>>(insts, dyn, static_cast<const dr_drace_env*>(env), static_cast<dr_drace_state*>(states),
                                                    actions, events, n_inst);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

} // namespace dr
