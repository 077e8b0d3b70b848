// PR-MPPI device pieces shared by the environments' kernels: model draws,
// the reduction of the three weighted averages into candidates, and the
// choice / temperature / plan bookkeeping of the update kernels
// (prmppi_kernels.cu:204-282, prmppi_controller.cu:227-346).
#pragma once

#include <hip/hip_runtime.h>

#include "dr_prkernels.h"
#include "dr_wave.h"

namespace dr {

__device__ __forceinline__ int model_branch(const PrInstDev& I, const PrDyn& Y, uint32_t step, int m)
{
    return I.sampled ? sample_branch(I.key0, I.key1, (uint32_t)m, step, Y.belief, I.B) : m;
}

// LDS of an update workgroup: candidates 1, 2 and the new robust plan
// ([3][T][4]), per-lane candidate costs and safeties, model branches,
// denominators, the choice.
struct PrUpdShared
{
    float c[3][DR_MAX_T * 4];
    float cc[2 * DR_MAX_MODELS], cs[2 * DR_MAX_MODELS];
    int br[DR_MAX_MODELS];
    double dn[3];
    int pick, use;
};

// Denominators and model branches (before the first barrier).
__device__ __forceinline__ void pr_begin(const PrInstDev& I, const PrDyn& Y, uint32_t step, const double* __restrict__ den, PrUpdShared& sh,
                                         int tid)
{
    if (tid < 3)
        sh.dn[tid] = sum_in_order(den + I.den_off + tid, I.n_chunks, 3);
    if (tid < I.M)
        sh.br[tid] = model_branch(I, Y, step, tid);
}

// The three candidates from the FP64 partials (after a barrier that
// publishes sh.dn).
__device__ __forceinline__ void pr_candidates(const PrInstDev& I, const double* __restrict__ part, const float* __restrict__ work,
                                              PrUpdShared& sh, int tid)
{
    // each lane sums 4 values at a time over the chunks in order, 8 chunks of
    // loads in flight together
    const int T = I.T, n = 3 * T * 4, nc = I.n_chunks;
    const int stride = (int)blockDim.x;
    const double* pp = part + I.part_off;
    for (int i0 = tid; i0 < n; i0 += 4 * stride)
    {
        double num[4] = {0.0, 0.0, 0.0, 0.0};
        for (int c0 = 0; c0 < nc; c0 += 8)
        {
            double v[4][8];
#pragma unroll
            for (int b = 0; b < 8; ++b)
#pragma unroll
                for (int k = 0; k < 4; ++k)
                {
                    const int i = i0 + k * stride;
                    v[k][b] = (c0 + b < nc && i < n) ? pp[(long long)(c0 + b) * n + i] : 0.0;
                }
#pragma unroll
            for (int b = 0; b < 8; ++b)
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    num[k] += v[k][b];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            const int i = i0 + k * stride;
            if (i >= n)
                break;
            const int j = i / (T * 4), r = i - j * T * 4;
            float u = work[I.plan_off + (j == 0 ? 0 : T * 4) + r];
            if (sh.dn[j] > 1e-30)
                u = (float)(num[k] / sh.dn[j]);
            sh.c[j][r] = u;
        }
    }
}

// Choice, flags, temperatures, action and the next plans (after a barrier
// that publishes sh.cc / sh.cs).
__device__ __forceinline__ void pr_finish(const PrInstDev& I, PrDyn& Y, uint32_t step, PrUpdShared& sh, float* __restrict__ work,
                                          float* __restrict__ plan, float* __restrict__ cand, int* __restrict__ models,
                                          float* __restrict__ actions, int inst, int tid)
{
    const int T = I.T, M = I.M;
    if (tid == 0)
    {
        float c[2];
        for (int which = 0; which < 2; ++which)
        {
            float a = 0.0f;
            if (I.sampled)
            {
                for (int m = 0; m < M; ++m)
                    a = a + sh.cc[which * M + m];
                a = a / (float)M;
            }
            else
                for (int m = 0; m < M; ++m)
                    a = fma_(Y.belief[m], sh.cc[which * M + m], a);
            c[which] = a;
        }
        const int pick = c[0] <= c[1] ? 0 : 1;
        int use = 1;
        for (int m = 0; m < M; ++m)
            if (sh.cs[pick * M + m] > 0.0f)
                use = 0;
        Y.cand_cost[0] = c[0];
        Y.cand_cost[1] = c[1];
        Y.use_nom = use;
        Y.reset_nom = c[1] < c[0] ? 1 : 0;
        for (int j = 0; j < 3; ++j)
        {
            const float nu = (float)sh.dn[j];
            Y.nu[j] = nu;
            if (I.adapt)
            {
                float l = Y.lambda[j];
                if (nu > 0.001f * (float)I.K && l > 0.01f)
                    l = l * 0.9f;
                else if (nu < 0.0005f * (float)I.K && l < 100.0f)
                    l = l * 1.2f;
                Y.lambda[j] = l;
            }
        }
        Y.step = step + 1u;
        sh.pick = pick;
        sh.use = use;
    }
    __syncthreads();
    const float* chosen = sh.c[sh.pick];
    if (tid < 4)
        actions[inst * 4 + tid] = sh.use ? chosen[tid] : sh.c[2][tid];
    if (tid < M)
        models[I.model_off + tid] = sh.br[tid];
    const int off = I.plan_off;
    for (int i = tid; i < T * 4; i += blockDim.x)
    {
        const int t = i >> 2, d = i & 3;
        plan[off + i] = chosen[i];
        plan[off + T * 4 + i] = sh.c[2][i];
        work[off + i] = (t + 1 < T) ? chosen[i + 4] : I.shift_fill[d];
        work[off + T * 4 + i] = (t + 1 < T) ? sh.c[2][i + 4] : I.shift_fill[d];
        cand[off + i] = sh.c[0][i];
        cand[off + T * 4 + i] = sh.c[1][i];
    }
}

} // namespace dr
