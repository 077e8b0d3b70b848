// Wave64 reductions and the fixed-order partial sums shared by the MPPI and
// PR-MPPI reduction kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace dr {

constexpr int WAVE_LANES = 64;

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
        v = fminf(v, __shfl_xor(v, m, WAVE_LANES));
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
        v += __shfl_xor(v, m, WAVE_LANES);
    return v;
}

// Sum of n values p[0], p[stride], ... in index order; loads are issued in
// batches of 16 so that their latencies overlap (same order as a plain loop).
__device__ __forceinline__ double sum_in_order(const double* __restrict__ p, int n, long long stride)
{
    double acc = 0.0;
    for (int c0 = 0; c0 < n; c0 += 16)
    {
        double v[16];
#pragma unroll
        for (int b = 0; b < 16; ++b)
            v[b] = (c0 + b < n) ? p[(long long)(c0 + b) * stride] : 0.0;
#pragma unroll
        for (int b = 0; b < 16; ++b)
            acc += v[b];
    }
    return acc;
}

} // namespace dr
