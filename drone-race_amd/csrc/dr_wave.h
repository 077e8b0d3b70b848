// Wave64 reductions and the fixed-order partial sums shared by the MPPI and
// PR-MPPI reduction kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace dr {

constexpr int WAVE_LANES = 64;

// fminf(v, the v of the lane that the DPP control CTRL selects); lanes of the
// rows that ROWS leaves out take min(v, +inf) = v.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_min(float v)
{
    const int inf = __float_as_int(__builtin_inff());
    return fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(inf, __float_as_int(v), CTRL, ROWS, 0xf, false)));
}

// Minimum over the wave, the same value in every lane, of full waves only:
// the pairing tree of wave_sum below (the butterfly's), read out of lane 63.
__device__ __forceinline__ float wave_min(float v)
{
    v = dpp_min<0xB1, 0xf>(v);  // quad_perm [1, 0, 3, 2]
    v = dpp_min<0x4E, 0xf>(v);  // quad_perm [2, 3, 0, 1]
    v = dpp_min<0x141, 0xf>(v); // row_half_mirror
    v = dpp_min<0x140, 0xf>(v); // row_mirror
    v = dpp_min<0x142, 0xa>(v); // row_bcast:15 into rows 1 and 3
    v = dpp_min<0x143, 0xc>(v); // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), WAVE_LANES - 1));
}

// v + the v of the lane that the DPP control CTRL selects (both 32-bit halves
// moved); lanes of the rows that ROWS leaves out add +0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_add(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
    return v + __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over the wave, the same value in every lane, of full waves only.  The
// pairing tree of a butterfly over lane masks 1, 2, 4, ..., 32 (each addition
// takes the same two operands, so the same bits), in DPP instead of LDS
// permutes: pairs within quads, then the other half of each 8 and of each 16
// lanes (mirrors: after the levels before, every lane of a half holds the
// same value), then rows 0 + 1 and 2 + 3 in rows 1 and 3 (row broadcast of
// lanes 15), then all four rows in row 3 (lane 31); lane 63's total is read
// out for every lane.
__device__ __forceinline__ double wave_sum(double v)
{
    v = dpp_add<0xB1, 0xf>(v);  // quad_perm [1, 0, 3, 2]
    v = dpp_add<0x4E, 0xf>(v);  // quad_perm [2, 3, 0, 1]
    v = dpp_add<0x141, 0xf>(v); // row_half_mirror
    v = dpp_add<0x140, 0xf>(v); // row_mirror
    v = dpp_add<0x142, 0xa>(v); // row_bcast:15 into rows 1 and 3
    v = dpp_add<0x143, 0xc>(v); // row_bcast:31 into rows 2 and 3
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, WAVE_LANES - 1);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), WAVE_LANES - 1);
    return __longlong_as_double(((long long)hi << 32) | (long long)lo);
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
