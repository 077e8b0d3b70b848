// Control-step kernels for gfx950 (wave64).  One control step of a batch is
//   k_rollout (or k_rollout_ws) -> k_wavg -> k_update
// covering every instance of the batch in each launch (tile tables built on
// the host from the per-instance descriptors).
//
// Reference kernels these replace (src/kernels.cu):
//   generateNoiseKernel :18-31 + fullRolloutKernel :34-171   -> k_rollout
//   computeMaskedMinCostsKernel :221-301                     -> k_rollout (tile minima) + k_wavg
//   weightedAverageKernelUnified :307-408                    -> k_wavg + k_update
//   clampNominalKernel :434-456 + host shift mppi_controller.cu:430-436 -> k_update
// and the closed-loop environment step of SimulationEngine::dynStep
// (src/engine.cpp:286-312)                                 -> k_update (advance)
#include <hip/hip_runtime.h>

#include "dr_kernels.h"

namespace dr {

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
        v = fminf(v, __shfl_xor(v, m, WAVE));
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
        v += __shfl_xor(v, m, WAVE);
    return v;
}

// One lane = one sample k of one instance.  Per step:
// Philox noise (4 normals) -> float4 store of sigma*z into the sample's noise
// row -> clamp(nominal[t] + eps) -> Euler step -> gate update -> running
// cost.  Noise is generated NB steps at a time: the NB Philox/Box-Muller
// chains are independent, so the scheduler interleaves them, which hides the
// quarter-rate multiplies and transcendental latencies at the one or two
// waves per SIMD that a single-batch rollout gives.  The nominal, the track
// and the opponents' predicted positions (one row per step) sit in LDS.
template <int BLOCK, int NB>
__global__ __launch_bounds__(BLOCK) void k_rollout(const InstDev* __restrict__ insts, InstDyn* __restrict__ dyn,
                                                   const RTile* __restrict__ tiles, int n_tiles, const dr_track* __restrict__ tracks,
                                                   const dr_drone_state* __restrict__ drones, const float* __restrict__ nominal,
                                                   float* __restrict__ costs, float4* __restrict__ noise, float* __restrict__ tmin,
                                                   QuadConst C, CostConst Q, dr_cost_cfg KC, int opp_stride)
{
    extern __shared__ float4 s_opp[]; // [T][opp_stride]
    __shared__ float s_nom[DR_MAX_T * 4];
    __shared__ dr_track s_track;
    __shared__ float s_min[BLOCK / WAVE];

    const int tid = threadIdx.x;
    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    {
        const RTile rt = tiles[tile];
        const InstDev& I = insts[rt.inst];
        const int K = I.K, T = I.T;
        const uint32_t step = dyn[rt.inst].step;
        const int n_opp = I.n_opp;

        __syncthreads(); // LDS of the previous tile is no longer read
        for (int i = tid; i < T * 4; i += BLOCK)
            s_nom[i] = nominal[I.nom_off + i];
        {
            const float* src = reinterpret_cast<const float*>(tracks + I.track);
            float* dst = reinterpret_cast<float*>(&s_track);
            for (int i = tid; i < (int)(sizeof(dr_track) / 4); i += BLOCK)
                dst[i] = src[i];
        }
        for (int i = tid; i < T * n_opp; i += BLOCK)
        {
            const int t = i / n_opp, j = i - t * n_opp;
            s_opp[t * opp_stride + j] = opp_predict(drones[I.opp[j]], C.dt * (float)(t + 1));
        }
        __syncthreads();

        const int k = rt.k0 + tid;
        float cost = __builtin_inff();
        if (k < K)
        {
            QState s = state_from_abi(drones[I.drone]);
            float acc = 0.0f, decay = 1.0f;
            bool stop = false;
            float4* nrow = noise + I.noise_off + (long long)k * T;
            for (int t0 = 0; t0 < T; t0 += NB)
            {
                float e[NB][4];
#pragma unroll
                for (int u = 0; u < NB; ++u)
                {
                    float z[4];
                    normal4(I.key0, I.key1, (uint32_t)k, (uint32_t)(t0 + u), step, STREAM_CONTROL, z);
#pragma unroll
                    for (int d = 0; d < 4; ++d)
                        e[u][d] = I.sigma[d] * z[d];
                }
#pragma unroll
                for (int u = 0; u < NB; ++u)
                    if (t0 + u < T)
                        nrow[t0 + u] = make_float4(e[u][0], e[u][1], e[u][2], e[u][3]);
#pragma unroll
                for (int u = 0; u < NB; ++u)
                {
                    const int t = t0 + u;
                    if (t < T && !stop)
                    {
                        float uin[4], uc[4];
                        uin[0] = s_nom[t * 4 + 0] + e[u][0];
                        uin[1] = s_nom[t * 4 + 1] + e[u][1];
                        uin[2] = s_nom[t * 4 + 2] + e[u][2];
                        uin[3] = s_nom[t * 4 + 3] + e[u][3];
                        clamp_control(C, uin, uc);
                        const float p0[3] = {s.p[0], s.p[1], s.p[2]};
                        quad_step(C, s, uc);
                        update_gate(s_track.gates, s_track.n_gates, KC.gate_traversal_margin, p0, s);
                        int term;
                        const float c = stage_cost(C, Q, KC, s_track, s, s_opp + t * opp_stride, n_opp, &term);
                        acc = fma_(c, decay, acc);
                        decay = decay * DECAY;
                        stop = (term != TERM_NONE);
                    }
                }
            }
            acc = acc + final_cost(KC, s_track, s);
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

// Warp-specialised rollout: a workgroup of two waves covers 64 samples.
// Wave 0 (producer) generates the Philox/Box-Muller noise of a chunk of CH
// steps into a double-buffered LDS ring and into the samples' noise rows in
// HBM; wave 1 (consumer) integrates the dynamics and costs of the previous
// chunk from the ring.  One barrier per chunk.  This gives every SIMD two
// independent instruction streams when the batch has only one sample per
// lane per SIMD (65536 samples over 1024 SIMDs), so dependent VALU chains
// no longer stall the issue slot.  Arithmetic is identical to k_rollout.
template <int CH>
__global__ __launch_bounds__(128) void k_rollout_ws(const InstDev* __restrict__ insts, InstDyn* __restrict__ dyn,
                                                    const RTile* __restrict__ tiles, int n_tiles, const dr_track* __restrict__ tracks,
                                                    const dr_drone_state* __restrict__ drones, const float* __restrict__ nominal,
                                                    float* __restrict__ costs, float4* __restrict__ noise, float* __restrict__ tmin,
                                                    QuadConst C, CostConst Q, dr_cost_cfg KC, int opp_stride)
{
    extern __shared__ float4 s_opp[]; // [T][opp_stride]
    __shared__ float4 s_eps[2][CH][WAVE];
    __shared__ float s_nom[DR_MAX_T * 4];
    __shared__ dr_track s_track;
    __shared__ float s_min[2];

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const bool producer = tid < WAVE;
    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    {
        const RTile rt = tiles[tile];
        const InstDev& I = insts[rt.inst];
        const int K = I.K, T = I.T;
        const uint32_t step = dyn[rt.inst].step;
        const int n_opp = I.n_opp;
        const uint32_t key0 = I.key0, key1 = I.key1;
        const float sg0 = I.sigma[0], sg1 = I.sigma[1], sg2 = I.sigma[2], sg3 = I.sigma[3];

        __syncthreads(); // LDS of the previous tile is no longer read
        for (int i = tid; i < T * 4; i += 128)
            s_nom[i] = nominal[I.nom_off + i];
        {
            const float* src = reinterpret_cast<const float*>(tracks + I.track);
            float* dst = reinterpret_cast<float*>(&s_track);
            for (int i = tid; i < (int)(sizeof(dr_track) / 4); i += 128)
                dst[i] = src[i];
        }
        for (int i = tid; i < T * n_opp; i += 128)
        {
            const int t = i / n_opp, j = i - t * n_opp;
            s_opp[t * opp_stride + j] = opp_predict(drones[I.opp[j]], C.dt * (float)(t + 1));
        }
        __syncthreads();

        const int k = rt.k0 + lane;
        const bool live = k < K;
        const int n_chunks = (T + CH - 1) / CH;
        float4* nrow = noise + I.noise_off + (long long)k * T;

        QState s;
        float acc = 0.0f, decay = 1.0f;
        bool stop = false;
        if (!producer && live)
            s = state_from_abi(drones[I.drone]);

        for (int i = 0; i <= n_chunks; ++i)
        {
            if (producer)
            {
                if (i < n_chunks && live)
                {
                    float4* ring = &s_eps[i & 1][0][0];
#pragma unroll
                    for (int u = 0; u < CH; ++u)
                    {
                        const int t = i * CH + u;
                        float z[4];
                        normal4(key0, key1, (uint32_t)k, (uint32_t)t, step, STREAM_CONTROL, z);
                        const float4 e = make_float4(sg0 * z[0], sg1 * z[1], sg2 * z[2], sg3 * z[3]);
                        ring[u * WAVE + lane] = e;
                        if (t < T)
                            nrow[t] = e;
                    }
                }
            }
            else if (i > 0 && live)
            {
                const int j = i - 1;
                const float4* ring = &s_eps[j & 1][0][0];
#pragma unroll
                for (int u = 0; u < CH; ++u)
                {
                    const int t = j * CH + u;
                    if (t < T && !stop)
                    {
                        const float4 e = ring[u * WAVE + lane];
                        float uin[4], uc[4];
                        uin[0] = s_nom[t * 4 + 0] + e.x;
                        uin[1] = s_nom[t * 4 + 1] + e.y;
                        uin[2] = s_nom[t * 4 + 2] + e.z;
                        uin[3] = s_nom[t * 4 + 3] + e.w;
                        clamp_control(C, uin, uc);
                        const float p0[3] = {s.p[0], s.p[1], s.p[2]};
                        quad_step(C, s, uc);
                        update_gate(s_track.gates, s_track.n_gates, KC.gate_traversal_margin, p0, s);
                        int term;
                        const float c = stage_cost(C, Q, KC, s_track, s, s_opp + t * opp_stride, n_opp, &term);
                        acc = fma_(c, decay, acc);
                        decay = decay * DECAY;
                        stop = (term != TERM_NONE);
                    }
                }
            }
            __syncthreads();
        }

        float cost = __builtin_inff();
        if (!producer && live)
        {
            acc = acc + final_cost(KC, s_track, s);
            costs[I.cost_off + k] = acc;
            cost = acc;
        }
        const float wm = wave_min(cost);
        if (lane == 0)
            s_min[tid / WAVE] = wm;
        __syncthreads();
        if (tid == 0)
            tmin[tile] = fminf(s_min[0], s_min[1]);
    }
}

// Four-stage pipelined rollout: a workgroup of four waves covers 64 samples,
// one wave per stage, stages joined by double-buffered LDS rings holding CH
// steps, one barrier per chunk:
//   wave 0  Philox4x32-10 words for (k, t, step)
//   wave 1  Box-Muller + sigma scaling; stores the noise row in HBM
//   wave 2  clamp(nominal + eps), Euler step, gate update (integrates every
//           step; the early stop is resolved by wave 3)
//   wave 3  running cost, discount, terminal test; after a terminal state the
//           cost and the state used by the terminal cost are frozen
// At chunk i wave w works on chunk i - w.  Four waves per 64 samples keep four
// independent instruction streams per SIMD on a single small batch.  The
// arithmetic of every sample is that of k_rollout.
template <int CH>
__global__ __launch_bounds__(256) void k_rollout_p4(const InstDev* __restrict__ insts, InstDyn* __restrict__ dyn,
                                                    const RTile* __restrict__ tiles, int n_tiles, const dr_track* __restrict__ tracks,
                                                    const dr_drone_state* __restrict__ drones, const float* __restrict__ nominal,
                                                    float* __restrict__ costs, float4* __restrict__ noise, float* __restrict__ tmin,
                                                    QuadConst C, CostConst Q, dr_cost_cfg KC, int opp_stride)
{
    extern __shared__ float4 s_opp[]; // [T][opp_stride]
    __shared__ uint4 r_w[2][CH][WAVE];
    __shared__ float4 r_e[2][CH][WAVE];
    __shared__ float4 r_s[2][CH][WAVE];
    __shared__ float s_nom[DR_MAX_T * 4];
    __shared__ dr_track s_track;
    __shared__ float s_min[4];

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int role = tid / WAVE;
    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    {
        const RTile rt = tiles[tile];
        const InstDev& I = insts[rt.inst];
        const int K = I.K, T = I.T;
        const uint32_t step = dyn[rt.inst].step;
        const int n_opp = I.n_opp;
        const uint32_t key0 = I.key0, key1 = I.key1;
        const float sg0 = I.sigma[0], sg1 = I.sigma[1], sg2 = I.sigma[2], sg3 = I.sigma[3];

        __syncthreads(); // LDS of the previous tile is no longer read
        for (int i = tid; i < T * 4; i += 256)
            s_nom[i] = nominal[I.nom_off + i];
        {
            const float* src = reinterpret_cast<const float*>(tracks + I.track);
            float* dst = reinterpret_cast<float*>(&s_track);
            for (int i = tid; i < (int)(sizeof(dr_track) / 4); i += 256)
                dst[i] = src[i];
        }
        for (int i = tid; i < T * n_opp; i += 256)
        {
            const int t = i / n_opp, j = i - t * n_opp;
            s_opp[t * opp_stride + j] = opp_predict(drones[I.opp[j]], C.dt * (float)(t + 1));
        }
        __syncthreads();

        const int k = rt.k0 + lane;
        const bool live = k < K;
        const int n_chunks = (T + CH - 1) / CH;

        QState s; // wave 2: integrated state; wave 3: state at the last costed step
        if (live)
            s = state_from_abi(drones[I.drone]);
        float acc = 0.0f, decay = 1.0f;
        bool stop = false;

        for (int i = 0; i < n_chunks + 3; ++i)
        {
            const int c = i - role;
            if (live && c >= 0 && c < n_chunks)
            {
                if (role == 0)
                {
#pragma unroll
                    for (int u = 0; u < CH; ++u)
                    {
                        U4 ctr;
                        ctr.x = (uint32_t)k;
                        ctr.y = (uint32_t)(c * CH + u);
                        ctr.z = step;
                        ctr.w = STREAM_CONTROL;
                        const U4 w = philox4x32_10(ctr, key0, key1);
                        r_w[c & 1][u][lane] = make_uint4(w.x, w.y, w.z, w.w);
                    }
                }
                else if (role == 1)
                {
                    float4* nrow = noise + I.noise_off + (long long)k * T;
#pragma unroll
                    for (int u = 0; u < CH; ++u)
                    {
                        const int t = c * CH + u;
                        const uint4 w = r_w[c & 1][u][lane];
                        float z[4];
                        box_muller(w.x, w.y, &z[0], &z[1]);
                        box_muller(w.z, w.w, &z[2], &z[3]);
                        const float4 e = make_float4(sg0 * z[0], sg1 * z[1], sg2 * z[2], sg3 * z[3]);
                        r_e[c & 1][u][lane] = e;
                        if (t < T)
                            nrow[t] = e;
                    }
                }
                else if (role == 2)
                {
#pragma unroll
                    for (int u = 0; u < CH; ++u)
                    {
                        const int t = c * CH + u;
                        if (t < T)
                        {
                            const float4 e = r_e[c & 1][u][lane];
                            float uin[4], uc[4];
                            uin[0] = s_nom[t * 4 + 0] + e.x;
                            uin[1] = s_nom[t * 4 + 1] + e.y;
                            uin[2] = s_nom[t * 4 + 2] + e.z;
                            uin[3] = s_nom[t * 4 + 3] + e.w;
                            clamp_control(C, uin, uc);
                            const float p0[3] = {s.p[0], s.p[1], s.p[2]};
                            quad_step(C, s, uc);
                            update_gate(s_track.gates, s_track.n_gates, KC.gate_traversal_margin, p0, s);
                            r_s[c & 1][u][lane] = make_float4(s.p[0], s.p[1], s.p[2], u2f((uint32_t)(s.gate | (s.laps << 8))));
                        }
                    }
                }
                else
                {
#pragma unroll
                    for (int u = 0; u < CH; ++u)
                    {
                        const int t = c * CH + u;
                        if (t < T && !stop)
                        {
                            const float4 v = r_s[c & 1][u][lane];
                            const uint32_t code = f2u(v.w);
                            s.p[0] = v.x;
                            s.p[1] = v.y;
                            s.p[2] = v.z;
                            s.gate = (int)(code & 0xffu);
                            s.laps = (int)code >> 8;
                            int term;
                            const float cst = stage_cost(C, Q, KC, s_track, s, s_opp + t * opp_stride, n_opp, &term);
                            acc = fma_(cst, decay, acc);
                            decay = decay * DECAY;
                            stop = (term != TERM_NONE);
                        }
                    }
                }
            }
            __syncthreads();
        }

        float cost = __builtin_inff();
        if (role == 3 && live)
        {
            acc = acc + final_cost(KC, s_track, s);
            costs[I.cost_off + k] = acc;
            cost = acc;
        }
        const float wm = wave_min(cost);
        if (lane == 0)
            s_min[role] = wm;
        __syncthreads();
        if (tid == 0)
            tmin[tile] = fminf(fminf(s_min[0], s_min[1]), fminf(s_min[2], s_min[3]));
    }
}

// Noise-free rollout of each instance's nominal from its current state (the
// plan's own cost, reported as traj_cost): one workgroup per instance, the
// opponents' predicted positions in LDS, one active lane.  It runs on a
// forked stream concurrently with the sampled rollout.
__global__ __launch_bounds__(64) void k_eval(const InstDev* __restrict__ insts, InstDyn* __restrict__ dyn,
                                             const dr_track* __restrict__ tracks, const dr_drone_state* __restrict__ drones,
                                             const float* __restrict__ nominal, QuadConst C, CostConst Q, dr_cost_cfg KC, int opp_stride)
{
    extern __shared__ float4 s_opp[];
    __shared__ float s_nom[DR_MAX_T * 4];
    __shared__ dr_track s_track;
    const int inst = blockIdx.x, tid = threadIdx.x;
    const InstDev& I = insts[inst];
    const int T = I.T, n_opp = I.n_opp;
    for (int i = tid; i < T * n_opp; i += 64)
    {
        const int t = i / n_opp, j = i - t * n_opp;
        s_opp[t * opp_stride + j] = opp_predict(drones[I.opp[j]], C.dt * (float)(t + 1));
    }
    for (int i = tid; i < T * 4; i += 64)
        s_nom[i] = nominal[I.nom_off + i];
    {
        const float* src = reinterpret_cast<const float*>(tracks + I.track);
        float* dst = reinterpret_cast<float*>(&s_track);
        for (int i = tid; i < (int)(sizeof(dr_track) / 4); i += 64)
            dst[i] = src[i];
    }
    __syncthreads();
    if (tid != 0)
        return;
    QState s = state_from_abi(drones[I.drone]);
    float acc = 0.0f, decay = 1.0f;
    for (int t = 0; t < T; ++t)
    {
        float uin[4], uc[4];
        for (int d = 0; d < 4; ++d)
            uin[d] = s_nom[t * 4 + d] + 0.0f;
        clamp_control(C, uin, uc);
        const float p0[3] = {s.p[0], s.p[1], s.p[2]};
        quad_step(C, s, uc);
        update_gate(s_track.gates, s_track.n_gates, KC.gate_traversal_margin, p0, s);
        int term;
        const float c = stage_cost(C, Q, KC, s_track, s, s_opp + t * opp_stride, n_opp, &term);
        acc = fma_(c, decay, acc);
        decay = decay * DECAY;
        if (term != TERM_NONE)
            break;
    }
    dyn[inst].traj_cost = acc + final_cost(KC, s_track, s);
}

hipError_t launch_eval(int n_inst, hipStream_t st, const InstDev* insts, InstDyn* dyn, const dr_track* tracks,
                       const dr_drone_state* drones, const float* nominal, const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC,
                       int max_T, int opp_stride)
{
    const size_t shm = (size_t)max_T * opp_stride * sizeof(float4);
    k_eval<<<n_inst, 64, shm, st>>>(insts, dyn, tracks, drones, nominal, C, Q, KC, opp_stride);
    return hipGetLastError();
}

hipError_t launch_wavg(int chunk, int n_wtiles, hipStream_t st, const InstDev* insts, InstDyn* dyn, const WTile* wtiles,
                       const float* costs, const float4* noise, const float* tmin, double* part, double* den)
{
    if (chunk < 1 || chunk > WAVG_MAX_CHUNK)
        return hipErrorInvalidValue;
    k_wavg<<<n_wtiles, WAVG_BLOCK, 0, st>>>(insts, dyn, wtiles, costs, noise, tmin, part, den, chunk);
    return hipGetLastError();
}

hipError_t launch_update(int n_inst, hipStream_t st, const InstDev* insts, InstDyn* dyn, const double* part, const double* den,
                         float* nominal, float* actions, const QuadConst& C, int advance, dr_drone_state* drones, const int* drone_inst,
                         int n_drones, const dr_track* tracks, float margin)
{
    k_update<<<n_inst, 256, 0, st>>>(insts, dyn, part, den, nominal, actions, C, advance, drones, drone_inst, n_drones, tracks, margin);
    return hipGetLastError();
}

} // namespace dr
