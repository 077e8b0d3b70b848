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

#include "dr_adjoint.h"
#include "dr_kernels.h"
#include "dr_wave.h"

namespace dr {

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
    // The stage of each wave rotates with the workgroup: the waves of a
    // workgroup go to the CU's SIMDs in order, so a fixed assignment would put
    // every workgroup's dynamics wave (the heaviest stage) on one SIMD.
    const int role = ((tid / WAVE) + (int)blockIdx.x) & 3;
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

// One element of a sample's noise row: sigma (x) the four normals of Philox
// counter (k, t, step, STREAM_CONTROL) under the instance key (Box-Muller on
// words x, y and on z, w), as every rollout makes it.
__device__ __forceinline__ float4 control_noise(uint32_t key0, uint32_t key1, uint32_t k, uint32_t t, uint32_t step, float sg0,
                                                float sg1, float sg2, float sg3)
{
    const U4 w = philox4x32_10(U4{k, t, step, STREAM_CONTROL}, key0, key1);
    float z[4];
    box_muller(w.x, w.y, &z[0], &z[1]);
    box_muller(w.z, w.w, &z[2], &z[3]);
    return make_float4(sg0 * z[0], sg1 * z[1], sg2 * z[2], sg3 * z[3]);
}

// The noise rows of every instance, made ahead of the rollout that reads
// them (rollout_split 4).  Worker `wk` of `n_wk` takes, instance by instance,
// the units wk, wk + n_wk, ... of K x Q units: unit (k, q) is the span of
// ceil(T / Q) steps q of sample k's row, Q = ceil(n_wk / K) so that every
// worker has a unit where the rows allow.  Within a unit the Philox rounds that read only the
// sample index are made once, and four steps are in flight per trip.  ahead:
// the step after the current one (dyn.step_ahead, written by k_wavg before
// the refinement); otherwise the current step.
__device__ void make_noise_rows(const InstDev* __restrict__ insts, const InstDyn* __restrict__ dyn, int n_inst, float4* __restrict__ noise,
                                int wk, int n_wk, bool ahead)
{
    for (int i = 0; i < n_inst; ++i)
    {
        const InstDev& I = insts[i];
        const uint32_t step = ahead ? dyn[i].step_ahead : dyn[i].step;
        const int T = I.T, K = I.K;
        const int Q = min((T + 3) / 4, (n_wk + K - 1) / K), span = (T + Q - 1) / Q;
        const float sg0 = I.sigma[0], sg1 = I.sigma[1], sg2 = I.sigma[2], sg3 = I.sigma[3];
        for (int u = wk; u < K * Q; u += n_wk)
        {
            const int k = u / Q, q = u - k * Q;
            const int t0 = q * span, t1 = min(T, t0 + span);
            float4* row = noise + I.noise_off + (long long)k * T;
            for (int t = t0; t < t1; t += 4)
            {
                float4 v[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    v[j] = control_noise(I.key0, I.key1, (uint32_t)k, (uint32_t)min(t + j, t1 - 1), step, sg0, sg1, sg2, sg3);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (t + j < t1)
                        row[t + j] = v[j];
            }
        }
    }
}

// The current step's noise rows of every instance, before a rollout that
// reads them when they were not made ahead (the first step of a run).
__global__ __launch_bounds__(256) void k_noise(const InstDev* __restrict__ insts, const InstDyn* __restrict__ dyn, int n_inst,
                                               float4* __restrict__ noise)
{
    make_noise_rows(insts, dyn, n_inst, noise, blockIdx.x * 256 + threadIdx.x, gridDim.x * 256, false);
}

// Eight-stage pipelined rollout for small sample counts: a workgroup of eight
// waves covers 64 samples, so that a launch with few tiles still gives every
// SIMD four instruction streams.  Stages and the chunk each works on at
// chunk period i (stages joined by LDS rings of CH steps, one barrier per
// period):
//   P   Philox4x32-10 words                                        chunk i
//   B   Box-Muller, sigma, noise row store, even / odd (two)      chunk i - 1
//   R   clamp(nominal + eps), body rates and attitude             chunk i - 2
//   X   velocity, position from the pre-step attitude, gate       chunk i - 3
//   C   running cost and terminal test                            chunk i - 4
//   S   discounted sum in step order, stop at the first terminal  chunk i - 5
// A chunk period lasts as long as its busiest wave (one wave issues about
// one VALU instruction per 8 cycles), so the work is dealt out by count:
// waves 0-3 each make the Philox words and the running cost of the steps
// u = w mod 4 of their chunks (the running cost with 7 opponents is the
// largest stage), waves 4-5 the Box-Muller of even / odd steps, wave 6 the
// attitude and the discounted sum, wave 7 the translation.  Wave w works as
// (w + blockIdx) mod 8; waves w and w + 4 share a SIMD, so each SIMD pairs a
// cost wave with a Box-Muller, attitude or translation wave whatever the
// rotation.  The arithmetic of every sample is that of k_rollout.
template <int CH>
__global__ __launch_bounds__(512, 4) void k_rollout_p8(const InstDev* __restrict__ insts, InstDyn* __restrict__ dyn,
                                                       const RTile* __restrict__ tiles, int n_tiles, const dr_track* __restrict__ tracks,
                                                       const dr_drone_state* __restrict__ drones, const float* __restrict__ nominal,
                                                       float* __restrict__ costs, float4* __restrict__ noise, float* __restrict__ tmin,
                                                       QuadConst C, CostConst Q, dr_cost_cfg KC, int opp_stride)
{
    static_assert(CH % 2 == 0, "even and odd steps of a chunk");
    extern __shared__ float4 s_opp[]; // [T][opp_stride]
    __shared__ uint4 r_w[2][CH][WAVE];  // P -> B
    __shared__ float4 r_e[2][CH][WAVE]; // B -> R: sigma * z
    __shared__ float4 r_q[2][CH][WAVE]; // R -> X: attitude at the start of the step
    __shared__ float r_f[2][CH][WAVE];  // R -> X: clamped collective thrust
    __shared__ float4 r_s[3][CH][WAVE]; // X -> C, S: position after the step, gate | laps << 8
    __shared__ float2 r_c[2][CH][WAVE]; // C -> S: running cost, terminal code
    __shared__ float s_nom[DR_MAX_T * 4];
    __shared__ dr_track s_track;

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    // 0-3 P + C of steps u = role mod 4, 4 B-even, 5 B-odd, 6 R + S, 7 X
    const int role = ((tid / WAVE) + (int)blockIdx.x) & 7;
    const int parity = role & 1; // B waves
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
        for (int i = tid; i < T * 4; i += 512)
            s_nom[i] = nominal[I.nom_off + i];
        {
            const float* src = reinterpret_cast<const float*>(tracks + I.track);
            float* dst = reinterpret_cast<float*>(&s_track);
            for (int i = tid; i < (int)(sizeof(dr_track) / 4); i += 512)
                dst[i] = src[i];
        }
        for (int i = tid; i < T * n_opp; i += 512)
        {
            const int t = i / n_opp, j = i - t * n_opp;
            s_opp[t * opp_stride + j] = opp_predict(drones[I.opp[j]], C.dt * (float)(t + 1));
        }
        __syncthreads();

        const int k = rt.k0 + lane;
        const bool live = k < K;
        const int n_chunks = (T + CH - 1) / CH;

        QState s; // R: attitude and rates; X: position, velocity, gate, laps
        if (live)
            s = state_from_abi(drones[I.drone]);
        // S: discounted sum, the state the terminal cost reads (position and gate code)
        float acc = 0.0f, decay = 1.0f;
        float4 last = make_float4(s.p[0], s.p[1], s.p[2], u2f((uint32_t)(s.gate | (s.laps << 8))));
        bool stop = false;

        for (int i = 0; i < n_chunks + 5; ++i)
        {
            if (live)
            {
                if (role < 4)
                {
                    if (i < n_chunks)
                    {
#pragma unroll
                        for (int u = role; u < CH; u += 4)
                        {
                            U4 ctr;
                            ctr.x = (uint32_t)k;
                            ctr.y = (uint32_t)(i * CH + u);
                            ctr.z = step;
                            ctr.w = STREAM_CONTROL;
                            const U4 w = philox4x32_10(ctr, key0, key1);
                            r_w[i & 1][u][lane] = make_uint4(w.x, w.y, w.z, w.w);
                        }
                    }
                    const int c = i - 4;
                    if (c >= 0 && c < n_chunks)
                    {
#pragma unroll
                        for (int u = role; u < CH; u += 4)
                        {
                            const int t = c * CH + u;
                            if (t < T)
                            {
                                const float4 v = r_s[c % 3][u][lane];
                                const uint32_t code = f2u(v.w);
                                QState x;
                                x.p[0] = v.x;
                                x.p[1] = v.y;
                                x.p[2] = v.z;
                                x.gate = (int)(code & 0xffu);
                                x.laps = (int)code >> 8;
                                int term;
                                const float cst = stage_cost(C, Q, KC, s_track, x, s_opp + t * opp_stride, n_opp, &term);
                                r_c[c & 1][u][lane] = make_float2(cst, u2f((uint32_t)term));
                            }
                        }
                    }
                }
                else if (role < 6)
                {
                    const int c = i - 1;
                    if (c >= 0 && c < n_chunks)
                    {
                        float4* nrow = noise + I.noise_off + (long long)k * T;
#pragma unroll
                        for (int u = parity; u < CH; u += 2)
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
                }
                else if (role == 6)
                {
                    const int c = i - 2;
                    if (c >= 0 && c < n_chunks)
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
                                r_q[c & 1][u][lane] = make_float4(s.q[0], s.q[1], s.q[2], s.q[3]);
                                r_f[c & 1][u][lane] = uc[0];
                                quad_attitude_step(C, s.q, s.w, uc);
                            }
                        }
                    }
                    // S: the chunk costed one period earlier, in step order
                    const int cs = i - 5;
                    if (cs >= 0 && cs < n_chunks)
                    {
#pragma unroll
                        for (int u = 0; u < CH; ++u)
                        {
                            const int t = cs * CH + u;
                            if (t < T && !stop)
                            {
                                const float2 ct = r_c[cs & 1][u][lane];
                                last = r_s[cs % 3][u][lane];
                                acc = fma_(ct.x, decay, acc);
                                decay = decay * DECAY;
                                stop = ((int)f2u(ct.y) != TERM_NONE);
                            }
                        }
                    }
                }
                else
                {
                    const int c = i - 3;
                    if (c >= 0 && c < n_chunks)
                    {
#pragma unroll
                        for (int u = 0; u < CH; ++u)
                        {
                            const int t = c * CH + u;
                            if (t < T)
                            {
                                const float4 q4 = r_q[c & 1][u][lane];
                                const float q[4] = {q4.x, q4.y, q4.z, q4.w};
                                const float uu[4] = {r_f[c & 1][u][lane], 0.0f, 0.0f, 0.0f}; // the translation reads the thrust only
                                const float p0[3] = {s.p[0], s.p[1], s.p[2]};
                                quad_translation_step(C, q, s.p, s.v, uu);
                                update_gate(s_track.gates, s_track.n_gates, KC.gate_traversal_margin, p0, s);
                                r_s[c % 3][u][lane] = make_float4(s.p[0], s.p[1], s.p[2], u2f((uint32_t)(s.gate | (s.laps << 8))));
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }

        if (role == 6)
        {
            float cost = __builtin_inff();
            if (live)
            {
                const uint32_t code = f2u(last.w);
                QState x;
                x.p[0] = last.x;
                x.p[1] = last.y;
                x.p[2] = last.z;
                x.gate = (int)(code & 0xffu);
                x.laps = (int)code >> 8;
                acc = acc + final_cost(KC, s_track, x);
                costs[I.cost_off + k] = acc;
                cost = acc;
            }
            const float wm = wave_min(cost);
            if (lane == 0)
                tmin[tile] = wm;
        }
    }
}

// Eight-wave rollout over noise rows made ahead (rollout_split 4: by the
// spare workgroups of the previous step's refinement, or by k_noise): no
// Philox or Box-Muller in the rollout, the rows are read back.  Stages and the
// chunk each works on at chunk period i (LDS rings of CH = 4 or 8 steps, one
// barrier per period):
//   R   noise rows (requested two chunks ahead), clamp(nominal + eps), body
//       rates and attitude                                            chunk i
//   X   velocity, position from the pre-step attitude, gate           chunk i - 1
//   C   running cost and terminal test of step u of the chunk         chunk i - 2
//   S   discounted sum in step order, stop at the first terminal      chunk i - 3
// Roles (wave w works as (w + blockIdx) mod 8; waves w and w + 4 share a
// SIMD): 0, 1, 4, 5 the running cost of steps u = 0, 1, 2, 3 (mod 4) of each
// chunk; 6 R and S; 7 X; 2 and 3 idle, so the SIMD pairs are {C, C}, {C, C},
// {-, R + S}, {-, X}.  A longer chunk halves the barriers and pipeline fill
// and drain per step.  The arithmetic of every sample is that of k_rollout.
template <int CH>
__global__ __launch_bounds__(512, 4) void k_rollout_p8nz(const InstDev* __restrict__ insts, InstDyn* __restrict__ dyn,
                                                         const RTile* __restrict__ tiles, int n_tiles, const dr_track* __restrict__ tracks,
                                                         const dr_drone_state* __restrict__ drones, const float* __restrict__ nominal,
                                                         float* __restrict__ costs, const float4* __restrict__ noise,
                                                         float* __restrict__ tmin, QuadConst C, CostConst Q, dr_cost_cfg KC, int opp_stride)
{
    static_assert(CH % 4 == 0, "four cost waves share each chunk");
    extern __shared__ float4 s_opp[]; // [T][opp_stride]
    __shared__ float4 r_q[2][CH][WAVE]; // R -> X: attitude at the start of the step
    __shared__ float r_f[2][CH][WAVE];  // R -> X: clamped collective thrust
    __shared__ float4 r_s[3][CH][WAVE]; // X -> C, S: position after the step, gate | laps << 8
    __shared__ float2 r_c[2][CH][WAVE]; // C -> S: running cost, terminal code
    __shared__ float s_nom[DR_MAX_T * 4];
    __shared__ dr_track s_track;

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int role = ((tid / WAVE) + (int)blockIdx.x) & 7;
    // the step of each chunk whose running cost a cost wave makes
    const int cu = role < 2 ? role : role - 2;
    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    {
        const RTile rt = tiles[tile];
        const InstDev& I = insts[rt.inst];
        const int K = I.K, T = I.T;
        const int n_opp = I.n_opp;

        __syncthreads(); // LDS of the previous tile is no longer read
        for (int i = tid; i < T * 4; i += 512)
            s_nom[i] = nominal[I.nom_off + i];
        {
            const float* src = reinterpret_cast<const float*>(tracks + I.track);
            float* dst = reinterpret_cast<float*>(&s_track);
            for (int i = tid; i < (int)(sizeof(dr_track) / 4); i += 512)
                dst[i] = src[i];
        }
        for (int i = tid; i < T * n_opp; i += 512)
        {
            const int t = i / n_opp, j = i - t * n_opp;
            s_opp[t * opp_stride + j] = opp_predict(drones[I.opp[j]], C.dt * (float)(t + 1));
        }
        __syncthreads();

        const int k = rt.k0 + lane;
        const bool live = k < K;
        const int n_chunks = (T + CH - 1) / CH;
        const float4* nrow = noise + I.noise_off + (long long)(live ? k : 0) * T;

        QState s; // R: attitude and rates; X: position, velocity, gate, laps
        if (live)
            s = state_from_abi(drones[I.drone]);
        // S: discounted sum, the state the terminal cost reads (position and gate code)
        float acc = 0.0f, decay = 1.0f;
        float4 last = make_float4(s.p[0], s.p[1], s.p[2], u2f((uint32_t)(s.gate | (s.laps << 8))));
        bool stop = false;
        // R: with CH = 4 the rows of the even chunks in ea, of the odd chunks
        // in eb, each requested two periods before it is used; with CH = 8
        // (periods twice as long, registers scarce) ea alone, one period ahead
        float4 ea[CH], eb[CH == 4 ? CH : 1];
        const auto request = [&](float4 (&e)[CH], int c) {
#pragma unroll
            for (int u = 0; u < CH; ++u)
                e[u] = nrow[min(c * CH + u, T - 1)];
        };
        const auto attitude = [&](const float4 (&e)[CH], int c) {
#pragma unroll
            for (int u = 0; u < CH; ++u)
            {
                const int t = c * CH + u;
                if (t < T)
                {
                    float uin[4], uc[4];
                    uin[0] = s_nom[t * 4 + 0] + e[u].x;
                    uin[1] = s_nom[t * 4 + 1] + e[u].y;
                    uin[2] = s_nom[t * 4 + 2] + e[u].z;
                    uin[3] = s_nom[t * 4 + 3] + e[u].w;
                    clamp_control(C, uin, uc);
                    r_q[c & 1][u][lane] = make_float4(s.q[0], s.q[1], s.q[2], s.q[3]);
                    r_f[c & 1][u][lane] = uc[0];
                    quad_attitude_step(C, s.q, s.w, uc);
                }
            }
        };
        if (live && role == 6)
        {
            request(ea, 0);
            if constexpr (CH == 4)
                request(eb, 1);
        }

        for (int i = 0; i < n_chunks + 3; ++i)
        {
            if (live)
            {
                if (role == 6)
                {
                    const int c = i;
                    if (c < n_chunks)
                    {
                        if constexpr (CH == 4)
                        {
                            if ((c & 1) == 0)
                            {
                                attitude(ea, c);
                                request(ea, c + 2);
                            }
                            else
                            {
                                attitude(eb, c);
                                request(eb, c + 2);
                            }
                        }
                        else
                        {
                            attitude(ea, c);
                            request(ea, c + 1);
                        }
                    }
                    // S: the chunk costed one period earlier, in step order
                    const int cs = i - 3;
                    if (cs >= 0 && cs < n_chunks)
                    {
#pragma unroll
                        for (int u = 0; u < CH; ++u)
                        {
                            const int t = cs * CH + u;
                            if (t < T && !stop)
                            {
                                const float2 ct = r_c[cs & 1][u][lane];
                                last = r_s[cs % 3][u][lane];
                                acc = fma_(ct.x, decay, acc);
                                decay = decay * DECAY;
                                stop = ((int)f2u(ct.y) != TERM_NONE);
                            }
                        }
                    }
                }
                else if (role == 7)
                {
                    const int c = i - 1;
                    if (c >= 0 && c < n_chunks)
                    {
#pragma unroll
                        for (int u = 0; u < CH; ++u)
                        {
                            const int t = c * CH + u;
                            if (t < T)
                            {
                                const float4 q4 = r_q[c & 1][u][lane];
                                const float q[4] = {q4.x, q4.y, q4.z, q4.w};
                                const float uu[4] = {r_f[c & 1][u][lane], 0.0f, 0.0f, 0.0f}; // the translation reads the thrust only
                                const float p0[3] = {s.p[0], s.p[1], s.p[2]};
                                quad_translation_step(C, q, s.p, s.v, uu);
                                update_gate(s_track.gates, s_track.n_gates, KC.gate_traversal_margin, p0, s);
                                r_s[c % 3][u][lane] = make_float4(s.p[0], s.p[1], s.p[2], u2f((uint32_t)(s.gate | (s.laps << 8))));
                            }
                        }
                    }
                }
                else if (role != 2 && role != 3)
                {
                    const int c = i - 2;
                    if (c >= 0 && c < n_chunks)
                    {
#pragma unroll
                        for (int u = cu; u < CH; u += 4)
                        {
                            const int t = c * CH + u;
                            if (t < T)
                            {
                                const float4 v = r_s[c % 3][u][lane];
                                const uint32_t code = f2u(v.w);
                                QState x;
                                x.p[0] = v.x;
                                x.p[1] = v.y;
                                x.p[2] = v.z;
                                x.gate = (int)(code & 0xffu);
                                x.laps = (int)code >> 8;
                                int term;
                                const float cst = stage_cost(C, Q, KC, s_track, x, s_opp + t * opp_stride, n_opp, &term);
                                r_c[c & 1][u][lane] = make_float2(cst, u2f((uint32_t)term));
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }

        if (role == 6)
        {
            float cost = __builtin_inff();
            if (live)
            {
                const uint32_t code = f2u(last.w);
                QState x;
                x.p[0] = last.x;
                x.p[1] = last.y;
                x.p[2] = last.z;
                x.gate = (int)(code & 0xffu);
                x.laps = (int)code >> 8;
                acc = acc + final_cost(KC, s_track, x);
                costs[I.cost_off + k] = acc;
                cost = acc;
            }
            const float wm = wave_min(cost);
            if (lane == 0)
                tmin[tile] = wm;
        }
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

// ---------------------------------------------------------------------------
// Adjoint rollout: J(U) and dJ/dU for control sequences of one instance,
// noise-free, from the instance's current drone state.  Checkpoints every
// ck steps (one state slot per checkpoint in LDS), each segment recomputed
// forward into LDS, then swept backward (dr_adjoint.h).  A state slot holds
// (p, q, v, w, gate, laps) and the discount weight of that state's running
// cost in four float4 words.

struct AdjCtx
{
    const dr_track* tr;
    const float4* opp; // [T][opp_stride] predicted opponent positions
    int opp_stride, n_opp, T, ck;
    QuadConst C;
    CostConst Q;
    dr_cost_cfg K;
};

__device__ __forceinline__ void put_state(float4* buf, int slot, int stride, const QState& s, float weight)
{
    buf[(slot * 4 + 0) * stride] = make_float4(s.p[0], s.p[1], s.p[2], s.q[0]);
    buf[(slot * 4 + 1) * stride] = make_float4(s.q[1], s.q[2], s.q[3], s.v[0]);
    buf[(slot * 4 + 2) * stride] = make_float4(s.v[1], s.v[2], s.w[0], s.w[1]);
    buf[(slot * 4 + 3) * stride] = make_float4(s.w[2], u2f((uint32_t)s.gate), u2f((uint32_t)s.laps), weight);
}

__device__ __forceinline__ QState get_state(const float4* buf, int slot, int stride, float* weight)
{
    QState s;
    const float4 a = buf[(slot * 4 + 0) * stride], b = buf[(slot * 4 + 1) * stride];
    const float4 c = buf[(slot * 4 + 2) * stride], d = buf[(slot * 4 + 3) * stride];
    s.p[0] = a.x; s.p[1] = a.y; s.p[2] = a.z; s.q[0] = a.w;
    s.q[1] = b.x; s.q[2] = b.y; s.q[3] = b.z; s.v[0] = b.w;
    s.v[1] = c.x; s.v[2] = c.y; s.w[0] = c.z; s.w[1] = c.w;
    s.w[2] = d.x; s.gate = (int)f2u(d.y); s.laps = (int)f2u(d.z);
    *weight = d.w;
    return s;
}

__device__ __forceinline__ void adj_step(const AdjCtx& A, QState& s, const float* U, int t, float uc[4])
{
    float uin[4] = {U[t * 4 + 0], U[t * 4 + 1], U[t * 4 + 2], U[t * 4 + 3]};
    clamp_control(A.C, uin, uc);
    const float p0[3] = {s.p[0], s.p[1], s.p[2]};
    quad_step(A.C, s, uc);
    update_gate(A.tr->gates, A.tr->n_gates, A.K.gate_traversal_margin, p0, s);
}

// One control sequence; buf = this lane's state slots (stride between words).
__device__ float cost_grad_lane(const AdjCtx& A, const QState& x0, const float* U, float* G, float4* buf, int stride)
{
    const int T = A.T, ck = A.ck;
    const int n_ck = (T + ck - 1) / ck;
    QState s = x0;
    float acc = 0.0f, decay = 1.0f, w_prev = 0.0f;
    int t_end = T - 1;
    for (int t = 0; t < T; ++t)
    {
        if (t % ck == 0)
            put_state(buf, t / ck, stride, s, w_prev);
        float uc[4];
        adj_step(A, s, U, t, uc);
        int term;
        const float c = stage_cost(A.C, A.Q, A.K, *A.tr, s, A.opp + t * A.opp_stride, A.n_opp, &term);
        acc = fma_(c, decay, acc);
        w_prev = decay;
        decay = decay * DECAY;
        if (term != TERM_NONE)
        {
            t_end = t;
            break;
        }
    }
    const float J = acc + final_cost(A.K, *A.tr, s);
    if (!G)
        return J;
    for (int i = 0; i < T * 4; ++i)
        G[i] = 0.0f;
    QGrad g;
    for (int i = 0; i < 3; ++i)
        g.p[i] = g.v[i] = g.w[i] = 0.0f;
    for (int i = 0; i < 4; ++i)
        g.q[i] = 0.0f;
    final_cost_grad(A.K, *A.tr, s, g.p);
    stage_cost_grad(A.C, A.Q, A.K, *A.tr, s, A.opp + t_end * A.opp_stride, A.n_opp, w_prev, g.p);
    for (int sg = t_end / ck; sg >= 0; --sg)
    {
        const int t0 = sg * ck;
        const int t1 = min(t0 + ck, t_end + 1);
        float w;
        QState x = get_state(buf, sg, stride, &w);
        for (int t = t0; t < t1; ++t)
        {
            put_state(buf, n_ck + (t - t0), stride, x, w);
            float uc[4];
            adj_step(A, x, U, t, uc);
            w = (t == 0) ? 1.0f : w * DECAY;
        }
        for (int t = t1 - 1; t >= t0; --t)
        {
            float wt;
            const QState xt = get_state(buf, n_ck + (t - t0), stride, &wt);
            float uin[4] = {U[t * 4 + 0], U[t * 4 + 1], U[t * 4 + 2], U[t * 4 + 3]}, uc[4], gu[4];
            clamp_control(A.C, uin, uc);
            quad_step_vjp(A.C, xt, uc, g, gu);
            G[t * 4 + 0] = gu[0] * clamp_grad(uin[0], A.C.thrust_min, A.C.thrust_max);
            G[t * 4 + 1] = gu[1] * clamp_grad(uin[1], -A.C.rate_max, A.C.rate_max);
            G[t * 4 + 2] = gu[2] * clamp_grad(uin[2], -A.C.rate_max, A.C.rate_max);
            G[t * 4 + 3] = gu[3] * clamp_grad(uin[3], -A.C.rate_max, A.C.rate_max);
            if (t >= 1)
                stage_cost_grad(A.C, A.Q, A.K, *A.tr, xt, A.opp + (t - 1) * A.opp_stride, A.n_opp, wt, g.p);
        }
    }
    return J;
}

__device__ void stage_inputs(const InstDev& I, const dr_track* tracks, const dr_drone_state* drones, const QuadConst& C, dr_track* s_track,
                             float4* s_opp, int opp_stride)
{
    const int T = I.T, n_opp = I.n_opp;
    for (int i = threadIdx.x; i < T * n_opp; i += blockDim.x)
    {
        const int t = i / n_opp, j = i - t * n_opp;
        s_opp[t * opp_stride + j] = opp_predict(drones[I.opp[j]], C.dt * (float)(t + 1));
    }
    const float* src = reinterpret_cast<const float*>(tracks + I.track);
    float* dst = reinterpret_cast<float*>(s_track);
    for (int i = threadIdx.x; i < (int)(sizeof(dr_track) / 4); i += blockDim.x)
        dst[i] = src[i];
}

// J and dJ/dU for n_seq sequences U[n_seq][T][4] of instance `inst`, one lane
// per sequence.  Dynamic LDS: opponent table, then lanes x slots state words.
__global__ void k_costgrad(const InstDev* __restrict__ insts, int inst, const dr_track* __restrict__ tracks,
                           const dr_drone_state* __restrict__ drones, const float* __restrict__ U, int n_seq, float* __restrict__ J,
                           float* __restrict__ G, int ck, QuadConst C, CostConst Q, dr_cost_cfg KC, int opp_stride)
{
    extern __shared__ float4 s_dyn[];
    __shared__ dr_track s_track;
    const InstDev& I = insts[inst];
    const int T = I.T;
    float4* s_opp = s_dyn;
    float4* s_buf = s_dyn + T * opp_stride;
    stage_inputs(I, tracks, drones, C, &s_track, s_opp, opp_stride);
    __syncthreads();
    const int seq = blockIdx.x * blockDim.x + threadIdx.x;
    if (seq >= n_seq)
        return;
    AdjCtx A{&s_track, s_opp, opp_stride, I.n_opp, T, ck, C, Q, KC};
    const QState x0 = state_from_abi(drones[I.drone]);
    J[seq] = cost_grad_lane(A, x0, U + (long long)seq * T * 4, G ? G + (long long)seq * T * 4 : nullptr, s_buf + threadIdx.x,
                            blockDim.x);
}

// Adjoint refinement of each instance's updated nominal: n_steps projected
// gradient steps U <- clamp(U - step * dJ/dU), one workgroup of REF_BLOCK
// lanes per instance, everything in LDS.  Every iterate's J is evaluated
// (one more forward pass after the last step) and the lowest-J sequence is
// kept, the plain MPPI update (iterate 0) on ties: a step that jumps over a
// discontinuity of the cost (a lost gate, a collision) is not taken.  J and
// dJ/dU are those of cost_grad_lane (the same dynamics, terminal rule and
// vector-Jacobian products); the work is spread over the lanes, the time
// recurrences included:
//  P  the iterate's controls are clamped, one lane per step.
//  A  the dynamics over all T steps, every state into the state table, as
//     block-wide scans (one lane per step): the body rates and the velocity
//     are affine recurrences, the position a prefix sum and the attitude a
//     product of unit quaternions, all associative in exact arithmetic, so
//     the states are quad_step's to rounding.  Then, one lane per (step,
//     gate) pair, the gate-crossing tests, and one lane per step the running
//     cost of x_{t+1} without the winner test and its dJ/dp (both read the
//     position only).
//  G  wave 0 runs the gate counter by searching the crossing bits for the
//     next gate's crossing (one ballot per 64 steps).
//  C  one lane per step t: the gate and laps of x_{t+1} (after the crossings
//     before it) into its state slot, the winner test on x_{t+1}, the
//     discounted running cost; the first terminal step t_end ends the
//     trajectory (cost_grad_lane's break); the last wave sums J while
//     waves 0-2 make the segment maps of D.
//  D  the reverse recurrence lam_t = A_t^T lam_{t+1} + g_t is linear in lam.
//     The steps 0..t_end are cut into n_seg <= REF_SEGS reverse segments
//     [a, b); per segment, 14 lanes carry lam_b = 0 with the g_t terms
//     (giving psi) and lam_b = e_i, i < 13, without them (giving column i of
//     Psi) down to a: lam_a = Psi lam_b + psi.  Each chain also keeps its
//     dJ/du_t at every step of the segment.
//  E  wave 0, lane i holding component i: lam_b of every segment, from the
//     terminal adjoint down to the first segment (one 13 x 13
//     matrix-vector product per segment).
//  F  one lane per step: dJ/du_t is the segment's psi chain plus its Psi
//     chains weighted by lam_b (the products are linear in lam), then the
//     projected step of u_t; steps after t_end have dJ/du = 0 (u_t is only
//     clamped).
// The forward states are reassociated (A) and the additions of C-F ordered
// differently from the one-lane sweep, so J and the gradient agree with
// cost_grad_lane to rounding, not bit for bit.
constexpr int REF_BLOCK = 256;
constexpr int REF_SEGS = 13; // 14 * REF_SEGS <= REF_BLOCK - WAVE: the last wave is free for J during D
constexpr int REF_COMP = 14; // affine part + 13 state directions
constexpr int REF_JAC4 = (QJAC + 3) / 4; // float4 slots per step of the Jacobian table

// Dynamic LDS of k_refine for horizons up to tcap: the float4 tables (state
// slots, sequence, dJ/dp, chain gradients, Jacobian blocks, best sequence)
// and seven 4-byte tables.
__host__ __device__ constexpr size_t refine_lds_bytes(int tcap)
{
    return sizeof(float4) * ((size_t)(tcap + 1) * 4 + (size_t)tcap * (3 + REF_COMP + REF_JAC4)) + 4 * 7 * (size_t)tcap;
}

// The running-cost weights DECAY^t as the forward passes form them (a running
// product, one multiply per step), tabulated at compile time.
struct DecayTable
{
    float w[DR_MAX_T];
};

constexpr DecayTable make_decay_table()
{
    DecayTable d{};
    float w = 1.0f;
    for (int t = 0; t < DR_MAX_T; ++t)
    {
        d.w[t] = w;
        w = w * DECAY;
    }
    return d;
}

__constant__ DecayTable c_decay = make_decay_table();

// Block-wide inclusive scans for the refinement's forward pass: lane t of
// the REF_BLOCK-lane workgroup holds element t (identity past the last
// element), and after the scan it holds e_0 o e_1 o ... o e_t for an
// associative op(earlier, later), and ex holds e_0 o ... o e_{t-1} (the
// identity at t = 0).  Within a wave: row shifts by 1, 2, 4, 8 and the row
// broadcasts of lanes 15 and 31 (DPP, no LDS), then a wave shift by one for
// the exclusive prefix; across waves: each wave's total through LDS (one
// barrier; each scan uses its own s_tot).  A lane whose shifted source lies
// outside its row receives the identity, and op(identity, x) = x.
template <int CTRL, int ROWS, int N>
__device__ __forceinline__ void dpp_shift(const float (&x)[N], const float (&id)[N], float (&out)[N])
{
#pragma unroll
    for (int i = 0; i < N; ++i)
        out[i] = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(id[i]), __float_as_int(x[i]), CTRL, ROWS, 0xf, false));
}

template <int N, class Op>
__device__ __forceinline__ void block_scan(float (&x)[N], float (&ex)[N], const float (&id)[N], Op op, float* s_tot)
{
    float y[N];
    dpp_shift<0x111, 0xf>(x, id, y); // row_shr:1
    op(y, x);
    dpp_shift<0x112, 0xf>(x, id, y); // row_shr:2
    op(y, x);
    dpp_shift<0x114, 0xf>(x, id, y); // row_shr:4
    op(y, x);
    dpp_shift<0x118, 0xf>(x, id, y); // row_shr:8
    op(y, x);
    dpp_shift<0x142, 0xa>(x, id, y); // row_bcast:15 into rows 1 and 3
    op(y, x);
    dpp_shift<0x143, 0xc>(x, id, y); // row_bcast:31 into rows 2 and 3
    op(y, x);
    dpp_shift<0x138, 0xf>(x, id, ex); // wave_shr:1: the exclusive prefix within the wave
    const int wave = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == WAVE - 1)
#pragma unroll
        for (int i = 0; i < N; ++i)
            s_tot[wave * N + i] = x[i];
    __syncthreads();
    if (wave > 0)
    {
        float pre[N];
#pragma unroll
        for (int i = 0; i < N; ++i)
            pre[i] = s_tot[i];
        for (int w = 1; w < wave; ++w)
        {
            float t[N];
#pragma unroll
            for (int i = 0; i < N; ++i)
                t[i] = s_tot[w * N + i];
            op(pre, t);
#pragma unroll
            for (int i = 0; i < N; ++i)
                pre[i] = t[i];
        }
        op(pre, x);
        op(pre, ex);
    }
}

// op(e, l): l <- the affine map "l after e", maps x -> A x + B_i (one A for
// the three axes): (A_l A_e, A_l B_e + B_l).
struct AffineOp
{
    __device__ void operator()(const float (&e)[4], float (&l)[4]) const
    {
        l[1] = fma_(l[0], e[1], l[1]);
        l[2] = fma_(l[0], e[2], l[2]);
        l[3] = fma_(l[0], e[3], l[3]);
        l[0] = l[0] * e[0];
    }
};

// op(e, l): l <- e (x) l (quaternion product, earlier factor on the left)
struct QuatOp
{
    __device__ void operator()(const float (&e)[4], float (&l)[4]) const
    {
        const float w = e[0] * l[0] - e[1] * l[1] - e[2] * l[2] - e[3] * l[3];
        const float x = e[0] * l[1] + e[1] * l[0] + e[2] * l[3] - e[3] * l[2];
        const float y = e[0] * l[2] - e[1] * l[3] + e[2] * l[0] + e[3] * l[1];
        const float z = e[0] * l[3] + e[1] * l[2] - e[2] * l[1] + e[3] * l[0];
        l[0] = w;
        l[1] = x;
        l[2] = y;
        l[3] = z;
    }
};

struct AddOp
{
    __device__ void operator()(const float (&e)[3], float (&l)[3]) const
    {
        l[0] = e[0] + l[0];
        l[1] = e[1] + l[1];
        l[2] = e[2] + l[2];
    }
};

__global__ __launch_bounds__(REF_BLOCK) void k_refine(const InstDev* __restrict__ insts, InstDyn* __restrict__ dyn,
                                                      const double* __restrict__ part, const double* __restrict__ den,
                                                      const dr_track* __restrict__ tracks, const dr_drone_state* __restrict__ drones,
                                                      float* __restrict__ nominal, int n_steps, float step, QuadConst C, CostConst Q,
                                                      dr_cost_cfg KC, float* __restrict__ actions, int advance,
                                                      const int* __restrict__ drone_inst, int n_drones,
                                                      dr_drone_state* __restrict__ drones_out, int n_inst, float4* __restrict__ noise,
                                                      int tcap)
{
    // Workgroups past the instances make the next step's noise rows (for a
    // rollout over noise made ahead) on the CUs the refinement leaves idle:
    // the refinement's workgroups come first in the grid and hold their CUs
    // alone (each takes most of a CU's LDS).
    if ((int)blockIdx.x >= n_inst)
    {
        make_noise_rows(insts, dyn, n_inst, noise, ((int)blockIdx.x - n_inst) * REF_BLOCK + (int)threadIdx.x,
                        ((int)gridDim.x - n_inst) * REF_BLOCK, true);
        return;
    }
    __shared__ dr_track s_track;
    __shared__ float s_psi[REF_SEGS * REF_COMP * 16];
    __shared__ float s_lam[REF_SEGS * 16]; // lam_b of each reverse segment
    __shared__ float4 s_op[MAX_OPP], s_ov[MAX_OPP];
    __shared__ float s_tot[4][(REF_BLOCK / WAVE) * 4]; // per-wave totals of the forward pass's four scans
    __shared__ int s_nx;
    __shared__ float s_jbest, s_fin, s_fgrad[3];
    __shared__ int s_tend, s_better;
    // The per-step tables, in dynamic LDS sized by the batch's longest horizon
    // tcap (refine_lds_bytes): at T = 100 two workgroups fit on one CU.
    extern __shared__ float4 s_steps[];
    float4* const s_x = s_steps;                 // (tcap + 1) * 4: state table, put_state layout
    float4* const s_u = s_x + (tcap + 1) * 4;    // the sequence being refined
    float4* const s_gp = s_u + tcap;             // dJ/dp of the running cost of x_{t+1}
    float4* const s_gu = s_gp + tcap;            // tcap * REF_COMP: per step t and chain c of its segment: dJ/du_t along chain c
    float4* const s_jac = s_gu + tcap * REF_COMP; // tcap * REF_JAC4: per step: the Jacobian blocks of quad_step_jacobian
    float4* const s_ubest = s_jac + tcap * REF_JAC4; // the lowest-cost sequence evaluated so far
    float* const s_w = reinterpret_cast<float*>(s_ubest + tcap); // running-cost weight of step t
    float* const s_cb = s_w + tcap;              // running cost of x_{t+1} without the winner test
    int* const s_tb = reinterpret_cast<int*>(s_cb + tcap); // its terminal code (collision or none)
    uint32_t* const s_cross = reinterpret_cast<uint32_t*>(s_tb + tcap); // bit g: the move x_t -> x_{t+1} crosses gate g
    int* const s_xstep = reinterpret_cast<int*>(s_cross + tcap); // gate crossings: step, gate, laps after
    int* const s_xgate = s_xstep + tcap;
    int* const s_xlaps = s_xgate + tcap;

    // The refinement is a chain of dependent phases: its waves take the issue
    // slots first where noise workgroups share the CU.
    __builtin_amdgcn_s_setprio(3);
    const InstDev& I = insts[blockIdx.x];
    const int T = I.T, n_opp = I.n_opp, tid = threadIdx.x;
    const float margin = KC.gate_traversal_margin;
    const uint32_t step0 = dyn[blockIdx.x].step; // read now, incremented at the end
    {
        const float* src = reinterpret_cast<const float*>(tracks + I.track);
        float* dst = reinterpret_cast<float*>(&s_track);
        for (int i = tid; i < (int)(sizeof(dr_track) / 4); i += REF_BLOCK)
            dst[i] = src[i];
    }
    if (tid < n_opp)
    {
        const dr_drone_state& o = drones[I.opp[tid]];
        s_op[tid] = make_float4(o.p[0], o.p[1], o.p[2], 0.0f);
        s_ov[tid] = make_float4(o.v[0], o.v[1], o.v[2], 0.0f);
    }
    {
        // the MPPI update (as k_update makes it): nominal + num/den over the
        // chunks in fixed order, box clamp; nu and the temperature rule.  Each
        // lane sums the denominator and its (up to 4) values together, eight
        // chunks of loads in flight at a time.
        constexpr int VPL = (DR_MAX_T * 4 + REF_BLOCK - 1) / REF_BLOCK;
        const int nc = I.n_chunks;
        const long long st = (long long)T * 4;
        const double* pd = den + I.den_off;
        const double* pp = part + I.part_off;
        double dn = 0.0, num[VPL];
        float nom[VPL];
#pragma unroll
        for (int k = 0; k < VPL; ++k)
        {
            const int i = tid + k * REF_BLOCK;
            num[k] = 0.0;
            nom[k] = i < T * 4 ? nominal[I.nom_off + i] : 0.0f;
        }
        for (int c0 = 0; c0 < nc; c0 += 8)
        {
            double vd[8], vp[VPL][8];
#pragma unroll
            for (int b = 0; b < 8; ++b)
            {
                vd[b] = c0 + b < nc ? pd[c0 + b] : 0.0;
#pragma unroll
                for (int k = 0; k < VPL; ++k)
                {
                    const int i = tid + k * REF_BLOCK;
                    vp[k][b] = (c0 + b < nc && i < T * 4) ? pp[(c0 + b) * st + i] : 0.0;
                }
            }
#pragma unroll
            for (int b = 0; b < 8; ++b)
            {
                dn += vd[b];
#pragma unroll
                for (int k = 0; k < VPL; ++k)
                    num[k] += vp[k][b];
            }
        }
        float* su = reinterpret_cast<float*>(s_u);
#pragma unroll
        for (int k = 0; k < VPL; ++k)
        {
            const int i = tid + k * REF_BLOCK;
            if (i < T * 4)
                su[i] = dn > 1e-30 ? nom[k] + (float)(num[k] / dn) : nom[k];
        }
        __syncthreads();
        for (int t = tid; t < T; t += REF_BLOCK)
        {
            const float4 u4 = s_u[t];
            const float in[4] = {u4.x, u4.y, u4.z, u4.w};
            float out[4];
            clamp_control(C, in, out);
            s_u[t] = make_float4(out[0], out[1], out[2], out[3]);
        }
        if (tid == 0)
        {
            InstDyn& Y = dyn[blockIdx.x];
            const float nu = (float)dn;
            Y.nu = nu;
            if (I.adapt)
            {
                if (nu > 0.001f * (float)I.K)
                    Y.lambda = Y.lambda * 0.9f;
                else if (nu < 0.0005f * (float)I.K)
                    Y.lambda = Y.lambda * 1.2f;
            }
        }
    }
    for (int t = tid; t < T; t += REF_BLOCK)
        s_w[t] = c_decay.w[t]; // the running-cost weights
    const QState x0 = state_from_abi(drones[I.drone]);
    float* const s_xf = reinterpret_cast<float*>(s_x); // the state table as floats: slot t at 16 t (p0 p1 p2 q0 | q1 q2 q3 v0 | v1 v2 w0 w1 | w2 gate laps -)
    __syncthreads();

    for (int it = 0;; ++it)
    {
        // The iterate's controls are already clamped (the MPPI update and every
        // projected step clamp them, and the clamp is idempotent), so the
        // forward pass integrates s_u as it is.
        // A: forward pass of the dynamics, every state into the state table, as
        // block-wide scans over the steps (lane t holds step t).  Each
        // recurrence is associative in exact arithmetic: the body rates and the
        // velocity are affine (w' = (1 - k) w + k u, v' = (1 - dt drag) v + dt (a_T
        // R(q) e3 - g e3)), the position a prefix sum (p' = p + dt v) and the
        // attitude a product of unit quaternions (q' = q (x) rs (1, h w): the
        // renormalisation factor (3 - |q|^2)/2 of quad_step is 1 on the unit
        // attitude).  So the states are quad_step's to rounding, and the pass is
        // four scans instead of a T-step serial chain.
        {
            // lane t: the state at step t from each scan's exclusive prefix (in
            // registers, for the next scan), the state at t + 1 from the
            // inclusive one (into the table)
            const int t = tid;
            const bool in = t < T;
            float* xs = s_xf + (t + 1) * 16; // slot t + 1
            const float4 u4 = in ? s_u[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (in)
                s_cross[t] = 0u; // crossing bits of this pass (set after the scans' barriers)
            const float id4[4] = {1.0f, 0.0f, 0.0f, 0.0f};
            float wt[3], qt[4], vt[3];
            {
                float x[4] = {1.0f, 0.0f, 0.0f, 0.0f}, ex[4];
                if (in)
                {
                    x[0] = 1.0f - C.rate_k;
                    x[1] = C.rate_k * u4.y;
                    x[2] = C.rate_k * u4.z;
                    x[3] = C.rate_k * u4.w;
                }
                block_scan(x, ex, id4, AffineOp(), s_tot[0]);
                for (int i = 0; i < 3; ++i)
                    wt[i] = fma_(ex[0], x0.w[i], ex[1 + i]);
                if (in)
                    for (int i = 0; i < 3; ++i)
                        xs[10 + i] = fma_(x[0], x0.w[i], x[1 + i]);
            }
            {
                float x[4] = {1.0f, 0.0f, 0.0f, 0.0f}, ex[4];
                if (in)
                {
                    const float4 inc = attitude_increment(C, wt);
                    x[0] = inc.w;
                    x[1] = inc.w * inc.x;
                    x[2] = inc.w * inc.y;
                    x[3] = inc.w * inc.z;
                }
                block_scan(x, ex, id4, QuatOp(), s_tot[1]);
                const float q0[4] = {x0.q[0], x0.q[1], x0.q[2], x0.q[3]};
                QuatOp()(q0, ex);
                for (int i = 0; i < 4; ++i)
                    qt[i] = ex[i];
                if (in)
                {
                    QuatOp()(q0, x);
                    for (int i = 0; i < 4; ++i)
                        xs[3 + i] = x[i];
                }
            }
            if (in && it < n_steps)
            {
                // the Jacobian blocks of step t for the reverse pass
                QState xq;
                for (int c = 0; c < 4; ++c)
                    xq.q[c] = qt[c];
                for (int c = 0; c < 3; ++c)
                    xq.w[c] = wt[c];
                const float uc[4] = {u4.x, u4.y, u4.z, u4.w};
                float J[REF_JAC4 * 4];
                quad_step_jacobian(C, xq, uc, J);
                J[REF_JAC4 * 4 - 1] = 0.0f;
                for (int k = 0; k < REF_JAC4; ++k)
                    s_jac[t * REF_JAC4 + k] = make_float4(J[4 * k], J[4 * k + 1], J[4 * k + 2], J[4 * k + 3]);
            }
            {
                float x[4] = {1.0f, 0.0f, 0.0f, 0.0f}, ex[4];
                if (in)
                {
                    float b3[3];
                    thrust_dir(qt, b3);
                    const float at = u4.x * C.inv_mass;
                    x[0] = fma_(-C.dt, C.drag, 1.0f);
                    x[1] = C.dt * (at * b3[0]);
                    x[2] = C.dt * (at * b3[1]);
                    x[3] = C.dt * fma_(at, b3[2], -C.gravity);
                }
                block_scan(x, ex, id4, AffineOp(), s_tot[2]);
                for (int i = 0; i < 3; ++i)
                    vt[i] = fma_(ex[0], x0.v[i], ex[1 + i]);
                if (in)
                    for (int i = 0; i < 3; ++i)
                        xs[7 + i] = fma_(x[0], x0.v[i], x[1 + i]);
            }
            {
                const float id3[3] = {0.0f, 0.0f, 0.0f};
                float x[3] = {0.0f, 0.0f, 0.0f}, ex[3];
                if (in)
                    for (int i = 0; i < 3; ++i)
                        x[i] = C.dt * vt[i];
                block_scan(x, ex, id3, AddOp(), s_tot[3]);
                if (in)
                    for (int i = 0; i < 3; ++i)
                        xs[i] = x0.p[i] + x[i];
            }
            if (t == 0)
            {
                for (int i = 0; i < 3; ++i)
                {
                    s_xf[i] = x0.p[i];
                    s_xf[7 + i] = x0.v[i];
                    s_xf[10 + i] = x0.w[i];
                }
                for (int i = 0; i < 4; ++i)
                    s_xf[3 + i] = x0.q[i];
            }
            __syncthreads();
        }
        // Per step, all lanes: every (step, gate) pair's crossing test of the
        // move x_t -> x_{t+1}; the running cost of x_{t+1} without the winner
        // test, and its dJ/dp (not needed on the last pass).  Both read the
        // position only.
        {
            const int ng = s_track.n_gates;
            for (int i = tid; i < T * ng; i += REF_BLOCK)
            {
                const int t = i / ng, gi = i - t * ng;
                const float p0[3] = {s_xf[t * 16], s_xf[t * 16 + 1], s_xf[t * 16 + 2]};
                const float p1[3] = {s_xf[(t + 1) * 16], s_xf[(t + 1) * 16 + 1], s_xf[(t + 1) * 16 + 2]};
                if (gate_crossed(s_track.gates[gi], margin, p0, p1))
                    atomicOr(&s_cross[t], 1u << gi);
            }
            const int Tw = (T + WAVE - 1) & ~(WAVE - 1); // cost and gradient lanes in separate waves
            for (int i = tid; i < 2 * Tw; i += REF_BLOCK)
            {
                const int kind = i / Tw, t = i - kind * Tw;
                if (t >= T || (kind > 0 && it == n_steps))
                    continue;
                const bool grad = kind == 1;
                QState x;
                x.p[0] = s_xf[(t + 1) * 16];
                x.p[1] = s_xf[(t + 1) * 16 + 1];
                x.p[2] = s_xf[(t + 1) * 16 + 2];
                const float tau = C.dt * (float)(t + 1);
                const auto opp_at = [&](int j) {
                    const float4 o = s_op[j], ov = s_ov[j];
                    return make_float4(fma_(ov.x, tau, o.x), fma_(ov.y, tau, o.y), fma_(ov.z, tau, o.z), 0.0f);
                };
                if (!grad)
                {
                    int term;
                    s_cb[t] = stage_cost_body_at(C, Q, KC, s_track, x, n_opp, opp_at, &term);
                    s_tb[t] = term;
                }
                else
                {
                    float gp[3] = {0.0f, 0.0f, 0.0f};
                    stage_cost_grad_at(C, Q, KC, s_track, x, n_opp, opp_at, s_w[t], gp);
                    s_gp[t].x = gp[0];
                    s_gp[t].y = gp[1];
                    s_gp[t].z = gp[2];
                }
            }
        }
        __syncthreads();
        if (tid == 0)
            s_tend = T - 1;
        // G: gate counter.  The crossing test of a step's move reads only the
        // two positions and the gate, so every (step, gate) pair was tested
        // during the forward pass.  Wave 0 runs the counter: from
        // step t the next crossing is the first t' >= t whose move crosses the
        // next gate (the counter cannot change in between), found with one
        // ballot per 64 steps; the crossings are listed, and every state slot
        // takes the gate and laps after the crossings before it -- exactly the
        // per-step update (next gate = gate + 1 mod n, a crossing of gate 0
        // counts a lap).
        if (tid < WAVE)
        {
            const int ng = s_track.n_gates;
            int gate = x0.gate, laps = x0.laps, t = 0, n_x = 0;
            while (t < T)
            {
                const int nx = gate + 1 == ng ? 0 : gate + 1;
                int found = -1;
                for (int c0 = t & ~(WAVE - 1); c0 < T && found < 0; c0 += WAVE)
                {
                    const int tt = c0 + tid;
                    const bool hit = tt >= t && tt < T && ((s_cross[tt] >> nx) & 1u) != 0u;
                    const unsigned long long m = __ballot(hit);
                    if (m)
                        found = c0 + __ffsll((long long)m) - 1;
                }
                if (found < 0)
                    break;
                gate = nx;
                laps += nx == 0 ? 1 : 0;
                if (tid == 0)
                {
                    s_xstep[n_x] = found;
                    s_xgate[n_x] = gate;
                    s_xlaps[n_x] = laps;
                }
                ++n_x;
                t = found + 1;
            }
            if (tid == 0)
                s_nx = n_x;
        }
        __syncthreads();
        // C: every state slot k takes the gate and laps after the crossings
        // before it (slot t + 1 by the lane of step t, slot 0 by lane 0); the
        // winner test on the running cost of every x_{t+1}, the first terminal
        // step, and the discounted cost for J; the last lane makes the terminal
        // cost of x_T and its dJ/dp meanwhile (the trajectory that runs the
        // whole horizon, the usual case)
        const auto gate_laps = [&](int k, int& gate, int& laps) {
            gate = x0.gate;
            laps = x0.laps;
            for (int i = 0; i < s_nx && s_xstep[i] < k; ++i)
            {
                gate = s_xgate[i];
                laps = s_xlaps[i];
            }
        };
        if (tid == 0)
        {
            s_x[3].y = u2f((uint32_t)x0.gate);
            s_x[3].z = u2f((uint32_t)x0.laps);
        }
        for (int t = tid; t < T; t += REF_BLOCK)
        {
            int gate, laps;
            gate_laps(t + 1, gate, laps);
            s_x[(t + 1) * 4 + 3].y = u2f((uint32_t)gate);
            s_x[(t + 1) * 4 + 3].z = u2f((uint32_t)laps);
            int term = s_tb[t];
            const float c = stage_cost_win(KC, s_track, laps, s_cb[t], &term);
            if (term != TERM_NONE)
                atomicMin(&s_tend, t);
            s_gp[t].w = c * s_w[t]; // discounted running cost
        }
        if (tid == REF_BLOCK - 1)
        {
            // x_T's gate and laps made here (lane T - 1 writes its slot in this phase)
            float w;
            QState s = get_state(s_x, T, 1, &w);
            gate_laps(T, s.gate, s.laps);
            s_fin = final_cost(KC, s_track, s);
            if (it < n_steps)
            {
                float g[3] = {0.0f, 0.0f, 0.0f};
                final_cost_grad(KC, s_track, s, g);
                for (int i = 0; i < 3; ++i)
                    s_fgrad[i] = g[i];
            }
        }
        __syncthreads();
        const int te = s_tend;
        const int R = (te + REF_SEGS) / REF_SEGS; // ceil((te + 1) / REF_SEGS)
        const int n_seg = (te + R) / R;           // ceil((te + 1) / R)
        // J of this sequence in the last wave (running costs up to te summed in
        // FP64, then the terminal cost; the plain MPPI update, pass 0, is the
        // default, a gradient step is kept only if its J is lower) and the
        // terminal adjoint, while waves 0-2 make the segment maps (D)
        if (tid >= REF_BLOCK - WAVE)
        {
            const int jl = tid - (REF_BLOCK - WAVE); // lane of the last wave
            double acc = 0.0;
            for (int t = jl; t <= te; t += 64)
                acc += (double)s_gp[t].w;
            acc = wave_sum(acc);
            if (jl == 0)
            {
                float w;
                const float fin = te == T - 1 ? s_fin : final_cost(KC, s_track, get_state(s_x, te + 1, 1, &w));
                const float J = (float)(acc + (double)fin);
                const bool better = it == 0 || J < s_jbest;
                if (better)
                    s_jbest = J;
                s_better = better;
            }
            if (tid == REF_BLOCK - 1 && it < n_steps)
            {
                // terminal adjoint: the final cost's dJ/dp at x_{te+1} (made in C
                // when te = T - 1) plus the running cost's there (the gradient lane
                // of step te made it from the same position, opponents and weight)
                float g[3] = {0.0f, 0.0f, 0.0f};
                if (te == T - 1)
                    for (int i = 0; i < 3; ++i)
                        g[i] = s_fgrad[i];
                else
                {
                    float w;
                    final_cost_grad(KC, s_track, get_state(s_x, te + 1, 1, &w), g);
                }
                const float4 gp = s_gp[te];
                float* lam = s_lam + (n_seg - 1) * 16;
                lam[0] = g[0] + gp.x;
                lam[1] = g[1] + gp.y;
                lam[2] = g[2] + gp.z;
                for (int i = 3; i < 13; ++i)
                    lam[i] = 0.0f;
            }
        }
        // D: per reverse segment, the affine map lam_a = Psi lam_b + psi
        else if (it < n_steps && tid < REF_COMP * n_seg)
        {
            const int seg = tid / REF_COMP, comp = tid - seg * REF_COMP;
            const int a = seg * R, b = min(a + R, te + 1);
            float lam[13];
#pragma unroll
            for (int i = 0; i < 13; ++i)
                lam[i] = (i + 1 == comp) ? 1.0f : 0.0f;
            // the step's Jacobian is loaded one step ahead (it does not depend
            // on lam); two steps per trip, alternating two register buffers
            const auto stepj = [&](int t, const float4 (&jq)[REF_JAC4]) {
                const float4 gp = s_gp[t > 0 ? t - 1 : 0];
                float gu[4];
                quad_step_jvp_t(C, reinterpret_cast<const float*>(jq), lam, gu);
                s_gu[t * REF_COMP + comp] = make_float4(gu[0], gu[1], gu[2], gu[3]);
                const bool add = comp == 0 && t >= 1;
                lam[0] = add ? lam[0] + gp.x : lam[0];
                lam[1] = add ? lam[1] + gp.y : lam[1];
                lam[2] = add ? lam[2] + gp.z : lam[2];
            };
            float4 ja[REF_JAC4], jb[REF_JAC4];
#pragma unroll
            for (int k = 0; k < REF_JAC4; ++k)
                ja[k] = s_jac[(b - 1) * REF_JAC4 + k];
            for (int t = b - 1; t >= a; t -= 2)
            {
                const int t1 = t - 1 >= a ? t - 1 : a;
#pragma unroll
                for (int k = 0; k < REF_JAC4; ++k)
                    jb[k] = s_jac[t1 * REF_JAC4 + k];
                __builtin_amdgcn_sched_barrier(0);
                stepj(t, ja);
                if (t - 1 < a)
                    break;
                const int t2 = t - 2 >= a ? t - 2 : a;
#pragma unroll
                for (int k = 0; k < REF_JAC4; ++k)
                    ja[k] = s_jac[t2 * REF_JAC4 + k];
                __builtin_amdgcn_sched_barrier(0);
                stepj(t - 1, jb);
            }
            const float* f = lam;
            float* out = s_psi + (seg * REF_COMP + comp) * 16;
#pragma unroll
            for (int i = 0; i < 13; ++i)
                out[i] = f[i];
        }
        __syncthreads();
        if (s_better)
            for (int t = tid; t < T; t += REF_BLOCK)
                s_ubest[t] = s_u[t];
        if (it == n_steps)
            break;
        // E: lam_b of every segment, last to first (wave 0; lane i = component i)
        if (tid < 64)
        {
            // lanes 13..63 compute copies of rows 0..15 (row 13..15: unused
            // padding) and are never read; every load is unconditional
            // (each segment's map is loaded one segment ahead: the loads do
            // not depend on lam)
            const int i = tid & 15;
            float lam = s_lam[(n_seg - 1) * 16 + i];
            float pc[REF_COMP];
#pragma unroll
            for (int c = 0; c < REF_COMP; ++c)
                pc[c] = s_psi[((n_seg - 1) * REF_COMP + c) * 16 + i];
            for (int sg = n_seg - 1; sg >= 1; --sg)
            {
                float pn[REF_COMP];
                const float* P = s_psi + (sg > 1 ? sg - 1 : sg) * REF_COMP * 16;
#pragma unroll
                for (int c = 0; c < REF_COMP; ++c)
                    pn[c] = P[c * 16 + i];
                __builtin_amdgcn_sched_barrier(0);
                // every lane takes the whole of lam_b from the segment's slot,
                // written by lanes 0..12 (four 16-byte LDS reads instead of
                // thirteen lane reads in a row; the slot was written by this
                // wave, so a wait on the write is all the ordering needed)
                const float4* lb = reinterpret_cast<const float4*>(s_lam + sg * 16);
                const float4 l0 = lb[0], l1 = lb[1], l2 = lb[2], l3 = lb[3];
                const float lj[13] = {l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, l1.z, l1.w, l2.x, l2.y, l2.z, l2.w, l3.x};
                float acc = pc[0];
#pragma unroll
                for (int j = 0; j < 13; ++j)
                    acc = fma_(pc[1 + j], lj[j], acc);
                lam = acc;
                if (tid < 13)
                    s_lam[(sg - 1) * 16 + tid] = lam;
                __builtin_amdgcn_s_waitcnt(0xc07f); // lgkmcnt(0): the slot is written before the next segment reads it
#pragma unroll
                for (int c = 0; c < REF_COMP; ++c)
                    pc[c] = pn[c];
            }
        }
        __syncthreads();
        // F: dJ/du_t of every step t <= te, one lane per step: the chains of
        // t's segment combined with the segment's lam_b (the vector-Jacobian
        // products are linear in lam), then the projected step of u_t
        for (int t = tid; t <= te; t += REF_BLOCK)
        {
            const float* lb = s_lam + (t / R) * 16;
            const float4* gc = s_gu + t * REF_COMP;
            float4 g = gc[0];
#pragma unroll
            for (int i = 0; i < 13; ++i)
            {
                const float li = lb[i];
                const float4 gi = gc[1 + i];
                g.x = fma_(li, gi.x, g.x);
                g.y = fma_(li, gi.y, g.y);
                g.z = fma_(li, gi.z, g.z);
                g.w = fma_(li, gi.w, g.w);
            }
            const float4 u4 = s_u[t];
            const float uin[4] = {u4.x, u4.y, u4.z, u4.w};
            const float G[4] = {g.x * clamp_grad(uin[0], C.thrust_min, C.thrust_max), g.y * clamp_grad(uin[1], -C.rate_max, C.rate_max),
                                g.z * clamp_grad(uin[2], -C.rate_max, C.rate_max), g.w * clamp_grad(uin[3], -C.rate_max, C.rate_max)};
            float in[4], out[4];
            for (int d = 0; d < 4; ++d)
                in[d] = fma_(-step, G[d], uin[d]);
            clamp_control(C, in, out);
            s_u[t] = make_float4(out[0], out[1], out[2], out[3]);
        }
        for (int t = te + 1 + tid; t < T; t += REF_BLOCK)
        {
            const float4 u4 = s_u[t];
            const float in[4] = {u4.x, u4.y, u4.z, u4.w};
            float out[4];
            clamp_control(C, in, out);
            s_u[t] = make_float4(out[0], out[1], out[2], out[3]);
        }
        __syncthreads();
    }
    __syncthreads();
    // the kept sequence's first control is the action, and the sequence
    // shifted left by one step (shift_fill at the end) the next nominal
    // (as k_update makes them)
    for (int t = tid; t < T; t += REF_BLOCK)
    {
        const float4 u4 = t + 1 < T ? s_ubest[t + 1] : make_float4(I.shift_fill[0], I.shift_fill[1], I.shift_fill[2], I.shift_fill[3]);
        float* u = nominal + I.nom_off + t * 4;
        u[0] = u4.x;
        u[1] = u4.y;
        u[2] = u4.z;
        u[3] = u4.w;
    }
    if (tid < 4)
    {
        const float4 a = s_ubest[0];
        actions[blockIdx.x * 4 + tid] = tid == 0 ? a.x : tid == 1 ? a.y : tid == 2 ? a.z : a.w;
    }
    if (tid == 0)
        dyn[blockIdx.x].step = step0 + 1u;
    if (tid == WAVE)
    {
        // Closed loop (as k_update): the drone integrates the action, the
        // drones without a controller (spread over the workgroups) keep their
        // velocity.  Other workgroups may still read this step's drones, so
        // the moved drones go to the other copy of the world.
        if (advance)
        {
            QState s = x0; // the drone as this launch read it
            const float4 a = s_ubest[0];
            const float act[4] = {a.x, a.y, a.z, a.w};
            advance_state(s, act, s_track, C, margin);
            dr_drone_state st;
            state_to_abi(s, st);
            st._pad = 0;
            drones_out[I.drone] = st;
            for (int d = blockIdx.x; d < n_drones; d += n_inst) // the refinement's workgroups only
                if (drone_inst[d] < 0)
                {
                    dr_drone_state o = drones[d];
                    for (int c = 0; c < 3; ++c)
                        o.p[c] = fma_(C.dt, o.v[c], o.p[c]);
                    drones_out[d] = o;
                }
        }
    }
}

// Weighted average over one chunk of samples [k0, k0 + n) and timesteps
// [t0, t1):
//   w_k = exp(-(c_k - min)/lambda), min over the instance's rollout tiles,
//   numerator[t][d] = sum_k w_k eps_k[t][d], denominator = sum_k w_k,
// both accumulated in FP64 so the summation order cannot move the FP32
// result.  The chunk's weights are computed once into LDS; lane (t, sub)
// then streams the float4 noise of step t of samples sub, sub+S, ... where
// S = floor(256/(t1-t0)) lanes share each step: the lanes of one sample read
// one contiguous (t1-t0)*16-byte piece of its row (coalesced) and no
// cross-lane reduction is needed until the final S-way sum.  Only the
// workgroups with t0 == 0 write the chunk's denominator.
__global__ __launch_bounds__(WAVG_BLOCK) void k_wavg(const InstDev* __restrict__ insts, InstDyn* __restrict__ dyn,
                                                     const WTile* __restrict__ wtiles, const float* __restrict__ costs,
                                                     const float4* __restrict__ noise, const float* __restrict__ tmin,
                                                     double* __restrict__ part, double* __restrict__ den, int chunk)
{
    __shared__ float s_w[WAVG_MAX_CHUNK];
    __shared__ double s_acc[WAVG_BLOCK][4];
    __shared__ float s_min[WAVG_BLOCK / WAVE];
    __shared__ double s_den[WAVG_BLOCK / WAVE];

    const int tid = threadIdx.x;
    const int wave = tid / WAVE, lane = tid & (WAVE - 1);
    const WTile wt = wtiles[blockIdx.x];
    const InstDev& I = insts[wt.inst];
    const int K = I.K, T = I.T;

    float m = __builtin_inff();
    for (int i = tid; i < I.rtile_count; i += WAVG_BLOCK)
        m = fminf(m, tmin[I.rtile_first + i]);
    m = wave_min(m);
    if (lane == 0)
        s_min[wave] = m;
    __syncthreads();
    float minC = s_min[0];
    for (int w = 1; w < WAVG_BLOCK / WAVE; ++w)
        minC = fminf(minC, s_min[w]);

    const float lam = dyn[wt.inst].lambda;
    const int k0 = wt.chunk * chunk;
    const int n = min(chunk, K - k0);
    double dsum = 0.0;
    for (int i = tid; i < n; i += WAVG_BLOCK)
    {
        float w = 0.0f;
        if (__builtin_isfinite(minC))
            w = expf_ieee(-(costs[I.cost_off + k0 + i] - minC) / lam);
        s_w[i] = w;
        dsum += (double)w;
    }
    dsum = wave_sum(dsum);
    if (lane == 0)
        s_den[wave] = dsum;
    __syncthreads();
    if (tid == 0 && wt.t0 == 0)
    {
        double d = s_den[0];
        for (int w = 1; w < WAVG_BLOCK / WAVE; ++w)
            d += s_den[w];
        den[I.den_off + wt.chunk] = d;
        if (wt.chunk == 0)
        {
            dyn[wt.inst].min_cost = minC;
            dyn[wt.inst].step_ahead = dyn[wt.inst].step + 1u;
        }
    }

    const int TG = wt.t1 - wt.t0;
    const int S = WAVG_BLOCK / TG;
    const int sub = tid / TG, tt = tid - sub * TG;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    if (sub < S)
    {
        const float4* base = noise + I.noise_off + (long long)k0 * T + wt.t0 + tt;
#pragma unroll 8
        for (int i = sub; i < n; i += S)
        {
            const float4 e = base[(long long)i * T];
            const double w = (double)s_w[i];
            a0 += w * (double)e.x;
            a1 += w * (double)e.y;
            a2 += w * (double)e.z;
            a3 += w * (double)e.w;
        }
    }
    s_acc[tid][0] = a0;
    s_acc[tid][1] = a1;
    s_acc[tid][2] = a2;
    s_acc[tid][3] = a3;
    __syncthreads();
    double* out = part + I.part_off + ((long long)wt.chunk * T + wt.t0) * 4;
    for (int i = tid; i < TG * 4; i += WAVG_BLOCK)
    {
        const int r = i >> 2, d = i & 3;
        double v = s_acc[r][d];
        for (int q = 1; q < S; ++q)
            v += s_acc[q * TG + r][d];
        out[i] = v;
    }
}

// Per instance: numerator and denominator over chunks in fixed order,
// nominal += num/den when den > 1e-30 (kernels.cu:402-403), box clamp of the
// whole nominal (kernels.cu:434-456), action = first control, shift left by
// one step with shift_fill at the end (mppi_controller.cu:430-436), nu and
// the optional temperature update (mppi_controller.cu:413-424).  With
// advance != 0 the closed loop follows (SimulationEngine::dynStep,
// engine.cpp:286-312): the instance's drone integrates its action with the
// same dynamics and gate logic, noise-free, and drones without a controller
// (distributed over the workgroups) keep their velocity.  With the adjoint
// refinement on, k_refine does all of this itself.
__global__ __launch_bounds__(512) void k_update(const InstDev* __restrict__ insts, InstDyn* __restrict__ dyn,
                                                const double* __restrict__ part, const double* __restrict__ den,
                                                float* __restrict__ nominal, float* __restrict__ actions, QuadConst C, int advance,
                                                dr_drone_state* __restrict__ drones, const int* __restrict__ drone_inst, int n_drones,
                                                const dr_track* __restrict__ tracks, float margin)
{
    __shared__ float s_u[DR_MAX_T * 4];
    const int inst = blockIdx.x;
    const InstDev& I = insts[inst];
    const int T = I.T, tid = threadIdx.x;

    const double dn = sum_in_order(den + I.den_off, I.n_chunks, 1);
    for (int i = tid; i < T * 4; i += blockDim.x)
    {
        const double num = sum_in_order(part + I.part_off + i, I.n_chunks, (long long)T * 4);
        float u = nominal[I.nom_off + i];
        if (dn > 1e-30)
            u = u + (float)(num / dn);
        s_u[i] = u;
    }
    __syncthreads();
    for (int t = tid; t < T; t += blockDim.x)
    {
        float in[4] = {s_u[t * 4 + 0], s_u[t * 4 + 1], s_u[t * 4 + 2], s_u[t * 4 + 3]};
        float out[4];
        clamp_control(C, in, out);
        for (int d = 0; d < 4; ++d)
            s_u[t * 4 + d] = out[d];
    }
    __syncthreads();
    if (tid < 4)
        actions[inst * 4 + tid] = s_u[tid];
    for (int i = tid; i < T * 4; i += blockDim.x)
    {
        const int t = i / 4, d = i & 3;
        nominal[I.nom_off + i] = (t + 1 < T) ? s_u[i + 4] : I.shift_fill[d];
    }
    if (tid == 0)
    {
        InstDyn& Y = dyn[inst];
        const float nu = (float)dn;
        Y.nu = nu;
        if (I.adapt)
        {
            if (nu > 0.001f * (float)I.K)
                Y.lambda = Y.lambda * 0.9f;
            else if (nu < 0.0005f * (float)I.K)
                Y.lambda = Y.lambda * 1.2f;
        }
        Y.step = Y.step + 1u;
        if (advance)
        {
            const float act[4] = {s_u[0], s_u[1], s_u[2], s_u[3]};
            advance_drone(drones[I.drone], act, tracks[I.track], C, margin);
            for (int d = inst; d < n_drones; d += gridDim.x)
                if (drone_inst[d] < 0)
                    for (int c = 0; c < 3; ++c)
                        drones[d].p[c] = fma_(C.dt, drones[d].v[c], drones[d].p[c]);
        }
    }
}

// Quality of the executed closed loop: after the drones have moved, the
// running cost of each controlled drone's new state (stage_cost, the other
// drones of its world where they now are), summed undiscounted over the
// steps, the number of those states in collision, and the race progress of
// the latest one.  One lane per instance.
__global__ __launch_bounds__(64) void k_exec(const InstDev* __restrict__ insts, InstDyn* __restrict__ dyn, int n_inst,
                                             const dr_track* __restrict__ tracks, const dr_drone_state* __restrict__ drones,
                                             QuadConst C, CostConst Q, dr_cost_cfg KC)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst)
        return;
    const InstDev& I = insts[i];
    const dr_track& tr = tracks[I.track];
    const QState s = state_from_abi(drones[I.drone]);
    const auto opp_at = [&](int j) {
        const dr_drone_state& o = drones[I.opp[j]];
        return make_float4(o.p[0], o.p[1], o.p[2], 0.0f);
    };
    int term;
    const float c = stage_cost_at(C, Q, KC, tr, s, I.n_opp, opp_at, &term);
    InstDyn& Y = dyn[i];
    Y.exec_cost = Y.exec_cost + c;
    Y.exec_progress = advance(tr.gates, tr.n_gates, s);
    if (term == TERM_LOSE)
        Y.exec_collisions = Y.exec_collisions + 1u;
}

// ---------------------------------------------------------------------------
// launch wrappers (host)

template <int BLOCK>
static void rollout_nb(int nb, int g, size_t shm, hipStream_t st, const InstDev* insts, InstDyn* dyn, const RTile* tiles, int n_tiles,
                       const dr_track* tracks, const dr_drone_state* drones, const float* nominal, float* costs, float4* noise,
                       float* tmin, const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC, int opp_stride)
{
    switch (nb)
    {
    case 1:
        k_rollout<BLOCK, 1><<<g, BLOCK, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                   opp_stride);
        break;
    case 2:
        k_rollout<BLOCK, 2><<<g, BLOCK, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                   opp_stride);
        break;
    default:
        k_rollout<BLOCK, 4><<<g, BLOCK, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                   opp_stride);
        break;
    }
}

hipError_t launch_rollout(int block, int nb, int split, int grid, hipStream_t st, const InstDev* insts, InstDyn* dyn,
                          const RTile* tiles, int n_tiles, const dr_track* tracks, const dr_drone_state* drones, const float* nominal,
                          float* costs, float4* noise, float* tmin, const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC,
                          int max_T, int opp_stride)
{
    if (nb != 1 && nb != 2 && nb != 4)
        return hipErrorInvalidValue;
    const int g = grid > 0 ? (grid < n_tiles ? grid : n_tiles) : n_tiles;
    const size_t shm = (size_t)max_T * opp_stride * sizeof(float4);
    if (split == 4)
    {
        // tiles of 64 samples, eight waves, over the noise rows made ahead;
        // chunk length = 4 * noise_batch steps
        if (nb == 1)
            k_rollout_p8nz<4><<<g, 512, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                   opp_stride);
        else if (nb == 2)
            k_rollout_p8nz<8><<<g, 512, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                   opp_stride);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    }
    if (split == 3)
    {
        // tiles of 64 samples, eight stage waves; chunk length = 2 * noise_batch steps
        switch (nb)
        {
        case 1:
            k_rollout_p8<2><<<g, 512, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                 opp_stride);
            break;
        case 2:
            k_rollout_p8<4><<<g, 512, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                 opp_stride);
            break;
        default:
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if (split == 2)
    {
        // tiles of 64 samples, four stage waves; chunk length = 2 * noise_batch steps
        switch (nb)
        {
        case 1:
            k_rollout_p4<2><<<g, 256, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                 opp_stride);
            break;
        case 2:
            k_rollout_p4<4><<<g, 256, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                 opp_stride);
            break;
        default:
            k_rollout_p4<8><<<g, 256, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                 opp_stride);
            break;
        }
        return hipGetLastError();
    }
    if (split)
    {
        // tiles of 64 samples; chunk length = 4 * noise_batch steps
        switch (nb)
        {
        case 1:
            k_rollout_ws<4><<<g, 128, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                 opp_stride);
            break;
        case 2:
            k_rollout_ws<8><<<g, 128, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                 opp_stride);
            break;
        default:
            k_rollout_ws<16><<<g, 128, shm, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC,
                                                  opp_stride);
            break;
        }
        return hipGetLastError();
    }
    switch (block)
    {
    case 64:
        rollout_nb<64>(nb, g, shm, st, insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC, opp_stride);
        break;
    case 128:
        rollout_nb<128>(nb, g, shm, st, insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC, opp_stride);
        break;
    case 256:
        rollout_nb<256>(nb, g, shm, st, insts, dyn, tiles, n_tiles, tracks, drones, nominal, costs, noise, tmin, C, Q, KC, opp_stride);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_exec(int n_inst, hipStream_t st, const InstDev* insts, InstDyn* dyn, const dr_track* tracks,
                       const dr_drone_state* drones, const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC)
{
    hipLaunchKernelGGL(k_exec, dim3((n_inst + 63) / 64), dim3(64), 0, st, insts, dyn, n_inst, tracks, drones, C, Q, KC);
    return hipGetLastError();
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
    k_update<<<n_inst, 512, 0, st>>>(insts, dyn, part, den, nominal, actions, C, advance, drones, drone_inst, n_drones, tracks, margin);
    return hipGetLastError();
}

// The refinement's dynamic LDS exceeds the default limit at long horizons:
// allowed up to DR_MAX_T on the current device (once per engine).
hipError_t refine_allow_lds()
{
    return hipFuncSetAttribute((const void*)k_refine, hipFuncAttributeMaxDynamicSharedMemorySize, (int)refine_lds_bytes(DR_MAX_T));
}

hipError_t launch_noise(int blocks, hipStream_t st, const InstDev* insts, const InstDyn* dyn, int n_inst, float4* noise)
{
    k_noise<<<blocks, 256, 0, st>>>(insts, dyn, n_inst, noise);
    return hipGetLastError();
}

hipError_t launch_costgrad(int lanes, hipStream_t st, const InstDev* insts, int inst, int T, int n_ck, int ck,
                           const dr_track* tracks, const dr_drone_state* drones, const float* U, int n_seq, float* J, float* G,
                           const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC, int opp_stride)
{
    const size_t shm = ((size_t)T * opp_stride + (size_t)lanes * (n_ck + ck) * 4) * sizeof(float4);
    k_costgrad<<<(n_seq + lanes - 1) / lanes, lanes, shm, st>>>(insts, inst, tracks, drones, U, n_seq, J, G, ck, C, Q, KC, opp_stride);
    return hipGetLastError();
}

hipError_t launch_refine(int n_inst, hipStream_t st, const InstDev* insts, InstDyn* dyn, const double* part, const double* den,
                         const dr_track* tracks, const dr_drone_state* drones, float* nominal, int n_steps, float step,
                         const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC, float* actions, int advance,
                         const int* drone_inst, int n_drones, dr_drone_state* drones_out, int n_cu, size_t lds_per_cu, float4* noise,
                         int tcap)
{
    if (tcap < 1 || tcap > DR_MAX_T)
        return hipErrorInvalidValue;
    // With noise made ahead (noise != nullptr), the workgroups past the
    // instances make the next step's noise rows on the CUs the refinement
    // leaves idle: as many as fit next to it (two per CU when the LDS of two
    // workgroups fits in one CU's, one otherwise).
    int noise_blocks = 0;
    if (noise)
    {
        static size_t static_lds = 0;
        if (!static_lds)
        {
            hipFuncAttributes fa;
            if (hipFuncGetAttributes(&fa, (const void*)k_refine) != hipSuccess)
                return hipErrorInvalidValue;
            static_lds = fa.sharedSizeBytes;
        }
        const size_t per_wg = static_lds + refine_lds_bytes(tcap);
        const int per_cu = 2 * per_wg <= lds_per_cu ? 2 : 1;
        noise_blocks = per_cu * n_cu > n_inst + 64 ? per_cu * n_cu - n_inst : 64;
    }
    k_refine<<<n_inst + noise_blocks, REF_BLOCK, refine_lds_bytes(tcap), st>>>(insts, dyn, part, den, tracks, drones, nominal, n_steps,
                                                                               step, C, Q, KC, actions, advance, drone_inst, n_drones,
                                                                               drones_out, n_inst, noise, tcap);
    return hipGetLastError();
}

} // namespace dr
