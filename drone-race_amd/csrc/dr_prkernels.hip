// PR-MPPI control-step kernels for gfx950 (wave64).  One step of a batch:
//   k_pr_rollout -> k_pr_agg -> k_pr_wavg -> k_pr_update -> k_pr_advance
// each covering every instance of the batch.
//
// Reference kernels these replace (src/prmppi_kernels.cu):
//   PRMPPIgenerateNoiseKernel :8-20 + PRMPPIsampleThetaValues :23-51
//     + PRMPPIfullRolloutKernel :55-130                    -> k_pr_rollout
//   PRMPPIcostAvgKernel :134-166 + computeMinCosts :169-199 -> k_pr_agg (+ minima in k_pr_wavg)
//   PRMPPIWeightedAverageKernel :204-282                   -> k_pr_wavg + k_pr_update
//   PRMPPIcomputeCandidateCostKernel :285-347, PRMPPIcomputeSafeCostKernel
//   :350-393 and the host choice prmppi_controller.cu:227-346 -> k_pr_update
// and the closed-loop step (SimulationEngine::dynStep, engine.cpp:286-312) -> k_pr_advance.
#include <hip/hip_runtime.h>

#include "dr_prcommon.h"
#include "dr_prkernels.h"
#include "dr_wave.h"

namespace dr {

// Opponent j of one model at one step: p + (v + shift) * tau, with the
// shifted velocity staged once per workgroup.
struct ModelOpp
{
    const float4 (*pv)[2];
    float tau;
    __device__ float4 operator()(int j) const
    {
        const float4 p = pv[j][0], v = pv[j][1];
        return make_float4(fma_(v.x, tau, p.x), fma_(v.y, tau, p.y), fma_(v.z, tau, p.z), 0.0f);
    }
};

__device__ __forceinline__ void stage_model_opp(float4 (*dst)[2], const dr_drone_state& o, const PrInstDev& I, int b, int j)
{
    dst[j][0] = make_float4(o.p[0], o.p[1], o.p[2], 0.0f);
    dst[j][1] = make_float4(o.v[0] + I.shift[b][j][0], o.v[1] + I.shift[b][j][1], o.v[2] + I.shift[b][j][2], 0.0f);
}

// A workgroup of W waves covers 64 samples of one instance and up to W of
// its 2*M rollouts per sample (wave w: plan (pair0 + w) / M, model
// (pair0 + w) % M).  The noise of the 64 samples is made once per chunk of
// PR_CH steps into a double-buffered LDS ring, each wave producing some of
// the chunk's steps (Philox + Box-Muller), and read by every rollout wave:
// the M models and both plans see the same noise (common random numbers,
// prmppi_kernels.cu:8-20, :107).  The workgroup with pair0 == 0 also writes
// the noise rows to HBM for the weighted averages.  One barrier per chunk;
// at chunk i the waves produce chunk i and integrate chunk i - 1.
__global__ __launch_bounds__(1024) void k_pr_rollout(const PrInstDev* __restrict__ insts, const PrDyn* __restrict__ dyn,
                                                     const PrTile* __restrict__ tiles, int n_tiles, const dr_track* __restrict__ tracks,
                                                     const dr_drone_state* __restrict__ drones, const float* __restrict__ work,
                                                     float2* __restrict__ pair, float4* __restrict__ noise, QuadConst C, CostConst Q,
                                                     dr_cost_cfg KC)
{
    __shared__ float4 s_eps[2][PR_CH][WAVE];
    __shared__ float s_nom[2 * DR_MAX_T * 4];
    __shared__ dr_track s_track;
    __shared__ float4 s_pv[PR_MAX_WAVES][MAX_OPP][2];

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1), wave = tid / WAVE;
    const int W = (int)blockDim.x / WAVE;
    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    {
        const PrTile pt = tiles[tile];
        const PrInstDev& I = insts[pt.inst];
        const PrDyn& Y = dyn[pt.inst];
        const int K = I.K, T = I.T, M = I.M, n_opp = I.n_opp;
        const uint32_t step = Y.step;
        const uint32_t key0 = I.key0, key1 = I.key1;
        const float sg0 = I.sigma[0], sg1 = I.sigma[1], sg2 = I.sigma[2], sg3 = I.sigma[3];

        __syncthreads(); // LDS of the previous tile is no longer read
        for (int i = tid; i < 2 * T * 4; i += blockDim.x)
            s_nom[i] = work[I.plan_off + i];
        {
            const float* src = reinterpret_cast<const float*>(tracks + I.track);
            float* dst = reinterpret_cast<float*>(&s_track);
            for (int i = tid; i < (int)(sizeof(dr_track) / 4); i += blockDim.x)
                dst[i] = src[i];
        }
        const bool rolls = wave < pt.npairs;
        const int pidx = pt.pair0 + wave;
        if (rolls && lane < n_opp)
            stage_model_opp(s_pv[wave], drones[I.opp[lane]], I, model_branch(I, Y, step, pidx % M), lane);
        __syncthreads();

        const int k = pt.k0 + lane;
        const bool live = k < K;
        const bool write_noise = pt.pair0 == 0;
        const int n_chunks = (T + PR_CH - 1) / PR_CH;
        float4* nrow = noise + I.noise_off + (long long)k * T;
        const float* nomp = s_nom + (rolls ? (pidx / M) : 0) * T * 4;

        QState s;
        float acc = 0.0f, decay = 1.0f, smax = -__builtin_inff();
        bool stop = false;
        if (rolls && live)
            s = state_from_abi(drones[I.drone]);

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
                        float z[4];
                        normal4(key0, key1, (uint32_t)k, (uint32_t)t, step, STREAM_CONTROL, z);
                        const float4 e = make_float4(sg0 * z[0], sg1 * z[1], sg2 * z[2], sg3 * z[3]);
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
#pragma unroll
                for (int u = 0; u < PR_CH; ++u)
                {
                    const int t = j * PR_CH + u;
                    if (t < T && !stop)
                    {
                        const float4 e = ring[u * WAVE + lane];
                        float uin[4], uc[4];
                        uin[0] = nomp[t * 4 + 0] + e.x;
                        uin[1] = nomp[t * 4 + 1] + e.y;
                        uin[2] = nomp[t * 4 + 2] + e.z;
                        uin[3] = nomp[t * 4 + 3] + e.w;
                        clamp_control(C, uin, uc);
                        const float p0[3] = {s.p[0], s.p[1], s.p[2]};
                        quad_step(C, s, uc);
                        update_gate(s_track.gates, s_track.n_gates, KC.gate_traversal_margin, p0, s);
                        const ModelOpp opp{s_pv[wave], C.dt * (float)(t + 1)};
                        float sf;
                        int term;
                        const float c = pr_stage_cost(C, Q, KC, s_track, s, n_opp, opp, &sf, &term);
                        acc = fma_(c, decay, acc);
                        smax = fmaxf(smax, sf);
                        decay = decay * DECAY;
                        stop = (term != TERM_NONE);
                    }
                }
            }
            __syncthreads();
        }
        if (rolls && live)
        {
            acc = acc + final_cost(KC, s_track, s);
            pair[I.pair_off + (long long)pidx * K + k] = make_float2(acc, smax);
        }
    }
}

// Per sample: model average of each plan's costs plus safety_weight when any
// model is unsafe, and the max safety (prmppi_kernels.cu:134-166); per
// workgroup the minima of nominal/full, robust/full and robust/safety
// (the first stage of prmppi_kernels.cu:169-199).
__global__ __launch_bounds__(PR_AGG_BLOCK) void k_pr_agg(const PrInstDev* __restrict__ insts, const PrDyn* __restrict__ dyn,
                                                         const PrATile* __restrict__ atiles, const float2* __restrict__ pair,
                                                         float4* __restrict__ agg, float4* __restrict__ ablk)
{
    __shared__ float s_m[3][PR_AGG_BLOCK / WAVE];
    const PrATile at = atiles[blockIdx.x];
    const PrInstDev& I = insts[at.inst];
    const PrDyn& Y = dyn[at.inst];
    const int K = I.K, M = I.M;
    const int tid = threadIdx.x, k = at.k0 + tid;
    float m0 = __builtin_inff(), m1 = __builtin_inff(), m2 = __builtin_inff();
    if (k < K)
    {
        float full[2], smax[2];
        for (int plan = 0; plan < 2; ++plan)
        {
            float avg = 0.0f, mx = -__builtin_inff();
            const float2* src = pair + I.pair_off + (long long)plan * M * K + k;
            for (int m = 0; m < M; ++m)
            {
                const float2 v = src[(long long)m * K];
                if (I.sampled)
                    avg = avg + v.x;
                else
                    avg = fma_(Y.belief[m], v.x, avg);
                mx = fmaxf(mx, v.y);
            }
            if (I.sampled)
                avg = avg / (float)M;
            full[plan] = mx > 0.0f ? avg + I.safety_weight : avg;
            smax[plan] = mx;
        }
        agg[I.agg_off + k] = make_float4(full[0], smax[0], full[1], smax[1]);
        m0 = full[0];
        m1 = full[1];
        m2 = smax[1];
    }
    m0 = wave_min(m0);
    m1 = wave_min(m1);
    m2 = wave_min(m2);
    if ((tid & (WAVE - 1)) == 0)
    {
        s_m[0][tid / WAVE] = m0;
        s_m[1][tid / WAVE] = m1;
        s_m[2][tid / WAVE] = m2;
    }
    __syncthreads();
    if (tid == 0)
    {
        float r[3];
        for (int j = 0; j < 3; ++j)
        {
            r[j] = s_m[j][0];
            for (int w = 1; w < PR_AGG_BLOCK / WAVE; ++w)
                r[j] = fminf(r[j], s_m[j][w]);
        }
        ablk[blockIdx.x] = make_float4(r[0], r[1], r[2], 0.0f);
    }
}

// The three weighted averages of prmppi_kernels.cu:204-282 over one chunk
// of samples and a group of steps, reading each noise row once:
//   set 0: w = exp(-(full_nom - min0)/lambda0), u = nom + eps   -> candidate 1
//   set 1: w = exp(-(full_rob - min1)/lambda1), u = rob + eps   -> candidate 2
//   set 2: w = exp(-(safe_rob - min2)/lambda2), u = rob + eps   -> new robust plan
// Numerators sum w*u and denominators sum w in FP64 (partials per chunk).
__global__ __launch_bounds__(PR_WAVG_BLOCK) void k_pr_wavg(const PrInstDev* __restrict__ insts, PrDyn* __restrict__ dyn,
                                                           const WTile* __restrict__ wtiles, const float4* __restrict__ agg,
                                                           const float4* __restrict__ noise, const float4* __restrict__ ablk,
                                                           const float* __restrict__ work, double* __restrict__ part,
                                                           double* __restrict__ den, int chunk)
{
    __shared__ float s_w[3][PR_WAVG_MAX_CHUNK];
    __shared__ double s_acc[PR_WAVG_BLOCK][12];
    __shared__ float s_min[3][PR_WAVG_BLOCK / WAVE];
    __shared__ double s_den[3][PR_WAVG_BLOCK / WAVE];

    const int tid = threadIdx.x;
    const int wave = tid / WAVE, lane = tid & (WAVE - 1);
    const WTile wt = wtiles[blockIdx.x];
    const PrInstDev& I = insts[wt.inst];
    const int K = I.K, T = I.T;

    float m0 = __builtin_inff(), m1 = __builtin_inff(), m2 = __builtin_inff();
    for (int i = tid; i < I.ablk_count; i += PR_WAVG_BLOCK)
    {
        const float4 b = ablk[I.ablk_first + i];
        m0 = fminf(m0, b.x);
        m1 = fminf(m1, b.y);
        m2 = fminf(m2, b.z);
    }
    m0 = wave_min(m0);
    m1 = wave_min(m1);
    m2 = wave_min(m2);
    if (lane == 0)
    {
        s_min[0][wave] = m0;
        s_min[1][wave] = m1;
        s_min[2][wave] = m2;
    }
    __syncthreads();
    float mins[3];
    for (int j = 0; j < 3; ++j)
    {
        mins[j] = s_min[j][0];
        for (int w = 1; w < PR_WAVG_BLOCK / WAVE; ++w)
            mins[j] = fminf(mins[j], s_min[j][w]);
    }
    const PrDyn& Y = dyn[wt.inst];
    const float lam[3] = {Y.lambda[0], Y.lambda[1], Y.lambda[2]};
    const int k0 = wt.chunk * chunk;
    const int n = min(chunk, K - k0);
    double ds[3] = {0.0, 0.0, 0.0};
    for (int i = tid; i < n; i += PR_WAVG_BLOCK)
    {
        const float4 a = agg[I.agg_off + k0 + i];
        const float c[3] = {a.x, a.z, a.w};
        for (int j = 0; j < 3; ++j)
        {
            float w = 0.0f;
            if (__builtin_isfinite(mins[j]))
                w = expf_ieee(-(c[j] - mins[j]) / lam[j]);
            s_w[j][i] = w;
            ds[j] += (double)w;
        }
    }
    for (int j = 0; j < 3; ++j)
    {
        const double v = wave_sum(ds[j]);
        if (lane == 0)
            s_den[j][wave] = v;
    }
    __syncthreads();
    if (tid < 3 && wt.t0 == 0)
    {
        double d = s_den[tid][0];
        for (int w = 1; w < PR_WAVG_BLOCK / WAVE; ++w)
            d += s_den[tid][w];
        den[I.den_off + wt.chunk * 3 + tid] = d;
        if (wt.chunk == 0)
            dyn[wt.inst].min_cost[tid] = mins[tid];
    }

    const int TG = wt.t1 - wt.t0;
    const int S = PR_WAVG_BLOCK / TG;
    const int sub = tid / TG, tt = tid - sub * TG;
    double a[12];
#pragma unroll
    for (int q = 0; q < 12; ++q)
        a[q] = 0.0;
    if (sub < S)
    {
        const int t = wt.t0 + tt;
        const float* nv = work + I.plan_off + t * 4;
        const float* rv = work + I.plan_off + T * 4 + t * 4;
        const float n0 = nv[0], n1 = nv[1], n2 = nv[2], n3 = nv[3];
        const float r0 = rv[0], r1 = rv[1], r2 = rv[2], r3 = rv[3];
        const float4* base = noise + I.noise_off + (long long)k0 * T + t;
#pragma unroll 4
        for (int i = sub; i < n; i += S)
        {
            const float4 e = base[(long long)i * T];
            const double w0 = (double)s_w[0][i], w1 = (double)s_w[1][i], w2 = (double)s_w[2][i];
            const double u0 = (double)(n0 + e.x), u1 = (double)(n1 + e.y), u2 = (double)(n2 + e.z), u3 = (double)(n3 + e.w);
            const double v0 = (double)(r0 + e.x), v1 = (double)(r1 + e.y), v2 = (double)(r2 + e.z), v3 = (double)(r3 + e.w);
            a[0] += w0 * u0;
            a[1] += w0 * u1;
            a[2] += w0 * u2;
            a[3] += w0 * u3;
            a[4] += w1 * v0;
            a[5] += w1 * v1;
            a[6] += w1 * v2;
            a[7] += w1 * v3;
            a[8] += w2 * v0;
            a[9] += w2 * v1;
            a[10] += w2 * v2;
            a[11] += w2 * v3;
        }
    }
#pragma unroll
    for (int q = 0; q < 12; ++q)
        s_acc[tid][q] = a[q];
    __syncthreads();
    double* out = part + I.part_off + (long long)wt.chunk * 3 * T * 4;
    for (int i = tid; i < 3 * TG * 4; i += PR_WAVG_BLOCK)
    {
        const int j = i / (TG * 4);
        const int rem = i - j * TG * 4;
        const int r = rem >> 2, d = rem & 3;
        double v = s_acc[r][j * 4 + d];
        for (int q = 1; q < S; ++q)
            v += s_acc[q * TG + r][j * 4 + d];
        out[(long long)j * T * 4 + (wt.t0 + r) * 4 + d] = v;
    }
}

// Per instance: the three candidates sum(w u)/sum(w) (the shifted plan where
// sum(w) <= 1e-30, prmppi_kernels.cu:268-281), the noise-free rollouts of
// candidates 1 and 2 under every model (lane which*M + m; cost with the
// safety penalty and max safety, prmppi_kernels.cu:285-393), the choice
// (cheaper model-averaged candidate becomes the nominal plan, robust action
// when any model finds the chosen plan unsafe), the three temperature
// updates (prmppi_controller.cu:328-346), and the plans for the next step:
// unshifted (plan) and shifted with shift_fill (work).
constexpr int UPD_CH = 8; // steps per chunk of the candidate rollouts' three-wave pipeline

__global__ __launch_bounds__(256) void k_pr_update(const PrInstDev* __restrict__ insts, PrDyn* __restrict__ dyn,
                                                   const double* __restrict__ part, const double* __restrict__ den,
                                                   float* __restrict__ work, float* __restrict__ plan, float* __restrict__ cand,
                                                   int* __restrict__ models, float* __restrict__ actions, const dr_track* __restrict__ tracks,
                                                   const dr_drone_state* __restrict__ drones, QuadConst C, CostConst Q, dr_cost_cfg KC)
{
    __shared__ PrUpdShared sh;
    __shared__ float4 s_pv[DR_MAX_MODELS][MAX_OPP][2];
    __shared__ dr_track s_track;
    __shared__ float4 s_ring[2][UPD_CH][2 * DR_MAX_MODELS]; // position, gate | laps << 8
    __shared__ float4 s_rq[2][UPD_CH][2 * DR_MAX_MODELS];   // attitude at the start of the step
    __shared__ float s_rf[2][UPD_CH][2 * DR_MAX_MODELS];    // clamped collective thrust
    __shared__ float4 s_rc[UPD_CH][2 * DR_MAX_MODELS];      // running cost, safety, terminal code of one chunk

    const int inst = blockIdx.x;
    const PrInstDev& I = insts[inst];
    PrDyn& Y = dyn[inst];
    const int T = I.T, M = I.M, n_opp = I.n_opp, tid = threadIdx.x;
    const uint32_t step = Y.step;

    pr_begin(I, Y, step, den, sh, tid);
    {
        const float* src = reinterpret_cast<const float*>(tracks + I.track);
        float* dst = reinterpret_cast<float*>(&s_track);
        for (int i = tid; i < (int)(sizeof(dr_track) / 4); i += blockDim.x)
            dst[i] = src[i];
    }
    __syncthreads();
    for (int i = tid; i < M * n_opp; i += blockDim.x)
    {
        const int m = i / n_opp, j = i - m * n_opp;
        stage_model_opp(s_pv[m], drones[I.opp[j]], I, sh.br[m], j);
    }
    pr_candidates(I, part, work, sh, tid);
    __syncthreads();

    // Noise-free rollouts, one (candidate, model) per lane l < 2M, in three
    // waves joined by double-buffered LDS rings of UPD_CH steps, one barrier
    // per chunk: lane l of wave 0 clamps the controls and integrates the body
    // rates and the attitude (the next control in a register), handing on each
    // step's pre-step attitude and thrust; lane l of wave 2, one chunk behind,
    // integrates velocity and position and runs the gate counter (the track's
    // gates in registers), writing each step's position, gate and laps; wave
    // 1, two chunks behind, evaluates the running cost, safety and terminal
    // test of every (step, candidate) pair of the chunk in parallel lanes, then
    // lane l sums its candidate's costs in step order.  The operations are
    // quad_step's and pr_stage_cost's.  The state integrates every step; the
    // cost, the safety and the state the terminal cost reads are frozen at the
    // first terminal step.
    {
        const int n_lanes = 2 * M;
        const int n_ch = (T + UPD_CH - 1) / UPD_CH;
        const bool att_lane = tid < n_lanes;
        const int cl = tid - WAVE;
        const bool cost_lane = cl >= 0 && cl < n_lanes;
        const int xl = tid - 2 * WAVE;
        const bool trans_lane = xl >= 0 && xl < n_lanes;
        QState s = state_from_abi(drones[I.drone]);
        GateRegs gr;
        const float4* U = nullptr;
        float4 un = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (att_lane)
        {
            U = reinterpret_cast<const float4*>(sh.c[tid / M]);
            un = U[0];
        }
        if (trans_lane)
            gr.init(s_track, s.gate);
        float acc = 0.0f, decay = 1.0f, smax = -__builtin_inff();
        QState fin = s;
        bool stop = false;
        for (int k = 0; k <= n_ch + 1; ++k)
        {
            if (att_lane && k < n_ch)
            {
                const int t0 = k * UPD_CH, t1 = min(t0 + UPD_CH, T);
                for (int t = t0; t < t1; ++t)
                {
                    const float4 u4 = un;
                    un = U[t + 1 < T ? t + 1 : t]; // issued every step, so never waited for early
                    float uin[4] = {u4.x + 0.0f, u4.y + 0.0f, u4.z + 0.0f, u4.w + 0.0f}, uc[4];
                    clamp_control(C, uin, uc);
                    s_rq[k & 1][t - t0][tid] = make_float4(s.q[0], s.q[1], s.q[2], s.q[3]);
                    s_rf[k & 1][t - t0][tid] = uc[0];
                    quad_attitude_step(C, s.q, s.w, uc);
                }
            }
            else if (trans_lane && k >= 1 && k <= n_ch)
            {
                const int t0 = (k - 1) * UPD_CH, t1 = min(t0 + UPD_CH, T);
                for (int t = t0; t < t1; ++t)
                {
                    const float4 q4 = s_rq[(k - 1) & 1][t - t0][xl];
                    const float q[4] = {q4.x, q4.y, q4.z, q4.w};
                    const float uu[4] = {s_rf[(k - 1) & 1][t - t0][xl], 0.0f, 0.0f, 0.0f}; // the translation reads the thrust only
                    const float p0[3] = {s.p[0], s.p[1], s.p[2]};
                    quad_translation_step(C, q, s.p, s.v, uu);
                    gr.step(KC.gate_traversal_margin, p0, s);
                    s_ring[(k - 1) & 1][t - t0][xl] = make_float4(s.p[0], s.p[1], s.p[2], u2f((uint32_t)s.gate | ((uint32_t)s.laps << 8)));
                }
            }
            else if (cl >= 0 && cl < WAVE && k >= 2)
            {
                const int t0 = (k - 2) * UPD_CH, nt = min(t0 + UPD_CH, T) - t0;
                for (int i = cl; i < nt * n_lanes; i += WAVE)
                {
                    const int to = i / n_lanes, c = i - to * n_lanes;
                    const float4 r = s_ring[(k - 2) & 1][to][c];
                    const uint32_t gl = f2u(r.w);
                    QState x;
                    x.p[0] = r.x;
                    x.p[1] = r.y;
                    x.p[2] = r.z;
                    x.gate = (int)(gl & 0xffu);
                    x.laps = (int)(gl >> 8);
                    const ModelOpp mo{s_pv[c - (c / M) * M], C.dt * (float)(t0 + to + 1)};
                    float sf;
                    int term;
                    const float cst = pr_stage_cost(C, Q, KC, s_track, x, n_opp, mo, &sf, &term);
                    s_rc[to][c] = make_float4(cst, sf, u2f((uint32_t)term), 0.0f);
                }
                if (cost_lane)
                    for (int to = 0; to < nt && !stop; ++to)
                    {
                        const float4 rc = s_rc[to][cl];
                        acc = fma_(rc.x, decay, acc);
                        smax = fmaxf(smax, rc.y);
                        decay = decay * DECAY;
                        const float4 r = s_ring[(k - 2) & 1][to][cl];
                        const uint32_t gl = f2u(r.w);
                        fin.p[0] = r.x;
                        fin.p[1] = r.y;
                        fin.p[2] = r.z;
                        fin.gate = (int)(gl & 0xffu);
                        fin.laps = (int)(gl >> 8);
                        stop = (int)f2u(rc.z) != TERM_NONE;
                    }
            }
            __syncthreads();
        }
        if (cost_lane)
        {
            acc = acc + final_cost(KC, s_track, fin);
            sh.cc[cl] = smax > 0.0f ? acc + I.safety_weight : acc;
            sh.cs[cl] = smax;
        }
    }
    __syncthreads();
    pr_finish(I, Y, step, sh, work, plan, cand, models, actions, inst, tid);
}

// Closed loop: each controlled drone integrates its instance's action;
// drones without a controller keep their velocity.
__global__ __launch_bounds__(64) void k_pr_advance(const PrInstDev* __restrict__ insts, const float* __restrict__ actions,
                                                   dr_drone_state* __restrict__ drones, const int* __restrict__ drone_inst, int n_drones,
                                                   const dr_track* __restrict__ tracks, QuadConst C, float margin)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n_drones)
        return;
    const int i = drone_inst[d];
    if (i >= 0)
        advance_drone(drones[d], actions + i * 4, tracks[insts[i].track], C, margin);
    else
        for (int c = 0; c < 3; ++c)
            drones[d].p[c] = fma_(C.dt, drones[d].v[c], drones[d].p[c]);
}

// ---------------------------------------------------------------------------
// launch wrappers (host)

hipError_t launch_pr_rollout(int waves, int grid, hipStream_t st, const PrInstDev* insts, const PrDyn* dyn, const PrTile* tiles, int n_tiles,
                             const dr_track* tracks, const dr_drone_state* drones, const float* work, float2* pair, float4* noise,
                             const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC)
{
    if (waves < 1 || waves > PR_MAX_WAVES)
        return hipErrorInvalidValue;
    const int g = grid > 0 ? (grid < n_tiles ? grid : n_tiles) : n_tiles;
    k_pr_rollout<<<g, waves * WAVE, 0, st>>>(insts, dyn, tiles, n_tiles, tracks, drones, work, pair, noise, C, Q, KC);
    return hipGetLastError();
}

hipError_t launch_pr_agg(int n_atiles, hipStream_t st, const PrInstDev* insts, const PrDyn* dyn, const PrATile* atiles, const float2* pair,
                         float4* agg, float4* ablk)
{
    k_pr_agg<<<n_atiles, PR_AGG_BLOCK, 0, st>>>(insts, dyn, atiles, pair, agg, ablk);
    return hipGetLastError();
}

hipError_t launch_pr_wavg(int chunk, int n_wtiles, hipStream_t st, const PrInstDev* insts, PrDyn* dyn, const WTile* wtiles,
                          const float4* agg, const float4* noise, const float4* ablk, const float* work, double* part, double* den)
{
    if (chunk < 1 || chunk > PR_WAVG_MAX_CHUNK)
        return hipErrorInvalidValue;
    k_pr_wavg<<<n_wtiles, PR_WAVG_BLOCK, 0, st>>>(insts, dyn, wtiles, agg, noise, ablk, work, part, den, chunk);
    return hipGetLastError();
}

hipError_t launch_pr_update(int n_inst, hipStream_t st, const PrInstDev* insts, PrDyn* dyn, const double* part, const double* den,
                            float* work, float* plan, float* cand, int* models, float* actions, const dr_track* tracks,
                            const dr_drone_state* drones, const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC)
{
    k_pr_update<<<n_inst, 256, 0, st>>>(insts, dyn, part, den, work, plan, cand, models, actions, tracks, drones, C, Q, KC);
    return hipGetLastError();
}

hipError_t launch_pr_advance(int n_drones, hipStream_t st, const PrInstDev* insts, const float* actions, dr_drone_state* drones,
                             const int* drone_inst, const dr_track* tracks, const QuadConst& C, float margin)
{
    k_pr_advance<<<(n_drones + 63) / 64, 64, 0, st>>>(insts, actions, drones, drone_inst, n_drones, tracks, C, margin);
    return hipGetLastError();
}

} // namespace dr
