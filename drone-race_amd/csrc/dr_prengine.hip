// Host side of the PR-MPPI engine: descriptor tables, device buffers, the
// per-step launch sequence, hipGraph capture and the dr_pr_* C ABI of
// include/dr_api.h.
//
// Reference host orchestration this replaces:
//   PRMPPIController::allocDevice / getControl  src/prmppi_controller.cu:47-104, 142-349
// The reference runs 8 launches and 8 blocking copies per controller and
// step, with the candidate choice and the safety check on the host; here one
// graph replay runs the whole step for every instance of the batch and the
// choice is made on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "dr_host.h"
#include "dr_prkernels.h"

namespace dr {

hipError_t launch_pr_rollout(int waves, int grid, hipStream_t st, const PrInstDev* insts, const PrDyn* dyn, const PrTile* tiles, int n_tiles,
                             const dr_track* tracks, const dr_drone_state* drones, const float* work, float2* pair, float4* noise,
                             const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC);
hipError_t launch_pr_agg(int n_atiles, hipStream_t st, const PrInstDev* insts, const PrDyn* dyn, const PrATile* atiles, const float2* pair,
                         float4* agg, float4* ablk);
hipError_t launch_pr_wavg(int chunk, int n_wtiles, hipStream_t st, const PrInstDev* insts, PrDyn* dyn, const WTile* wtiles,
                          const float4* agg, const float4* noise, const float4* ablk, const float* work, double* part, double* den);
hipError_t launch_pr_update(int n_inst, hipStream_t st, const PrInstDev* insts, PrDyn* dyn, const double* part, const double* den,
                            float* work, float* plan, float* cand, int* models, float* actions, const dr_track* tracks,
                            const dr_drone_state* drones, const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC);
hipError_t launch_pr_advance(int n_drones, hipStream_t st, const PrInstDev* insts, const float* actions, dr_drone_state* drones,
                             const int* drone_inst, const dr_track* tracks, const QuadConst& C, float margin);
hipError_t launch_prsr_rollout(int waves, int grid, hipStream_t st, const PrInstDev* insts, const PrDyn* dyn, const PrTile* tiles, int n_tiles,
                               const dr_sr_env* env, const dr_sr_cost& KC, const dr_sr_state* states, const float* work, float2* pair,
                               float4* noise);
hipError_t launch_prsr_update(int n_inst, hipStream_t st, const PrInstDev* insts, PrDyn* dyn, const double* part, const double* den,
                              float* work, float* plan, float* cand, int* models, float* actions, const dr_sr_env* env, const dr_sr_cost& KC,
                              const dr_sr_state* states);
hipError_t launch_prsr_advance(int n_inst, hipStream_t st, const PrInstDev* insts, PrDyn* dyn, const dr_sr_env* env, dr_sr_state* states,
                               const float* actions, int* events);

} // namespace dr

using namespace dr;

namespace {

int n_models_of(const dr_pr_instance& s) { return s.sampled ? s.n_models : s.n_branches; }

// Plan shifted one step left, shift_fill in the freed last step
// (prmppi_controller.cu:163-175 fills zero).
void shift_plan(const float* src, float* dst, int T, const float* fill)
{
    for (int t = 0; t < T; ++t)
        for (int d = 0; d < 4; ++d)
            dst[t * 4 + d] = (t + 1 < T) ? src[(t + 1) * 4 + d] : fill[d];
}

// Plans of D control dimensions <-> the engine's 4-wide layout (D = 2 for
// StratRace2D: lanes z, w stay zero).
void widen(const float* src, float* dst, int T, int D)
{
    for (int t = 0; t < T; ++t)
        for (int d = 0; d < 4; ++d)
            dst[t * 4 + d] = d < D ? src[t * D + d] : 0.0f;
}

void narrow(const float* src, float* dst, int T, int D)
{
    for (int t = 0; t < T; ++t)
        for (int d = 0; d < D; ++d)
            dst[t * D + d] = src[t * 4 + d];
}

} // namespace

struct dr_prengine
{
    int device = 0;
    hipStream_t stream = nullptr;
    QuadConst C{};
    CostConst Q{};
    dr_cost_cfg cost{};
    std::vector<dr_track> tracks;
    std::vector<dr_pr_instance> inst;
    int env_kind = 0; // 0 quadrotor race, 1 StratRace2D
    dr_sr_env srenv{};
    dr_sr_cost srcost{};
    std::vector<dr_prsr_instance> srinst;
    std::vector<int> drone_world;
    std::vector<int> drone_inst;
    int n_drones = 0;

    dr_launch_cfg launch{};
    std::vector<PrInstDev> idev;
    std::vector<PrTile> rtiles;
    std::vector<PrATile> atiles;
    std::vector<WTile> wtiles;
    long long noise_total = 0, pair_total = 0, part_total = 0;
    int agg_total = 0, plan_total = 0, model_total = 0, den_total = 0;

    DevBuf<PrInstDev> d_inst;
    DevBuf<PrDyn> d_dyn;
    DevBuf<PrTile> d_rtiles;
    DevBuf<PrATile> d_atiles;
    DevBuf<WTile> d_wtiles;
    DevBuf<dr_track> d_tracks;
    DevBuf<dr_drone_state> d_drones;
    DevBuf<int> d_drone_inst;
    DevBuf<float> d_work, d_plan, d_cand;
    DevBuf<float4> d_noise;
    DevBuf<float2> d_pair;
    DevBuf<float4> d_agg, d_ablk;
    DevBuf<double> d_part, d_den;
    DevBuf<int> d_models;
    DevBuf<float> d_actions;
    DevBuf<dr_sr_env> d_srenv;
    DevBuf<dr_sr_state> d_srstates;
    DevBuf<int> d_events;

    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    int graph_steps[2] = {0, 0};

    bool timing = false;
    std::vector<hipEvent_t> events;
    int timed_steps = 0;
    float last_ms[5] = {0, 0, 0, 0, 0};

    ~dr_prengine()
    {
        drop_graphs();
        for (auto ev : events)
            (void)hipEventDestroy(ev);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    void drop_graphs()
    {
        for (int i = 0; i < 2; ++i)
        {
            if (graph_exec[i])
                (void)hipGraphExecDestroy(graph_exec[i]);
            graph_exec[i] = nullptr;
            graph_steps[i] = 0;
        }
    }

    // Rollout waves per workgroup: the configured block, or every rollout of
    // a sample group of the widest instance (at most 16 waves).
    int waves() const
    {
        if (launch.rollout_block > 0)
            return launch.rollout_block / 64;
        int w = 1;
        for (const PrInstDev& I : idev)
            w = std::max(w, 2 * I.M);
        return std::min(w, PR_MAX_WAVES);
    }
    int dim() const { return env_kind ? 2 : 4; } // control dimensions of the plans the caller sees
    int build_layout();
    int enqueue_step(bool advance, int timed_index);
    int run(int n_steps, bool advance);
};

int dr_prengine::build_layout()
{
    const int n = (int)inst.size();
    const int chunk = launch.reduce_chunk;
    idev.assign(n, PrInstDev{});
    noise_total = pair_total = part_total = 0;
    agg_total = plan_total = model_total = den_total = 0;
    for (int i = 0; i < n; ++i)
    {
        PrInstDev& I = idev[i];
        if (env_kind == 1)
        {
            const dr_prsr_instance& r = srinst[i];
            I.K = r.n_samples;
            I.T = r.n_timesteps;
            I.track = 0;
            I.drone = i;
            I.key0 = (uint32_t)(r.seed & 0xffffffffu);
            I.key1 = (uint32_t)(r.seed >> 32);
            I.sigma[0] = I.sigma[1] = r.sigma;
            I.sigma[2] = I.sigma[3] = 0.0f;
            I.shift_fill[0] = r.shift_fill[0];
            I.shift_fill[1] = r.shift_fill[1];
            I.shift_fill[2] = I.shift_fill[3] = 0.0f;
            I.adapt = r.adapt_lambda;
            I.n_opp = 0;
            I.B = srenv.n_models;
            I.sampled = r.sampled ? 1 : 0;
            I.M = r.sampled ? r.n_draws : srenv.n_models;
            I.safety_weight = r.safety_weight;
            I.world_theta = r.true_theta;
        }
        else
        {
            const dr_pr_instance& s = inst[i];
            I.K = s.n_samples;
            I.T = s.n_timesteps;
            I.track = s.track;
            I.drone = s.drone;
            I.key0 = (uint32_t)(s.seed & 0xffffffffu);
            I.key1 = (uint32_t)(s.seed >> 32);
            for (int d = 0; d < 4; ++d)
            {
                I.sigma[d] = s.sigma[d];
                I.shift_fill[d] = s.shift_fill[d];
            }
            I.adapt = s.adapt_lambda;
            I.n_opp = 0;
            for (int d = 0; d < n_drones; ++d)
                if (d != s.drone && drone_world[d] == drone_world[s.drone])
                    I.opp[I.n_opp++] = d;
            I.M = n_models_of(s);
            I.B = s.n_branches;
            I.sampled = s.sampled ? 1 : 0;
            I.safety_weight = s.safety_weight;
            for (int b = 0; b < DR_MAX_BRANCHES; ++b)
                for (int j = 0; j < MAX_OPP; ++j)
                    for (int c = 0; c < 3; ++c)
                        I.shift[b][j][c] = s.vel_shift[b][j][c];
        }
        I.noise_off = noise_total;
        noise_total += (long long)I.K * I.T;
        I.pair_off = pair_total;
        pair_total += 2LL * I.M * I.K;
        I.agg_off = agg_total;
        agg_total += I.K;
        I.plan_off = plan_total;
        plan_total += 2 * I.T * 4;
        I.model_off = model_total;
        model_total += I.M;
        I.n_chunks = (I.K + chunk - 1) / chunk;
        I.part_off = part_total;
        part_total += (long long)I.n_chunks * 3 * I.T * 4;
        I.den_off = den_total;
        den_total += I.n_chunks * 3;
    }

    // Rollout tiles, longest work (T * rollouts per sample) first.
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return (long long)idev[a].T * idev[a].M > (long long)idev[b].T * idev[b].M; });
    const int W = waves();
    rtiles.clear();
    atiles.clear();
    wtiles.clear();
    for (int oi = 0; oi < n; ++oi)
    {
        const int i = order[oi];
        PrInstDev& I = idev[i];
        I.rtile_first = (int)rtiles.size();
        for (int k0 = 0; k0 < I.K; k0 += 64)
            for (int p0 = 0; p0 < 2 * I.M; p0 += W)
                rtiles.push_back(PrTile{i, k0, p0, std::min(W, 2 * I.M - p0)});
        I.rtile_count = (int)rtiles.size() - I.rtile_first;
    }
    for (int i = 0; i < n; ++i)
    {
        PrInstDev& I = idev[i];
        I.ablk_first = (int)atiles.size();
        for (int k0 = 0; k0 < I.K; k0 += PR_AGG_BLOCK)
            atiles.push_back(PrATile{i, k0});
        I.ablk_count = (int)atiles.size() - I.ablk_first;
    }
    for (int oi = 0; oi < n; ++oi)
    {
        const int i = order[oi];
        const PrInstDev& I = idev[i];
        const int tg = launch.reduce_tgroup;
        for (int c = 0; c < I.n_chunks; ++c)
            for (int t0 = 0; t0 < I.T; t0 += tg)
                wtiles.push_back(WTile{i, c, t0, std::min(I.T, t0 + tg)});
    }

    HIPCHK(d_inst.alloc(n));
    HIPCHK(hipMemcpy(d_inst.p, idev.data(), n * sizeof(PrInstDev), hipMemcpyHostToDevice));
    HIPCHK(d_rtiles.alloc(rtiles.size()));
    HIPCHK(hipMemcpy(d_rtiles.p, rtiles.data(), rtiles.size() * sizeof(PrTile), hipMemcpyHostToDevice));
    HIPCHK(d_atiles.alloc(atiles.size()));
    HIPCHK(hipMemcpy(d_atiles.p, atiles.data(), atiles.size() * sizeof(PrATile), hipMemcpyHostToDevice));
    HIPCHK(d_wtiles.alloc(wtiles.size()));
    HIPCHK(hipMemcpy(d_wtiles.p, wtiles.data(), wtiles.size() * sizeof(WTile), hipMemcpyHostToDevice));
    HIPCHK(d_ablk.alloc(atiles.size()));
    HIPCHK(d_part.alloc(part_total));
    HIPCHK(d_den.alloc(den_total));
    drop_graphs();
    return DR_OK;
}

int dr_prengine::enqueue_step(bool advance, int timed_index)
{
    const int n = (int)inst.size();
    hipEvent_t* ev = (timed_index >= 0) ? &events[(size_t)timed_index * 5] : nullptr;
    if (ev)
        HIPCHK(hipEventRecord(ev[0], stream));
    if (env_kind == 1)
        HIPCHK(launch_prsr_rollout(waves(), launch.rollout_grid, stream, d_inst.p, d_dyn.p, d_rtiles.p, (int)rtiles.size(), d_srenv.p, srcost,
                                   d_srstates.p, d_work.p, d_pair.p, d_noise.p));
    else
        HIPCHK(launch_pr_rollout(waves(), launch.rollout_grid, stream, d_inst.p, d_dyn.p, d_rtiles.p, (int)rtiles.size(), d_tracks.p,
                                 d_drones.p, d_work.p, d_pair.p, d_noise.p, C, Q, cost));
    if (ev)
        HIPCHK(hipEventRecord(ev[1], stream));
    HIPCHK(launch_pr_agg((int)atiles.size(), stream, d_inst.p, d_dyn.p, d_atiles.p, d_pair.p, d_agg.p, d_ablk.p));
    if (ev)
        HIPCHK(hipEventRecord(ev[2], stream));
    HIPCHK(launch_pr_wavg(launch.reduce_chunk, (int)wtiles.size(), stream, d_inst.p, d_dyn.p, d_wtiles.p, d_agg.p, d_noise.p, d_ablk.p,
                          d_work.p, d_part.p, d_den.p));
    if (ev)
        HIPCHK(hipEventRecord(ev[3], stream));
    if (env_kind == 1)
    {
        HIPCHK(launch_prsr_update(n, stream, d_inst.p, d_dyn.p, d_part.p, d_den.p, d_work.p, d_plan.p, d_cand.p, d_models.p, d_actions.p,
                                  d_srenv.p, srcost, d_srstates.p));
        if (advance)
            HIPCHK(launch_prsr_advance(n, stream, d_inst.p, d_dyn.p, d_srenv.p, d_srstates.p, d_actions.p, d_events.p));
    }
    else
    {
        HIPCHK(launch_pr_update(n, stream, d_inst.p, d_dyn.p, d_part.p, d_den.p, d_work.p, d_plan.p, d_cand.p, d_models.p, d_actions.p,
                                d_tracks.p, d_drones.p, C, Q, cost));
        if (advance)
            HIPCHK(launch_pr_advance(n_drones, stream, d_inst.p, d_actions.p, d_drones.p, d_drone_inst.p, d_tracks.p, C,
                                     cost.gate_traversal_margin));
    }
    if (ev)
        HIPCHK(hipEventRecord(ev[4], stream));
    return DR_OK;
}

int dr_prengine::run(int n_steps, bool advance)
{
    if (n_steps < 0)
        return host_fail(DR_EINVAL, "n_steps < 0");
    timed_steps = 0;
    if (timing)
    {
        const size_t need = (size_t)n_steps * 5;
        while (events.size() < need)
        {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            events.push_back(ev);
        }
        for (int s = 0; s < n_steps; ++s)
            if (int rc = enqueue_step(advance, s))
                return rc;
        timed_steps = n_steps;
        return DR_OK;
    }
    if (!launch.use_graph)
    {
        for (int s = 0; s < n_steps; ++s)
            if (int rc = enqueue_step(advance, -1))
                return rc;
        return DR_OK;
    }
    const int gi = advance ? 1 : 0;
    const int spg = std::max(1, launch.steps_per_graph);
    if (!graph_exec[gi] || graph_steps[gi] != spg)
    {
        if (graph_exec[gi])
            (void)hipGraphExecDestroy(graph_exec[gi]);
        graph_exec[gi] = nullptr;
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = DR_OK;
        for (int s = 0; s < spg && rc == DR_OK; ++s)
            rc = enqueue_step(advance, -1);
        hipError_t ce = hipStreamEndCapture(stream, &g);
        if (rc)
        {
            if (g)
                (void)hipGraphDestroy(g);
            return rc;
        }
        if (ce != hipSuccess)
            return host_fail(DR_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        hipError_t ie = hipGraphInstantiate(&graph_exec[gi], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess)
            return host_fail(DR_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        graph_steps[gi] = spg;
    }
    int done = 0;
    while (done + spg <= n_steps)
    {
        HIPCHK(hipGraphLaunch(graph_exec[gi], stream));
        done += spg;
    }
    for (; done < n_steps; ++done)
        if (int rc = enqueue_step(advance, -1))
            return rc;
    return DR_OK;
}

// ---------------------------------------------------------------------------
// C ABI

namespace {

int check_pr_inst(dr_prengine* e, int i)
{
    if (!e)
        return host_fail(DR_EINVAL, "NULL handle");
    if (i < 0 || i >= (int)e->inst.size())
        return host_fail(DR_EINVAL, "instance index out of range");
    return DR_OK;
}

bool valid_belief(const float* b, int n)
{
    for (int i = 0; i < n; ++i)
        if (!(b[i] >= 0.0f) || !std::isfinite(b[i]))
            return false;
    return true;
}

int valid_launch(const dr_launch_cfg& c)
{
    if (c.rollout_block != 0 && (c.rollout_block < 64 || c.rollout_block > 64 * PR_MAX_WAVES || c.rollout_block % 64 != 0))
        return host_fail(DR_EINVAL, "rollout_block must be 0 (auto) or a multiple of 64 in [64, 1024]");
    if (c.rollout_grid < 0)
        return host_fail(DR_EINVAL, "rollout_grid must be >= 0");
    if (c.reduce_chunk < 64 || c.reduce_chunk > PR_WAVG_MAX_CHUNK || c.reduce_chunk % 64 != 0)
        return host_fail(DR_EINVAL, "reduce_chunk must be a multiple of 64 in [64, 2048]");
    if (c.reduce_tgroup < 1 || c.reduce_tgroup > PR_WAVG_BLOCK)
        return host_fail(DR_EINVAL, "reduce_tgroup must be in [1, 256]");
    if (c.steps_per_graph < 1)
        return host_fail(DR_EINVAL, "steps_per_graph must be >= 1");
    return DR_OK;
}

template <class T>
int copy_out(dr_prengine* e, T* dst, const T* src, size_t n)
{
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

template <class T>
int copy_in(dr_prengine* e, T* dst, const T* src, size_t n)
{
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

} // namespace

extern "C" {

int dr_pr_create(const dr_quad_params* params, const dr_cost_cfg* cost, const dr_track* tracks, int n_tracks, const dr_pr_instance* instances,
                 int n_instances, const int32_t* drone_world, int n_drones, dr_prengine** out)
{
    if (!out)
        return host_fail(DR_EINVAL, "out is NULL");
    *out = nullptr;
    if (!params || !cost || !tracks || !instances || !drone_world)
        return host_fail(DR_EINVAL, "NULL argument");
    if (n_tracks < 1 || n_instances < 1 || n_drones < 1)
        return host_fail(DR_EINVAL, "need at least one track, instance and drone");
    if (!(params->mass > 0.0f) || !(params->dt > 0.0f) || !(params->drone_radius > 0.0f))
        return host_fail(DR_EINVAL, "mass, dt and drone_radius must be > 0");
    if (!(params->thrust_max >= params->thrust_min) || !(params->rate_max >= 0.0f))
        return host_fail(DR_EINVAL, "control limits are inconsistent");
    if (!(cost->opp_prox_radius > 0.0f) || !(cost->boundary_threshold_factor > 0.0f))
        return host_fail(DR_EINVAL, "opp_prox_radius and boundary_threshold_factor must be > 0");
    for (int i = 0; i < n_tracks; ++i)
        if (int rc = check_track(tracks[i], i))
            return rc;
    std::vector<int> inst_of_drone(n_drones, -1);
    for (int i = 0; i < n_instances; ++i)
    {
        const dr_pr_instance& s = instances[i];
        const std::string w = "instance " + std::to_string(i) + ": ";
        if (s.n_samples < 1 || s.n_samples > (1 << 24))
            return host_fail(DR_EINVAL, w + "n_samples must be in [1, 2^24]");
        if (s.n_timesteps < 1 || s.n_timesteps > DR_MAX_T)
            return host_fail(DR_EINVAL, w + "n_timesteps must be in [1, 256]");
        if (s.track < 0 || s.track >= n_tracks)
            return host_fail(DR_EINVAL, w + "track index out of range");
        if (s.drone < 0 || s.drone >= n_drones)
            return host_fail(DR_EINVAL, w + "drone index out of range");
        if (inst_of_drone[s.drone] >= 0)
            return host_fail(DR_EINVAL, w + "drone already controlled by another instance");
        if (!(s.lambda > 0.0f))
            return host_fail(DR_EINVAL, w + "lambda must be > 0");
        if (s.n_branches < 1 || s.n_branches > DR_MAX_BRANCHES)
            return host_fail(DR_EINVAL, w + "n_branches must be in [1, 8]");
        if (s.sampled != 0 && s.sampled != 1)
            return host_fail(DR_EINVAL, w + "sampled must be 0 or 1");
        if (s.sampled && (s.n_models < 1 || s.n_models > DR_MAX_MODELS))
            return host_fail(DR_EINVAL, w + "n_models must be in [1, 16]");
        if (!valid_belief(s.belief, s.n_branches))
            return host_fail(DR_EINVAL, w + "belief entries must be finite and >= 0");
        if (!std::isfinite(s.safety_weight))
            return host_fail(DR_EINVAL, w + "safety_weight must be finite");
        inst_of_drone[s.drone] = i;
    }
    for (int d = 0; d < n_drones; ++d)
    {
        int cnt = 0;
        for (int x = 0; x < n_drones; ++x)
            cnt += drone_world[x] == drone_world[d];
        if (cnt > DR_MAX_WORLD_DRONES)
            return host_fail(DR_EINVAL, "a world holds more than 8 drones");
    }

    dr_prengine* e = new (std::nothrow) dr_prengine();
    if (!e)
        return host_fail(DR_ENOMEM, "host allocation failed");
    auto bail = [&](int rc) {
        delete e;
        return rc;
    };
    if (hipGetDevice(&e->device) != hipSuccess)
        return bail(host_fail(DR_EHIP, "no HIP device"));
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(host_fail(DR_EHIP, "stream creation failed"));
    e->C = make_quad_const(*params);
    e->Q = make_cost_const(*params, *cost);
    e->cost = *cost;
    e->tracks.assign(tracks, tracks + n_tracks);
    e->inst.assign(instances, instances + n_instances);
    e->drone_world.assign(drone_world, drone_world + n_drones);
    e->drone_inst = inst_of_drone;
    e->n_drones = n_drones;
    e->launch = dr_launch_cfg{0, 0, 1024, 1, 1, 1, 16, 0};

    if (int rc = e->build_layout())
        return bail(rc);
    auto chk = [&](hipError_t x, const char* what) {
        if (x != hipSuccess)
        {
            host_fail(x == hipErrorOutOfMemory ? DR_ENOMEM : DR_EHIP, std::string(what) + ": " + hipGetErrorString(x));
            return false;
        }
        return true;
    };
    if (!chk(e->d_dyn.alloc(n_instances), "alloc dyn") || !chk(e->d_tracks.alloc(n_tracks), "alloc tracks") ||
        !chk(e->d_drones.alloc(n_drones), "alloc drones") || !chk(e->d_drone_inst.alloc(n_drones), "alloc drone_inst") ||
        !chk(e->d_work.alloc(e->plan_total), "alloc work") || !chk(e->d_plan.alloc(e->plan_total), "alloc plan") ||
        !chk(e->d_cand.alloc(e->plan_total), "alloc cand") || !chk(e->d_noise.alloc(e->noise_total), "alloc noise") ||
        !chk(e->d_pair.alloc(e->pair_total), "alloc pair") || !chk(e->d_agg.alloc(e->agg_total), "alloc agg") ||
        !chk(e->d_models.alloc(e->model_total), "alloc models") || !chk(e->d_actions.alloc((size_t)n_instances * 4), "alloc actions"))
        return bail(DR_ENOMEM);
    std::vector<PrDyn> dyn(n_instances);
    for (int i = 0; i < n_instances; ++i)
    {
        std::memset(&dyn[i], 0, sizeof(PrDyn));
        for (int j = 0; j < 3; ++j)
            dyn[i].lambda[j] = instances[i].lambda;
        for (int b = 0; b < instances[i].n_branches; ++b)
            dyn[i].belief[b] = instances[i].belief[b];
        dyn[i].use_nom = 1;
    }
    // Both plans start at zero (prmppi_controller.cu:33-37); the plans used
    // by the first step are those shifted.
    std::vector<float> plan(e->plan_total, 0.0f), work(e->plan_total, 0.0f);
    for (int i = 0; i < n_instances; ++i)
    {
        const PrInstDev& I = e->idev[i];
        for (int p = 0; p < 2; ++p)
            shift_plan(&plan[I.plan_off + p * I.T * 4], &work[I.plan_off + p * I.T * 4], I.T, I.shift_fill);
    }
    std::vector<dr_drone_state> dr(n_drones);
    for (auto& s : dr)
    {
        std::memset(&s, 0, sizeof(s));
        s.q[0] = 1.0f;
    }
    if (!chk(hipMemcpy(e->d_dyn.p, dyn.data(), dyn.size() * sizeof(PrDyn), hipMemcpyHostToDevice), "copy dyn") ||
        !chk(hipMemcpy(e->d_tracks.p, tracks, n_tracks * sizeof(dr_track), hipMemcpyHostToDevice), "copy tracks") ||
        !chk(hipMemcpy(e->d_drones.p, dr.data(), n_drones * sizeof(dr_drone_state), hipMemcpyHostToDevice), "copy drones") ||
        !chk(hipMemcpy(e->d_drone_inst.p, inst_of_drone.data(), n_drones * sizeof(int), hipMemcpyHostToDevice), "copy drone_inst") ||
        !chk(hipMemcpy(e->d_plan.p, plan.data(), plan.size() * sizeof(float), hipMemcpyHostToDevice), "copy plan") ||
        !chk(hipMemcpy(e->d_work.p, work.data(), work.size() * sizeof(float), hipMemcpyHostToDevice), "copy work") ||
        !chk(hipMemcpy(e->d_cand.p, plan.data(), plan.size() * sizeof(float), hipMemcpyHostToDevice), "copy cand"))
        return bail(DR_EHIP);
    *out = e;
    return DR_OK;
}

void dr_pr_destroy(dr_prengine* e)
{
    if (e)
    {
        (void)hipStreamSynchronize(e->stream);
        delete e;
    }
}

int dr_pr_set_drones(dr_prengine* e, const dr_drone_state* states, int n_drones)
{
    if (e && e->env_kind == 1)
        return host_fail(DR_ESTATE, "set_drones: a StratRace2D engine takes dr_pr_set_sr_states");
    if (!e || !states || n_drones != e->n_drones)
        return host_fail(DR_EINVAL, "set_drones: bad handle, pointer or count");
    if (int rc = check_drones(states, n_drones))
        return rc;
    return copy_in(e, e->d_drones.p, states, (size_t)n_drones);
}

int dr_pr_get_drones(dr_prengine* e, dr_drone_state* states, int n_drones)
{
    if (e && e->env_kind == 1)
        return host_fail(DR_ESTATE, "get_drones: a StratRace2D engine takes dr_pr_get_sr_states");
    if (!e || !states || n_drones != e->n_drones)
        return host_fail(DR_EINVAL, "get_drones: bad handle, pointer or count");
    return copy_out(e, states, e->d_drones.p, (size_t)n_drones);
}

int dr_pr_set_nominals(dr_prengine* e, int instance, const float* nom, const float* rob, int n_floats)
{
    if (int rc = check_pr_inst(e, instance))
        return rc;
    const PrInstDev& I = e->idev[instance];
    const int D = e->dim();
    if (!nom || !rob || n_floats != I.T * D)
        return host_fail(DR_EINVAL, "set_nominals: expected two plans of T*" + std::to_string(D) + " floats");
    std::vector<float> plan(2 * I.T * 4), work(2 * I.T * 4);
    widen(nom, plan.data(), I.T, D);
    widen(rob, plan.data() + I.T * 4, I.T, D);
    for (int p = 0; p < 2; ++p)
        shift_plan(&plan[p * I.T * 4], &work[p * I.T * 4], I.T, I.shift_fill);
    if (int rc = copy_in(e, e->d_plan.p + I.plan_off, plan.data(), plan.size()))
        return rc;
    return copy_in(e, e->d_work.p + I.plan_off, work.data(), work.size());
}

int dr_pr_get_nominals(dr_prengine* e, int instance, float* nom, float* rob, int n_floats)
{
    if (int rc = check_pr_inst(e, instance))
        return rc;
    const PrInstDev& I = e->idev[instance];
    const int D = e->dim();
    if (!nom || !rob || n_floats != I.T * D)
        return host_fail(DR_EINVAL, "get_nominals: expected two plans of T*" + std::to_string(D) + " floats");
    std::vector<float> plan(2 * I.T * 4);
    if (int rc = copy_out(e, plan.data(), e->d_plan.p + I.plan_off, plan.size()))
        return rc;
    narrow(plan.data(), nom, I.T, D);
    narrow(plan.data() + I.T * 4, rob, I.T, D);
    return DR_OK;
}

int dr_pr_set_belief(dr_prengine* e, int instance, const float* belief, int n_branches)
{
    if (int rc = check_pr_inst(e, instance))
        return rc;
    if (!belief || n_branches != e->idev[instance].B)
        return host_fail(DR_EINVAL, "set_belief: expected n_branches values");
    if (!valid_belief(belief, n_branches))
        return host_fail(DR_EINVAL, "set_belief: entries must be finite and >= 0");
    float b[DR_MAX_BRANCHES] = {0};
    std::memcpy(b, belief, n_branches * sizeof(float));
    return copy_in(e, e->d_dyn.p[instance].belief, b, (size_t)DR_MAX_BRANCHES);
}

int dr_pr_get_belief(dr_prengine* e, int instance, float* belief, int n_branches)
{
    if (int rc = check_pr_inst(e, instance))
        return rc;
    if (!belief || n_branches != e->idev[instance].B)
        return host_fail(DR_EINVAL, "get_belief: expected n_branches values");
    float b[DR_MAX_BRANCHES];
    if (int rc = copy_out(e, b, e->d_dyn.p[instance].belief, (size_t)DR_MAX_BRANCHES))
        return rc;
    std::memcpy(belief, b, n_branches * sizeof(float));
    return DR_OK;
}

int dr_pr_step_async(dr_prengine* e, int n_steps, int advance)
{
    if (!e)
        return host_fail(DR_EINVAL, "NULL handle");
    return e->run(n_steps, advance != 0);
}

int dr_pr_synchronize(dr_prengine* e)
{
    if (!e)
        return host_fail(DR_EINVAL, "NULL handle");
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->timed_steps > 0)
    {
        float acc[5] = {0, 0, 0, 0, 0};
        for (int s = 0; s < e->timed_steps; ++s)
        {
            hipEvent_t* ev = &e->events[(size_t)s * 5];
            for (int q = 0; q < 4; ++q)
            {
                float v;
                HIPCHK(hipEventElapsedTime(&v, ev[q], ev[q + 1]));
                acc[q] += v;
            }
            float t;
            HIPCHK(hipEventElapsedTime(&t, ev[0], ev[4]));
            acc[4] += t;
        }
        std::memcpy(e->last_ms, acc, sizeof(acc));
        e->timed_steps = 0;
    }
    return DR_OK;
}

int dr_pr_step(dr_prengine* e, int n_steps, int advance, float* actions)
{
    if (int rc = dr_pr_step_async(e, n_steps, advance))
        return rc;
    if (int rc = dr_pr_synchronize(e))
        return rc;
    if (actions)
    {
        const int n = (int)e->inst.size();
        std::vector<float> a((size_t)n * 4);
        if (int rc = copy_out(e, a.data(), e->d_actions.p, a.size()))
            return rc;
        narrow(a.data(), actions, n, e->dim());
    }
    return DR_OK;
}

int dr_pr_get_diag(dr_prengine* e, dr_pr_diag* diag, int n_instances)
{
    if (!e || !diag || n_instances != (int)e->inst.size())
        return host_fail(DR_EINVAL, "get_diag: bad handle, pointer or count");
    std::vector<PrDyn> dyn(n_instances);
    if (int rc = copy_out(e, dyn.data(), e->d_dyn.p, dyn.size()))
        return rc;
    for (int i = 0; i < n_instances; ++i)
    {
        dr_pr_diag& d = diag[i];
        std::memset(&d, 0, sizeof(d));
        for (int j = 0; j < 3; ++j)
        {
            d.min_costs[j] = dyn[i].min_cost[j];
            d.nu[j] = dyn[i].nu[j];
            d.lambda[j] = dyn[i].lambda[j];
        }
        d.cand_costs[0] = dyn[i].cand_cost[0];
        d.cand_costs[1] = dyn[i].cand_cost[1];
        d.use_nom_plan = dyn[i].use_nom;
        d.reset_nom = dyn[i].reset_nom;
        d.step = dyn[i].step;
    }
    return DR_OK;
}

int dr_pr_get_noise(dr_prengine* e, int instance, float* noise, size_t n)
{
    if (int rc = check_pr_inst(e, instance))
        return rc;
    const PrInstDev& I = e->idev[instance];
    const int D = e->dim();
    if (!noise || n != (size_t)I.K * I.T * D)
        return host_fail(DR_EINVAL, "get_noise: expected K*T*" + std::to_string(D) + " floats");
    if (D == 4)
        return copy_out(e, reinterpret_cast<float4*>(noise), e->d_noise.p + I.noise_off, (size_t)I.K * I.T);
    std::vector<float> rows((size_t)I.K * I.T * 4);
    if (int rc = copy_out(e, reinterpret_cast<float4*>(rows.data()), e->d_noise.p + I.noise_off, (size_t)I.K * I.T))
        return rc;
    narrow(rows.data(), noise, I.K * I.T, D);
    return DR_OK;
}

int dr_pr_get_pair_costs(dr_prengine* e, int instance, float* out, int n_floats)
{
    if (int rc = check_pr_inst(e, instance))
        return rc;
    const PrInstDev& I = e->idev[instance];
    if (!out || (long long)n_floats != 2LL * 2 * I.M * I.K)
        return host_fail(DR_EINVAL, "get_pair_costs: expected 2*M*K*2 floats");
    return copy_out(e, reinterpret_cast<float2*>(out), e->d_pair.p + I.pair_off, (size_t)2 * I.M * I.K);
}

int dr_pr_get_aggregates(dr_prengine* e, int instance, float* agg, int n_floats)
{
    if (int rc = check_pr_inst(e, instance))
        return rc;
    const PrInstDev& I = e->idev[instance];
    if (!agg || n_floats != I.K * 4)
        return host_fail(DR_EINVAL, "get_aggregates: expected K*4 floats");
    return copy_out(e, reinterpret_cast<float4*>(agg), e->d_agg.p + I.agg_off, (size_t)I.K);
}

int dr_pr_get_candidates(dr_prengine* e, int instance, float* cand1, float* cand2, int n_floats)
{
    if (int rc = check_pr_inst(e, instance))
        return rc;
    const PrInstDev& I = e->idev[instance];
    const int D = e->dim();
    if (!cand1 || !cand2 || n_floats != I.T * D)
        return host_fail(DR_EINVAL, "get_candidates: expected two plans of T*" + std::to_string(D) + " floats");
    std::vector<float> c(2 * I.T * 4);
    if (int rc = copy_out(e, c.data(), e->d_cand.p + I.plan_off, c.size()))
        return rc;
    narrow(c.data(), cand1, I.T, D);
    narrow(c.data() + I.T * 4, cand2, I.T, D);
    return DR_OK;
}

int dr_pr_get_models(dr_prengine* e, int instance, int32_t* branch_of_model, int n_models)
{
    if (int rc = check_pr_inst(e, instance))
        return rc;
    const PrInstDev& I = e->idev[instance];
    if (!branch_of_model || n_models != I.M)
        return host_fail(DR_EINVAL, "get_models: expected M values");
    return copy_out(e, branch_of_model, e->d_models.p + I.model_off, (size_t)n_models);
}

int dr_pr_set_launch(dr_prengine* e, const dr_launch_cfg* cfg)
{
    if (!e || !cfg)
        return host_fail(DR_EINVAL, "NULL argument");
    if (int rc = valid_launch(*cfg))
        return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    e->launch = *cfg;
    return e->build_layout();
}

// Times the candidate launch shapes on the engine's own workload (trials
// steps each after one warm-up step, no closed-loop advance) and keeps the
// fastest; the plans and per-instance state the timing steps change are
// restored.  Candidates: rollout waves per workgroup (every rollout of a
// sample group, or 2, 4, 8) x weighted-average chunk x timestep group.
int dr_pr_autotune(dr_prengine* e, int trials, dr_launch_cfg* best, float* best_ms)
{
    if (!e || trials < 1)
        return host_fail(DR_EINVAL, "autotune: bad handle or trials");
    HIPCHK(hipStreamSynchronize(e->stream));
    const size_t n = e->inst.size();  // one PrDyn per instance; PR-StratRace keeps placeholders in inst
    std::vector<float> work(e->plan_total), plan(e->plan_total);
    std::vector<PrDyn> dyn(n);
    HIPCHK(hipMemcpy(work.data(), e->d_work.p, work.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(plan.data(), e->d_plan.p, plan.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(dyn.data(), e->d_dyn.p, dyn.size() * sizeof(PrDyn), hipMemcpyDeviceToHost));
    const dr_launch_cfg saved = e->launch;
    dr_launch_cfg best_cfg = saved;
    float best_t = 1e30f;
    std::vector<dr_launch_cfg> cands;
    for (int rb : {0, 128, 256, 512})
        for (int ch : {512, 1024, 2048})
            for (int tg : {16, 32})
            {
                dr_launch_cfg c = saved;
                c.rollout_block = rb;
                c.reduce_chunk = ch;
                c.reduce_tgroup = tg;
                c.use_graph = 1;
                c.steps_per_graph = 1;
                cands.push_back(c);
            }
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    int rc = DR_OK;
    for (const dr_launch_cfg& c : cands)
    {
        rc = dr_pr_set_launch(e, &c);
        if (rc)
            break;
        e->drop_graphs();
        rc = e->run(1, false);
        if (rc)
            break;
        float ms = 0;
        if (hipEventRecord(a, e->stream) != hipSuccess)
            rc = host_fail(DR_EHIP, "autotune: hipEventRecord failed");
        if (!rc)
            rc = e->run(trials, false);
        if (!rc && (hipEventRecord(b, e->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                    hipEventElapsedTime(&ms, a, b) != hipSuccess))
            rc = host_fail(DR_EHIP, "autotune: event timing failed");
        if (rc)
            break;
        ms /= (float)trials;
        if (ms < best_t)
        {
            best_t = ms;
            best_cfg = c;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc)
    {
        dr_pr_set_launch(e, &saved);
        e->drop_graphs();
        return rc;
    }
    best_cfg.steps_per_graph = saved.steps_per_graph;
    if ((rc = dr_pr_set_launch(e, &best_cfg)))
        return rc;
    e->drop_graphs();
    HIPCHK(hipMemcpy(e->d_work.p, work.data(), work.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_plan.p, plan.data(), plan.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_dyn.p, dyn.data(), dyn.size() * sizeof(PrDyn), hipMemcpyHostToDevice));
    if (best)
        *best = best_cfg;
    if (best_ms)
        *best_ms = best_t;
    return DR_OK;
}

int dr_pr_get_launch(dr_prengine* e, dr_launch_cfg* cfg)
{
    if (!e || !cfg)
        return host_fail(DR_EINVAL, "NULL argument");
    *cfg = e->launch;
    return DR_OK;
}

int dr_pr_set_timing(dr_prengine* e, int on)
{
    if (!e)
        return host_fail(DR_EINVAL, "NULL handle");
    e->timing = on != 0;
    return DR_OK;
}

int dr_pr_last_timing(dr_prengine* e, float* ms, int n)
{
    if (!e || !ms || n != 5)
        return host_fail(DR_EINVAL, "last_timing: expected 5 floats");
    std::memcpy(ms, e->last_ms, sizeof(e->last_ms));
    return DR_OK;
}

int dr_prsr_create(const dr_sr_env* env, const dr_sr_cost* cost, const dr_prsr_instance* instances, int n_instances, dr_prengine** out)
{
    if (!out)
        return host_fail(DR_EINVAL, "out is NULL");
    *out = nullptr;
    if (!env || !cost || !instances)
        return host_fail(DR_EINVAL, "NULL argument");
    if (n_instances < 1)
        return host_fail(DR_EINVAL, "need at least one instance");
    if (int rc = check_sr_env(*env))
        return rc;
    for (int i = 0; i < n_instances; ++i)
    {
        const dr_prsr_instance& r = instances[i];
        const std::string w = "instance " + std::to_string(i) + ": ";
        if (r.n_samples < 1 || r.n_samples > (1 << 24))
            return host_fail(DR_EINVAL, w + "n_samples must be in [1, 2^24]");
        if (r.n_timesteps < 1 || r.n_timesteps > DR_MAX_T)
            return host_fail(DR_EINVAL, w + "n_timesteps must be in [1, 256]");
        if (r.sampled != 0 && r.sampled != 1)
            return host_fail(DR_EINVAL, w + "sampled must be 0 or 1");
        if (r.sampled && (r.n_draws < 1 || r.n_draws > DR_MAX_MODELS))
            return host_fail(DR_EINVAL, w + "n_draws must be in [1, 16]");
        if (!valid_belief(r.belief, env->n_models))
            return host_fail(DR_EINVAL, w + "belief entries must be finite and >= 0");
        if (!(r.lambda > 0.0f))
            return host_fail(DR_EINVAL, w + "lambda must be > 0");
        if (!(r.sigma >= 0.0f) || !std::isfinite(r.sigma) || !std::isfinite(r.safety_weight))
            return host_fail(DR_EINVAL, w + "sigma and safety_weight must be finite (sigma >= 0)");
        if (r.true_theta < 0 || r.true_theta >= env->n_models)
            return host_fail(DR_EINVAL, w + "true_theta out of range");
    }
    dr_prengine* e = new (std::nothrow) dr_prengine();
    if (!e)
        return host_fail(DR_ENOMEM, "host allocation failed");
    auto bail = [&](int rc) {
        delete e;
        return rc;
    };
    if (hipGetDevice(&e->device) != hipSuccess)
        return bail(host_fail(DR_EHIP, "no HIP device"));
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(host_fail(DR_EHIP, "stream creation failed"));
    e->env_kind = 1;
    e->srenv = *env;
    e->srcost = *cost;
    e->srinst.assign(instances, instances + n_instances);
    e->inst.assign(n_instances, dr_pr_instance{});
    e->n_drones = 0;
    e->launch = dr_launch_cfg{0, 0, 1024, 1, 1, 1, 16, 0};
    if (int rc = e->build_layout())
        return bail(rc);
    auto chk = [&](hipError_t x, const char* what) {
        if (x != hipSuccess)
        {
            host_fail(x == hipErrorOutOfMemory ? DR_ENOMEM : DR_EHIP, std::string(what) + ": " + hipGetErrorString(x));
            return false;
        }
        return true;
    };
    if (!chk(e->d_dyn.alloc(n_instances), "alloc dyn") || !chk(e->d_srenv.alloc(1), "alloc env") ||
        !chk(e->d_srstates.alloc(n_instances), "alloc states") || !chk(e->d_events.alloc(n_instances), "alloc events") ||
        !chk(e->d_work.alloc(e->plan_total), "alloc work") || !chk(e->d_plan.alloc(e->plan_total), "alloc plan") ||
        !chk(e->d_cand.alloc(e->plan_total), "alloc cand") || !chk(e->d_noise.alloc(e->noise_total), "alloc noise") ||
        !chk(e->d_pair.alloc(e->pair_total), "alloc pair") || !chk(e->d_agg.alloc(e->agg_total), "alloc agg") ||
        !chk(e->d_models.alloc(e->model_total), "alloc models") || !chk(e->d_actions.alloc((size_t)n_instances * 4), "alloc actions"))
        return bail(DR_ENOMEM);
    std::vector<PrDyn> dyn(n_instances);
    for (int i = 0; i < n_instances; ++i)
    {
        std::memset(&dyn[i], 0, sizeof(PrDyn));
        for (int j = 0; j < 3; ++j)
            dyn[i].lambda[j] = instances[i].lambda;
        for (int b = 0; b < env->n_models; ++b)
            dyn[i].belief[b] = instances[i].belief[b];
        dyn[i].use_nom = 1;
    }
    std::vector<float> plan(e->plan_total, 0.0f), work(e->plan_total, 0.0f);
    for (int i = 0; i < n_instances; ++i)
    {
        const PrInstDev& I = e->idev[i];
        for (int p = 0; p < 2; ++p)
            shift_plan(&plan[I.plan_off + p * I.T * 4], &work[I.plan_off + p * I.T * 4], I.T, I.shift_fill);
    }
    std::vector<dr_sr_state> st(n_instances);
    std::memset(st.data(), 0, st.size() * sizeof(dr_sr_state));
    std::vector<int> ev(n_instances, 0);
    if (!chk(hipMemcpy(e->d_dyn.p, dyn.data(), dyn.size() * sizeof(PrDyn), hipMemcpyHostToDevice), "copy dyn") ||
        !chk(hipMemcpy(e->d_srenv.p, env, sizeof(dr_sr_env), hipMemcpyHostToDevice), "copy env") ||
        !chk(hipMemcpy(e->d_srstates.p, st.data(), st.size() * sizeof(dr_sr_state), hipMemcpyHostToDevice), "copy states") ||
        !chk(hipMemcpy(e->d_events.p, ev.data(), ev.size() * sizeof(int), hipMemcpyHostToDevice), "copy events") ||
        !chk(hipMemcpy(e->d_plan.p, plan.data(), plan.size() * sizeof(float), hipMemcpyHostToDevice), "copy plan") ||
        !chk(hipMemcpy(e->d_work.p, work.data(), work.size() * sizeof(float), hipMemcpyHostToDevice), "copy work") ||
        !chk(hipMemcpy(e->d_cand.p, plan.data(), plan.size() * sizeof(float), hipMemcpyHostToDevice), "copy cand"))
        return bail(DR_EHIP);
    *out = e;
    return DR_OK;
}

int dr_pr_set_sr_states(dr_prengine* e, const dr_sr_state* states, int n)
{
    if (!e || !states || e->env_kind != 1 || n != (int)e->inst.size())
        return host_fail(DR_EINVAL, "set_sr_states: needs a StratRace2D engine and one state per instance");
    if (int rc = copy_in(e, e->d_srstates.p, states, (size_t)n))
        return rc;
    std::vector<int> ev(n, 0);
    return copy_in(e, e->d_events.p, ev.data(), ev.size());
}

int dr_pr_get_sr_states(dr_prengine* e, dr_sr_state* states, int n)
{
    if (!e || !states || e->env_kind != 1 || n != (int)e->inst.size())
        return host_fail(DR_EINVAL, "get_sr_states: needs a StratRace2D engine and one state per instance");
    return copy_out(e, states, e->d_srstates.p, (size_t)n);
}

int dr_pr_get_events(dr_prengine* e, int32_t* events, int n)
{
    if (!e || !events || e->env_kind != 1 || n != (int)e->inst.size())
        return host_fail(DR_EINVAL, "get_events: needs a StratRace2D engine and one value per instance");
    return copy_out(e, events, e->d_events.p, (size_t)n);
}

} // extern "C"
