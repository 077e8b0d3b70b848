// Host side of the control engine: descriptor tables, device buffers, the
// per-step launch sequence, hipGraph capture, the launch-shape autotuner and
// the C ABI of include/dr_api.h.
//
// Reference host orchestration this replaces:
//   MPPIController::allocDevice / getControl  src/mppi_controller.cu:143-210, 250-466
// The reference runs ~11 launches and ~10 blocking copies per controller and
// step on the default stream; here one graph replay runs the step for every
// instance of the batch, and nothing is copied unless the caller asks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "dr_host.h"
#include "dr_kernels.h"

namespace dr {

hipError_t launch_rollout(int block, int nb, int split, int grid, hipStream_t st, const InstDev* insts, InstDyn* dyn,
                          const RTile* tiles, int n_tiles, const dr_track* tracks, const dr_drone_state* drones, const float* nominal,
                          float* costs, float4* noise, float* tmin, const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC,
                          int max_T, int opp_stride);
hipError_t launch_eval(int n_inst, hipStream_t st, const InstDev* insts, InstDyn* dyn, const dr_track* tracks,
                       const dr_drone_state* drones, const float* nominal, const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC,
                       int max_T, int opp_stride);
hipError_t launch_exec(int n_inst, hipStream_t st, const InstDev* insts, InstDyn* dyn, const dr_track* tracks,
                       const dr_drone_state* drones, const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC);
hipError_t launch_wavg(int chunk, int n_wtiles, hipStream_t st, const InstDev* insts, InstDyn* dyn, const WTile* wtiles,
                       const float* costs, const float4* noise, const float* tmin, double* part, double* den);
hipError_t launch_update(int n_inst, hipStream_t st, const InstDev* insts, InstDyn* dyn, const double* part, const double* den,
                         float* nominal, float* actions, const QuadConst& C, int advance, dr_drone_state* drones, const int* drone_inst,
                         int n_drones, const dr_track* tracks, float margin);
hipError_t launch_costgrad(int lanes, hipStream_t st, const InstDev* insts, int inst, int T, int n_ck, int ck,
                           const dr_track* tracks, const dr_drone_state* drones, const float* U, int n_seq, float* J, float* G,
                           const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC, int opp_stride);
hipError_t launch_refine(int n_inst, hipStream_t st, const InstDev* insts, InstDyn* dyn, const double* part, const double* den,
                         const dr_track* tracks, const dr_drone_state* drones, float* nominal, int n_steps, float step,
                         const QuadConst& C, const CostConst& Q, const dr_cost_cfg& KC, float* actions, int advance,
                         const int* drone_inst, int n_drones, dr_drone_state* drones_out, int n_cu, size_t lds_per_cu, float4* noise,
                         int tcap);
hipError_t launch_noise(int blocks, hipStream_t st, const InstDev* insts, const InstDyn* dyn, int n_inst, float4* noise);
hipError_t refine_allow_lds();

} // namespace dr

using namespace dr;

namespace {

thread_local std::string g_err;

// HIP events per timed step: start, after the rollout, after the weighted
// average, before and after the adjoint refinement, end.
constexpr int kStepEvents = 6;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

bool valid_block(int b) { return b == 64 || b == 128 || b == 256; }
bool valid_chunk(int c) { return c >= 64 && c <= WAVG_MAX_CHUNK && c % 64 == 0; }
bool valid_nb(int b) { return b == 1 || b == 2 || b == 4; }

} // namespace

namespace dr {

int host_fail(int code, const std::string& msg) { return fail(code, msg); }

int check_track(const dr_track& t, int idx)
{
    if (t.n_gates < 2 || t.n_gates > DR_MAX_GATES)
        return fail(DR_EINVAL, "track " + std::to_string(idx) + ": n_gates must be in [2, 16]");
    for (int g = 0; g < t.n_gates; ++g)
        if (!(t.gates[g].radius > 0.0f))
            return fail(DR_EINVAL, "track " + std::to_string(idx) + ": gate radius must be > 0");
    return DR_OK;
}

int check_drones(const dr_drone_state* s, int n)
{
    for (int d = 0; d < n; ++d)
    {
        const dr_drone_state& x = s[d];
        const float vals[13] = {x.p[0], x.p[1], x.p[2], x.q[0], x.q[1], x.q[2], x.q[3], x.v[0], x.v[1], x.v[2], x.w[0], x.w[1], x.w[2]};
        for (float v : vals)
            if (!std::isfinite(v))
                return fail(DR_EINVAL, "drone " + std::to_string(d) + ": state is not finite");
        const double q2 = (double)x.q[0] * x.q[0] + (double)x.q[1] * x.q[1] + (double)x.q[2] * x.q[2] + (double)x.q[3] * x.q[3];
        if (!(std::fabs(q2 - 1.0) <= 1e-2))
            return fail(DR_EINVAL, "drone " + std::to_string(d) + ": attitude quaternion must be a unit quaternion");
    }
    return DR_OK;
}

} // namespace dr

struct dr_engine
{
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr; // forked stream of the nominal's noise-free rollout
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool eval = false;
    bool exec_cost = false; // k_exec after every closed-loop step
    int refine_steps = 0; // adjoint gradient steps per control step (0: plain MPPI)
    float refine_step = 0.05f;
    int refine_ck = 10;
    // rollout_split 4 (with the refinement on): the noise rows the next
    // rollout reads are made by the refinement's spare workgroups; true while
    // the noise buffer holds the rows of every instance's current step
    bool noise_ready = false;
    int n_cu = 256;
    size_t lds_per_cu = 160 * 1024;
    DevBuf<float> d_gU, d_gJ, d_gG; // scratch of dr_engine_cost_grad
    QuadConst C{};
    CostConst Q{};
    dr_cost_cfg cost{};
    int max_T = 1;
    int opp_stride = 1;
    std::vector<dr_track> tracks;
    std::vector<dr_instance> inst;
    std::vector<int> drone_world;
    std::vector<int> drone_inst;
    int n_drones = 0;

    dr_launch_cfg launch{};
    std::vector<InstDev> idev;
    std::vector<RTile> rtiles;
    std::vector<WTile> wtiles;
    long long noise_total = 0;
    long long part_total = 0;
    int den_total = 0;
    int cost_total = 0;
    int nom_total = 0;

    DevBuf<InstDev> d_inst;
    DevBuf<InstDyn> d_dyn;
    DevBuf<RTile> d_rtiles;
    DevBuf<WTile> d_wtiles;
    DevBuf<dr_track> d_tracks;
    // Two copies of the world's drone states.  The plain step moves the drones
    // in place; the refined step reads copy `cur` and writes the moved drones
    // into the other copy (the refinement's workgroups read every drone while
    // others finish), which then becomes current.
    DevBuf<dr_drone_state> d_drones;
    int cur = 0;
    DevBuf<int> d_drone_inst;
    DevBuf<float> d_nominal;
    DevBuf<float> d_costs;
    DevBuf<float4> d_noise;
    DevBuf<float> d_tmin;
    DevBuf<double> d_part;
    DevBuf<double> d_den;
    DevBuf<float> d_actions;

    // graphs by (advance, parity of the drone copy they start from, size):
    // size 0 holds launch.steps_per_graph steps, size 1 one step (the
    // remainder of a run, and runs shorter than steps_per_graph)
    hipGraphExec_t graph_exec[4][2] = {};
    int graph_steps[4][2] = {};

    bool timing = false;
    std::vector<hipEvent_t> events;
    int timed_steps = 0;
    float last_ms[5] = {0, 0, 0, 0, 0}; // rollout, reduce, update (+ advance), refinement, step

    ~dr_engine()
    {
        drop_graphs();
        for (auto ev : events)
            (void)hipEventDestroy(ev);
        if (ev_fork)
            (void)hipEventDestroy(ev_fork);
        if (ev_join)
            (void)hipEventDestroy(ev_join);
        if (side)
            (void)hipStreamDestroy(side);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    void drop_graphs()
    {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 2; ++j)
            {
                if (graph_exec[i][j])
                    (void)hipGraphExecDestroy(graph_exec[i][j]);
                graph_exec[i][j] = nullptr;
                graph_steps[i][j] = 0;
            }
    }

    dr_drone_state* drones_at(int p) const { return d_drones.p + (size_t)p * n_drones; }
    // whether a step moves the drones into the other copy
    bool flips(bool advance) const { return refine_steps > 0 && advance; }
    // whether the rollout reads noise rows made ahead (rollout_split 4 needs
    // the refinement, whose idle CUs make them; without it split 4 runs as 3)
    bool noise_ahead() const { return launch.rollout_split == 4 && refine_steps > 0; }
    int rollout_split() const { return launch.rollout_split == 4 && refine_steps == 0 ? 3 : launch.rollout_split; }

    int build_layout();
    int enqueue_step(bool advance, int timed_index, int p);
    hipGraphExec_t graph(bool advance, int p, int spg);
    int build_graph(bool advance, int p, int spg);
    int run(int n_steps, bool advance);
};

int dr_engine::build_layout()
{
    const int rb = launch.rollout_split ? WAVE : launch.rollout_block;
    const int chunk = launch.reduce_chunk;
    const int n = (int)inst.size();

    // Rollout tiles: instances in decreasing horizon so the longest tiles
    // are dispatched first (longest-processing-time order), each instance's
    // tiles contiguous so that its minimum is a contiguous range.
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return inst[a].n_timesteps > inst[b].n_timesteps; });

    idev.assign(n, InstDev{});
    rtiles.clear();
    wtiles.clear();
    noise_total = 0;
    part_total = 0;
    den_total = 0;
    cost_total = 0;
    nom_total = 0;
    for (int i = 0; i < n; ++i)
    {
        const dr_instance& s = inst[i];
        InstDev& I = idev[i];
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
        I.noise_off = noise_total;
        noise_total += (long long)I.T * I.K;
        I.cost_off = cost_total;
        cost_total += I.K;
        I.nom_off = nom_total;
        nom_total += I.T * 4;
        I.n_chunks = (I.K + chunk - 1) / chunk;
        I.part_off = part_total;
        part_total += (long long)I.n_chunks * I.T * 4;
        I.den_off = den_total;
        den_total += I.n_chunks;
    }
    for (int oi = 0; oi < n; ++oi)
    {
        const int i = order[oi];
        InstDev& I = idev[i];
        I.rtile_first = (int)rtiles.size();
        for (int k0 = 0; k0 < I.K; k0 += rb)
            rtiles.push_back(RTile{i, k0});
        I.rtile_count = (int)rtiles.size() - I.rtile_first;
    }
    for (int oi = 0; oi < n; ++oi)
    {
        const int i = order[oi];
        const InstDev& I = idev[i];
        const int tg = launch.reduce_tgroup;
        for (int c = 0; c < I.n_chunks; ++c)
            for (int t0 = 0; t0 < I.T; t0 += tg)
                wtiles.push_back(WTile{i, c, t0, std::min(I.T, t0 + tg)});
    }
    max_T = 1;
    opp_stride = 1;
    for (const InstDev& I : idev)
    {
        max_T = std::max(max_T, I.T);
        opp_stride = std::max(opp_stride, I.n_opp);
    }

    HIPCHK(d_inst.alloc(n));
    HIPCHK(hipMemcpy(d_inst.p, idev.data(), n * sizeof(InstDev), hipMemcpyHostToDevice));
    HIPCHK(d_rtiles.alloc(rtiles.size()));
    HIPCHK(hipMemcpy(d_rtiles.p, rtiles.data(), rtiles.size() * sizeof(RTile), hipMemcpyHostToDevice));
    HIPCHK(d_wtiles.alloc(wtiles.size()));
    HIPCHK(hipMemcpy(d_wtiles.p, wtiles.data(), wtiles.size() * sizeof(WTile), hipMemcpyHostToDevice));
    HIPCHK(d_tmin.alloc(rtiles.size()));
    HIPCHK(d_part.alloc(part_total));
    HIPCHK(d_den.alloc(den_total));
    drop_graphs();
    return DR_OK;
}

// One control step of the batch starting from drone copy p.
int dr_engine::enqueue_step(bool advance, int timed_index, int p)
{
    const int n = (int)inst.size();
    dr_drone_state* const drones = drones_at(p);
    dr_drone_state* const moved = flips(advance) ? drones_at(p ^ 1) : drones;
    hipEvent_t* ev = (timed_index >= 0) ? &events[(size_t)timed_index * kStepEvents] : nullptr;
    if (ev)
        HIPCHK(hipEventRecord(ev[0], stream));
    if (eval)
    {
        HIPCHK(hipEventRecord(ev_fork, stream));
        HIPCHK(hipStreamWaitEvent(side, ev_fork, 0));
        HIPCHK(launch_eval(n, side, d_inst.p, d_dyn.p, d_tracks.p, drones, d_nominal.p, C, Q, cost, max_T, opp_stride));
        HIPCHK(hipEventRecord(ev_join, side));
    }
    HIPCHK(launch_rollout(launch.rollout_block, launch.noise_batch, rollout_split(), launch.rollout_grid, stream, d_inst.p,
                          d_dyn.p, d_rtiles.p, (int)rtiles.size(), d_tracks.p, drones, d_nominal.p, d_costs.p, d_noise.p, d_tmin.p, C, Q,
                          cost, max_T, opp_stride));
    if (ev)
        HIPCHK(hipEventRecord(ev[1], stream));
    HIPCHK(launch_wavg(launch.reduce_chunk, (int)wtiles.size(), stream, d_inst.p, d_dyn.p, d_wtiles.p, d_costs.p, d_noise.p, d_tmin.p,
                       d_part.p, d_den.p));
    if (ev)
        HIPCHK(hipEventRecord(ev[2], stream));
    if (eval)
        HIPCHK(hipStreamWaitEvent(stream, ev_join, 0));
    if (refine_steps > 0)
    {
        // the MPPI update, the action, the shift and the closed-loop step run
        // inside the refinement kernel
        if (ev)
            HIPCHK(hipEventRecord(ev[3], stream));
        // with noise made ahead, the refinement launch also makes the next
        // step's noise rows on the CUs it leaves idle
        HIPCHK(launch_refine(n, stream, d_inst.p, d_dyn.p, d_part.p, d_den.p, d_tracks.p, drones, d_nominal.p, refine_steps,
                             refine_step, C, Q, cost, d_actions.p, advance ? 1 : 0, d_drone_inst.p, n_drones, moved, n_cu, lds_per_cu,
                             noise_ahead() ? d_noise.p : nullptr, max_T));
        if (ev)
            HIPCHK(hipEventRecord(ev[4], stream));
    }
    else
    {
        if (ev)
        {
            HIPCHK(hipEventRecord(ev[3], stream));
            HIPCHK(hipEventRecord(ev[4], stream));
        }
        HIPCHK(launch_update(n, stream, d_inst.p, d_dyn.p, d_part.p, d_den.p, d_nominal.p, d_actions.p, C, advance ? 1 : 0, drones,
                             d_drone_inst.p, n_drones, d_tracks.p, cost.gate_traversal_margin));
    }
    if (exec_cost && advance)
        HIPCHK(launch_exec(n, stream, d_inst.p, d_dyn.p, d_tracks.p, moved, C, Q, cost));
    if (ev)
        HIPCHK(hipEventRecord(ev[5], stream));
    return DR_OK;
}

// The graph of spg control steps that starts from drone copy p (captured once
// per launch shape).
hipGraphExec_t dr_engine::graph(bool advance, int p, int spg)
{
    const int gi = (advance ? 2 : 0) + p, sz = spg == 1 ? 1 : 0;
    return graph_steps[gi][sz] == spg ? graph_exec[gi][sz] : nullptr;
}

int dr_engine::build_graph(bool advance, int p, int spg)
{
    const int gi = (advance ? 2 : 0) + p, sz = spg == 1 ? 1 : 0;
    hipGraphExec_t& ge = graph_exec[gi][sz];
    if (ge && graph_steps[gi][sz] == spg)
        return DR_OK;
    if (ge)
        (void)hipGraphExecDestroy(ge);
    ge = nullptr;
    graph_steps[gi][sz] = 0;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = DR_OK;
    for (int s = 0; s < spg && rc == DR_OK; ++s)
    {
        rc = enqueue_step(advance, -1, p);
        p ^= flips(advance) ? 1 : 0;
    }
    hipError_t ce = hipStreamEndCapture(stream, &g);
    if (rc)
    {
        if (g)
            (void)hipGraphDestroy(g);
        return rc;
    }
    if (ce != hipSuccess)
        return fail(DR_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess)
        return fail(DR_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    graph_steps[gi][sz] = spg;
    return DR_OK;
}

int dr_engine::run(int n_steps, bool advance)
{
    if (n_steps < 0)
        return fail(DR_EINVAL, "n_steps < 0");
    timed_steps = 0;
    if (!noise_ahead())
        noise_ready = false; // the rollouts write the current step's rows themselves
    else if (!noise_ready && n_steps > 0)
    {
        // the first step's rows; every step then makes the next step's
        HIPCHK(launch_noise(n_cu * 4, stream, d_inst.p, d_dyn.p, (int)inst.size(), d_noise.p));
        noise_ready = true;
    }
    if (timing)
    {
        const size_t need = (size_t)n_steps * kStepEvents;
        while (events.size() < need)
        {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            events.push_back(ev);
        }
        for (int s = 0; s < n_steps; ++s)
        {
            int rc = enqueue_step(advance, s, cur);
            if (rc)
                return rc;
            cur ^= flips(advance) ? 1 : 0;
        }
        timed_steps = n_steps;
        return DR_OK;
    }
    if (!launch.use_graph)
    {
        for (int s = 0; s < n_steps; ++s)
        {
            int rc = enqueue_step(advance, -1, cur);
            if (rc)
                return rc;
            cur ^= flips(advance) ? 1 : 0;
        }
        return DR_OK;
    }
    // graphs of steps_per_graph steps, the remainder as one-step graphs; a
    // graph of an odd number of flipping steps ends on the other drone copy,
    // whose graph the next launch uses
    const int spg = std::max(1, launch.steps_per_graph);
    int done = 0;
    while (done < n_steps)
    {
        const int sz = n_steps - done >= spg ? spg : 1;
        if (int rc = build_graph(advance, cur, sz))
            return rc;
        HIPCHK(hipGraphLaunch(graph(advance, cur, sz), stream));
        done += sz;
        cur ^= (flips(advance) && (sz & 1)) ? 1 : 0;
    }
    return DR_OK;
}

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

const char* dr_last_error(void) { return g_err.c_str(); }

int dr_version(void) { return DR_ABI_VERSION; }

int dr_engine_create(const dr_quad_params* params, const dr_cost_cfg* cost, const dr_track* tracks, int n_tracks,
                     const dr_instance* instances, int n_instances, const int32_t* drone_world, int n_drones, dr_engine** out)
{
    if (!out)
        return fail(DR_EINVAL, "out is NULL");
    *out = nullptr;
    if (!params || !cost || !tracks || !instances || !drone_world)
        return fail(DR_EINVAL, "NULL argument");
    if (n_tracks < 1 || n_instances < 1 || n_drones < 1)
        return fail(DR_EINVAL, "need at least one track, instance and drone");
    if (!(params->mass > 0.0f) || !(params->dt > 0.0f) || !(params->drone_radius > 0.0f))
        return fail(DR_EINVAL, "mass, dt and drone_radius must be > 0");
    if (!(params->thrust_max >= params->thrust_min) || !(params->rate_max >= 0.0f))
        return fail(DR_EINVAL, "control limits are inconsistent");
    if (!(cost->opp_prox_radius > 0.0f) || !(cost->boundary_threshold_factor > 0.0f))
        return fail(DR_EINVAL, "opp_prox_radius and boundary_threshold_factor must be > 0");
    for (int i = 0; i < n_tracks; ++i)
    {
        int rc = check_track(tracks[i], i);
        if (rc)
            return rc;
    }
    std::vector<int> inst_of_drone(n_drones, -1);
    for (int i = 0; i < n_instances; ++i)
    {
        const dr_instance& s = instances[i];
        const std::string w = "instance " + std::to_string(i) + ": ";
        if (s.n_samples < 1 || s.n_samples > (1 << 24))
            return fail(DR_EINVAL, w + "n_samples must be in [1, 2^24]");
        if (s.n_timesteps < 1 || s.n_timesteps > DR_MAX_T)
            return fail(DR_EINVAL, w + "n_timesteps must be in [1, 256]");
        if (s.track < 0 || s.track >= n_tracks)
            return fail(DR_EINVAL, w + "track index out of range");
        if (s.drone < 0 || s.drone >= n_drones)
            return fail(DR_EINVAL, w + "drone index out of range");
        if (inst_of_drone[s.drone] >= 0)
            return fail(DR_EINVAL, w + "drone already controlled by another instance");
        if (!(s.lambda > 0.0f))
            return fail(DR_EINVAL, w + "lambda must be > 0");
        inst_of_drone[s.drone] = i;
    }
    for (int d = 0; d < n_drones; ++d)
    {
        int cnt = 0;
        for (int e = 0; e < n_drones; ++e)
            cnt += drone_world[e] == drone_world[d];
        if (cnt > DR_MAX_WORLD_DRONES)
            return fail(DR_EINVAL, "a world holds more than 8 drones");
    }

    dr_engine* e = new (std::nothrow) dr_engine();
    if (!e)
        return fail(DR_ENOMEM, "host allocation failed");
    auto bail = [&](int rc) {
        delete e;
        return rc;
    };
    if (hipGetDevice(&e->device) != hipSuccess)
        return bail(fail(DR_EHIP, "no HIP device"));
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, e->device) == hipSuccess && prop.multiProcessorCount > 0)
        {
            e->n_cu = prop.multiProcessorCount;
            if (prop.maxSharedMemoryPerMultiProcessor > 0)
                e->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
        }
    }
    if (refine_allow_lds() != hipSuccess)
        return bail(fail(DR_EHIP, "hipFuncSetAttribute (refinement LDS) failed"));
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming) != hipSuccess)
        return bail(fail(DR_EHIP, "stream or event creation failed"));
    e->C = make_quad_const(*params);
    e->Q = make_cost_const(*params, *cost);
    e->cost = *cost;
    e->tracks.assign(tracks, tracks + n_tracks);
    e->inst.assign(instances, instances + n_instances);
    e->drone_world.assign(drone_world, drone_world + n_drones);
    e->drone_inst = inst_of_drone;
    e->n_drones = n_drones;
    e->launch = dr_launch_cfg{128, 0, 1024, 1, 1, 1, 16, 1};

    int rc = e->build_layout();
    if (rc)
        return bail(rc);
    auto chk = [&](hipError_t x, const char* what) {
        if (x != hipSuccess)
        {
            fail(x == hipErrorOutOfMemory ? DR_ENOMEM : DR_EHIP, std::string(what) + ": " + hipGetErrorString(x));
            return false;
        }
        return true;
    };
    if (!chk(e->d_dyn.alloc(n_instances), "alloc dyn") || !chk(e->d_tracks.alloc(n_tracks), "alloc tracks") ||
        !chk(e->d_drones.alloc((size_t)n_drones * 2), "alloc drones") || !chk(e->d_drone_inst.alloc(n_drones), "alloc drone_inst") ||
        !chk(e->d_nominal.alloc(e->nom_total), "alloc nominal") || !chk(e->d_costs.alloc(e->cost_total), "alloc costs") ||
        !chk(e->d_noise.alloc(e->noise_total), "alloc noise") || !chk(e->d_actions.alloc((size_t)n_instances * 4), "alloc actions"))
        return bail(DR_ENOMEM);
    std::vector<InstDyn> dyn(n_instances);
    for (int i = 0; i < n_instances; ++i)
    {
        dyn[i] = InstDyn{};
        dyn[i].lambda = instances[i].lambda;
    }
    std::vector<float> nom(e->nom_total, 0.0f);
    std::vector<dr_drone_state> dr(n_drones);
    for (auto& s : dr)
    {
        std::memset(&s, 0, sizeof(s));
        s.q[0] = 1.0f;
    }
    if (!chk(hipMemcpy(e->d_dyn.p, dyn.data(), dyn.size() * sizeof(InstDyn), hipMemcpyHostToDevice), "copy dyn") ||
        !chk(hipMemcpy(e->d_tracks.p, tracks, n_tracks * sizeof(dr_track), hipMemcpyHostToDevice), "copy tracks") ||
        !chk(hipMemcpy(e->d_drones.p, dr.data(), n_drones * sizeof(dr_drone_state), hipMemcpyHostToDevice), "copy drones") ||
        !chk(hipMemcpy(e->d_drone_inst.p, inst_of_drone.data(), n_drones * sizeof(int), hipMemcpyHostToDevice), "copy drone_inst") ||
        !chk(hipMemcpy(e->d_nominal.p, nom.data(), nom.size() * sizeof(float), hipMemcpyHostToDevice), "copy nominal"))
        return bail(DR_EHIP);
    *out = e;
    return DR_OK;
}

void dr_engine_destroy(dr_engine* e)
{
    if (e)
    {
        (void)hipStreamSynchronize(e->stream);
        delete e;
    }
}

int dr_engine_set_drones(dr_engine* e, const dr_drone_state* states, int n_drones)
{
    if (!e || !states || n_drones != e->n_drones)
        return fail(DR_EINVAL, "set_drones: bad handle, pointer or count");
    if (int rc = check_drones(states, n_drones))
        return rc;
    HIPCHK(hipMemcpyAsync(e->drones_at(e->cur), states, n_drones * sizeof(dr_drone_state), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

int dr_engine_get_drones(dr_engine* e, dr_drone_state* states, int n_drones)
{
    if (!e || !states || n_drones != e->n_drones)
        return fail(DR_EINVAL, "get_drones: bad handle, pointer or count");
    HIPCHK(hipMemcpyAsync(states, e->drones_at(e->cur), n_drones * sizeof(dr_drone_state), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

static int check_inst(dr_engine* e, int i)
{
    if (!e)
        return fail(DR_EINVAL, "NULL handle");
    if (i < 0 || i >= (int)e->inst.size())
        return fail(DR_EINVAL, "instance index out of range");
    return DR_OK;
}

int dr_engine_set_nominal(dr_engine* e, int instance, const float* nominal, int n_floats)
{
    if (int rc = check_inst(e, instance))
        return rc;
    const InstDev& I = e->idev[instance];
    if (!nominal || n_floats != I.T * 4)
        return fail(DR_EINVAL, "set_nominal: expected T*4 floats");
    HIPCHK(hipMemcpyAsync(e->d_nominal.p + I.nom_off, nominal, n_floats * sizeof(float), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

int dr_engine_get_nominal(dr_engine* e, int instance, float* nominal, int n_floats)
{
    if (int rc = check_inst(e, instance))
        return rc;
    const InstDev& I = e->idev[instance];
    if (!nominal || n_floats != I.T * 4)
        return fail(DR_EINVAL, "get_nominal: expected T*4 floats");
    HIPCHK(hipMemcpyAsync(nominal, e->d_nominal.p + I.nom_off, n_floats * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

int dr_engine_set_step(dr_engine* e, int instance, uint32_t step)
{
    if (int rc = check_inst(e, instance))
        return rc;
    InstDyn y;
    HIPCHK(hipMemcpyAsync(&y, e->d_dyn.p + instance, sizeof(InstDyn), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    y.step = step;
    HIPCHK(hipMemcpyAsync(e->d_dyn.p + instance, &y, sizeof(InstDyn), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->noise_ready = false;
    return DR_OK;
}

int dr_engine_step_async(dr_engine* e, int n_steps, int advance)
{
    if (!e)
        return fail(DR_EINVAL, "NULL handle");
    return e->run(n_steps, advance != 0);
}

int dr_engine_synchronize(dr_engine* e)
{
    if (!e)
        return fail(DR_EINVAL, "NULL handle");
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->timed_steps > 0)
    {
        float acc[5] = {0, 0, 0, 0, 0};
        for (int s = 0; s < e->timed_steps; ++s)
        {
            hipEvent_t* ev = &e->events[(size_t)s * kStepEvents];
            float a, b, u1, r, u2, t;
            HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
            HIPCHK(hipEventElapsedTime(&b, ev[1], ev[2]));
            HIPCHK(hipEventElapsedTime(&u1, ev[2], ev[3]));
            HIPCHK(hipEventElapsedTime(&r, ev[3], ev[4]));
            HIPCHK(hipEventElapsedTime(&u2, ev[4], ev[5]));
            HIPCHK(hipEventElapsedTime(&t, ev[0], ev[5]));
            acc[0] += a;
            acc[1] += b;
            acc[2] += u1 + u2;
            acc[3] += r;
            acc[4] += t;
        }
        std::memcpy(e->last_ms, acc, sizeof(acc));
        e->timed_steps = 0;
    }
    return DR_OK;
}

int dr_engine_step(dr_engine* e, int n_steps, int advance, float* actions)
{
    int rc = dr_engine_step_async(e, n_steps, advance);
    if (rc)
        return rc;
    rc = dr_engine_synchronize(e);
    if (rc)
        return rc;
    if (actions)
        return dr_engine_get_actions(e, actions, (int)e->inst.size() * 4);
    return DR_OK;
}

int dr_engine_get_costs(dr_engine* e, int instance, float* costs, int n)
{
    if (int rc = check_inst(e, instance))
        return rc;
    const InstDev& I = e->idev[instance];
    if (!costs || n != I.K)
        return fail(DR_EINVAL, "get_costs: expected K floats");
    HIPCHK(hipMemcpyAsync(costs, e->d_costs.p + I.cost_off, n * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

int dr_engine_get_noise(dr_engine* e, int instance, float* noise, size_t n)
{
    if (int rc = check_inst(e, instance))
        return rc;
    const InstDev& I = e->idev[instance];
    if (!noise || n != (size_t)I.T * I.K * 4)
        return fail(DR_EINVAL, "get_noise: expected T*K*4 floats");
    HIPCHK(hipMemcpyAsync(noise, e->d_noise.p + I.noise_off, n * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

int dr_engine_get_actions(dr_engine* e, float* actions, int n_floats)
{
    if (!e || !actions || n_floats != (int)e->inst.size() * 4)
        return fail(DR_EINVAL, "get_actions: expected n_instances*4 floats");
    HIPCHK(hipMemcpyAsync(actions, e->d_actions.p, n_floats * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

int dr_engine_get_diag(dr_engine* e, dr_diag* diag, int n_instances)
{
    if (!e || !diag || n_instances != (int)e->inst.size())
        return fail(DR_EINVAL, "get_diag: expected n_instances entries");
    std::vector<InstDyn> y(n_instances);
    HIPCHK(hipMemcpyAsync(y.data(), e->d_dyn.p, n_instances * sizeof(InstDyn), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int i = 0; i < n_instances; ++i)
    {
        diag[i].min_cost = y[i].min_cost;
        diag[i].nu = y[i].nu;
        diag[i].lambda = y[i].lambda;
        diag[i].traj_cost = y[i].traj_cost;
        diag[i].step = y[i].step;
        diag[i].exec_collisions = y[i].exec_collisions;
        diag[i].exec_cost = y[i].exec_cost;
        diag[i].exec_progress = y[i].exec_progress;
    }
    return DR_OK;
}

int dr_engine_set_launch(dr_engine* e, const dr_launch_cfg* cfg)
{
    if (!e || !cfg)
        return fail(DR_EINVAL, "NULL argument");
    dr_launch_cfg c = *cfg;
    if (c.rollout_block == 0)
        c.rollout_block = 128;
    if (c.reduce_chunk == 0)
        c.reduce_chunk = 1024;
    if (c.reduce_tgroup == 0)
        c.reduce_tgroup = 16;
    if (c.noise_batch == 0)
        c.noise_batch = 1;
    if (c.steps_per_graph == 0)
        c.steps_per_graph = 1;
    if (!valid_block(c.rollout_block) || !valid_chunk(c.reduce_chunk) || !valid_nb(c.noise_batch) || c.rollout_grid < 0 ||
        c.steps_per_graph < 1 || c.steps_per_graph > 64 || c.reduce_tgroup < 1 || c.reduce_tgroup > DR_MAX_T ||
        (c.rollout_split < 0 || c.rollout_split > 4) || (c.rollout_split >= 3 && c.noise_batch > 2))
        return fail(DR_EINVAL, "set_launch: rollout_block in {64,128,256}, reduce_chunk a multiple of 64 in [64,4096], "
                               "noise_batch in {1,2,4} ({1,2} with rollout_split 3 or 4), steps_per_graph in [1,64], "
                               "reduce_tgroup in [1,256], rollout_split in {0,1,2,3,4}");
    HIPCHK(hipStreamSynchronize(e->stream));
    const bool relayout = c.rollout_block != e->launch.rollout_block || c.reduce_chunk != e->launch.reduce_chunk ||
                          c.reduce_tgroup != e->launch.reduce_tgroup || c.rollout_split != e->launch.rollout_split;
    e->launch = c;
    e->noise_ready = false;
    e->drop_graphs();
    if (relayout)
        return e->build_layout();
    return DR_OK;
}

int dr_engine_get_launch(dr_engine* e, dr_launch_cfg* cfg)
{
    if (!e || !cfg)
        return fail(DR_EINVAL, "NULL argument");
    *cfg = e->launch;
    return DR_OK;
}

int dr_engine_set_timing(dr_engine* e, int on)
{
    if (!e)
        return fail(DR_EINVAL, "NULL handle");
    e->timing = on != 0;
    return DR_OK;
}

int dr_engine_set_eval(dr_engine* e, int on)
{
    if (!e)
        return fail(DR_EINVAL, "NULL handle");
    HIPCHK(hipStreamSynchronize(e->stream));
    e->eval = on != 0;
    e->drop_graphs();
    return DR_OK;
}

static const size_t kAdjLdsBudget = 64 * 1024;

int dr_engine_set_exec_cost(dr_engine* e, int on)
{
    if (!e)
        return fail(DR_EINVAL, "NULL handle");
    HIPCHK(hipStreamSynchronize(e->stream));
    if (on)
    {
        const int n = (int)e->inst.size();
        std::vector<InstDyn> y(n);
        HIPCHK(hipMemcpy(y.data(), e->d_dyn.p, n * sizeof(InstDyn), hipMemcpyDeviceToHost));
        for (InstDyn& d : y)
        {
            d.exec_cost = 0.0f;
            d.exec_progress = 0.0f;
            d.exec_collisions = 0u;
        }
        HIPCHK(hipMemcpy(e->d_dyn.p, y.data(), n * sizeof(InstDyn), hipMemcpyHostToDevice));
    }
    if (e->exec_cost != (on != 0))
    {
        e->exec_cost = on != 0;
        e->drop_graphs();
    }
    return DR_OK;
}

int dr_engine_set_refine(dr_engine* e, int n_grad_steps, float step, int ckpt_every)
{
    if (!e)
        return fail(DR_EINVAL, "NULL handle");
    if (n_grad_steps < 0 || n_grad_steps > 100 || !(step >= 0.0f) || ckpt_every < 1 || ckpt_every > DR_MAX_T)
        return fail(DR_EINVAL, "set_refine: n_grad_steps in [0,100], step >= 0, ckpt_every in [1,256]");
    HIPCHK(hipStreamSynchronize(e->stream));
    e->refine_steps = n_grad_steps;
    e->refine_step = step;
    e->refine_ck = ckpt_every;
    e->noise_ready = false;
    e->drop_graphs();
    return DR_OK;
}

int dr_engine_cost_grad_device(dr_engine* e, int instance, const float* U, int n_seq, size_t n_floats, float* J, float* dJdU,
                               int ckpt_every, void* stream)
{
    if (int rc = check_inst(e, instance))
        return rc;
    if (!U || !J || n_seq < 1 || ckpt_every < 1)
        return fail(DR_EINVAL, "cost_grad: bad pointers or sizes");
    const InstDev& I = e->idev[instance];
    if (n_floats != (size_t)n_seq * I.T * 4)
        return fail(DR_EINVAL, "cost_grad: n_floats must be n_seq * T * 4 of the instance (U and dJdU are [n_seq][T][4])");
    const int ck = std::min(ckpt_every, I.T);
    const int n_ck = (I.T + ck - 1) / ck;
    const size_t table = (size_t)I.T * e->opp_stride * sizeof(float4);
    const size_t per_lane = (size_t)(n_ck + ck) * 4 * sizeof(float4);
    int lanes = 64;
    while (lanes > 1 && table + lanes * per_lane > kAdjLdsBudget)
        lanes /= 2;
    if (table + lanes * per_lane > kAdjLdsBudget)
        return fail(DR_EINVAL, "cost_grad: horizon and checkpoint interval need too much LDS");
    hipStream_t st = stream ? (hipStream_t)stream : e->stream;
    HIPCHK(launch_costgrad(lanes, st, e->d_inst.p, instance, I.T, n_ck, ck, e->d_tracks.p, e->drones_at(e->cur), U, n_seq, J, dJdU, e->C, e->Q,
                           e->cost, e->opp_stride));
    return DR_OK;
}

int dr_engine_cost_grad(dr_engine* e, int instance, const float* U, int n_seq, size_t n_floats, float* J, float* dJdU, int ckpt_every)
{
    if (int rc = check_inst(e, instance))
        return rc;
    if (!U || !J || n_seq < 1)
        return fail(DR_EINVAL, "cost_grad: bad pointers or sizes");
    const size_t n = (size_t)n_seq * e->idev[instance].T * 4;
    if (n_floats != n)
        return fail(DR_EINVAL, "cost_grad: n_floats must be n_seq * T * 4 of the instance (U and dJdU are [n_seq][T][4])");
    if (e->d_gU.n < n)
    {
        HIPCHK(e->d_gU.alloc(n));
        HIPCHK(e->d_gG.alloc(n));
    }
    if (e->d_gJ.n < (size_t)n_seq)
        HIPCHK(e->d_gJ.alloc(n_seq));
    HIPCHK(hipMemcpyAsync(e->d_gU.p, U, n * sizeof(float), hipMemcpyHostToDevice, e->stream));
    int rc = dr_engine_cost_grad_device(e, instance, e->d_gU.p, n_seq, n, e->d_gJ.p, dJdU ? e->d_gG.p : nullptr, ckpt_every, nullptr);
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(J, e->d_gJ.p, n_seq * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    if (dJdU)
        HIPCHK(hipMemcpyAsync(dJdU, e->d_gG.p, n * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

// Staging buffer of dr_engine_cost_grad_noalloc: one per GPU, held for the
// life of the process, never allocated or freed by a call.  Host code uses
// only its per-device address from hipGetSymbolAddress (the array's name is a
// host-side stand-in, not a device address).
static __device__ __align__(256) float scratch[DR_COSTGRAD_NOALLOC_MAX_FLOATS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// The address of this device's staging buffer, looked up once per device with
// that device current.
int scratch_on(int device, float** out)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(DR_EINVAL, "cost_grad_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device])
    {
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        if (cur != device)
            HIPCHK(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device)
            HIPCHK(hipSetDevice(cur));
        if (err != hipSuccess || !p)
            return fail(DR_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
        g_scratch_ptr[device] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[device];
    return DR_OK;
}

} // namespace

int dr_engine_cost_grad_noalloc(dr_engine* e, int instance, const float* U, int n_seq, size_t n_floats, float* J, float* dJdU,
                                int ckpt_every)
{
    if (int rc = check_inst(e, instance))
        return rc;
    if (!U || !J || n_seq < 1)
        return fail(DR_EINVAL, "cost_grad_noalloc: bad pointers or sizes");
    const size_t n = (size_t)n_seq * e->idev[instance].T * 4;
    if (n_floats != n)
        return fail(DR_EINVAL, "cost_grad_noalloc: n_floats must be n_seq * T * 4 of the instance (U and dJdU are [n_seq][T][4])");
    if (2 * n + (size_t)n_seq > (size_t)DR_COSTGRAD_NOALLOC_MAX_FLOATS)
        return fail(DR_EINVAL, "cost_grad_noalloc: n_seq * T * 8 + n_seq exceeds the staging buffer of " +
                                   std::to_string(DR_COSTGRAD_NOALLOC_MAX_FLOATS) + " floats");
    float* buf = nullptr;
    if (int rc = scratch_on(e->device, &buf))
        return rc;
    float* dU = buf;
    float* dG = buf + n;
    float* dJ = buf + 2 * n;
    HIPCHK(hipMemcpyAsync(dU, U, n * sizeof(float), hipMemcpyHostToDevice, e->stream));
    int rc = dr_engine_cost_grad_device(e, instance, dU, n_seq, n, dJ, dJdU ? dG : nullptr, ckpt_every, nullptr);
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(J, dJ, n_seq * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    if (dJdU)
        HIPCHK(hipMemcpyAsync(dJdU, dG, n * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

int dr_engine_last_timing(dr_engine* e, float* ms, int n)
{
    if (!e || !ms || n < 1 || n > 5)
        return fail(DR_EINVAL, "last_timing: bad handle, pointer or count (1..5)");
    std::memcpy(ms, e->last_ms, n * sizeof(float));
    return DR_OK;
}

int dr_engine_autotune(dr_engine* e, int trials, dr_launch_cfg* best, float* best_ms)
{
    if (!e || trials < 1)
        return fail(DR_EINVAL, "autotune: bad handle or trials");
    // Save the dynamic state the timing runs will overwrite.
    HIPCHK(hipStreamSynchronize(e->stream));
    const int n = (int)e->inst.size();
    std::vector<float> nom(e->nom_total);
    std::vector<InstDyn> dyn(n);
    HIPCHK(hipMemcpy(nom.data(), e->d_nominal.p, nom.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(dyn.data(), e->d_dyn.p, dyn.size() * sizeof(InstDyn), hipMemcpyDeviceToHost));

    const dr_launch_cfg saved = e->launch;
    dr_launch_cfg best_cfg = saved;
    float best_t = 1e30f;
    // Candidate shapes: the rollout structures (one lane per sample, noise
    // wave + dynamics wave, four- and eight-stage pipelines, and with the
    // refinement on the eight-wave pipeline over noise made ahead) x noise
    // batch x weighted-average chunk x timestep group.
    std::vector<dr_launch_cfg> cands;
    for (int split : {0, 1, 2, 3, 4})
        for (int nb : {1, 2, 4})
            for (int ch : {512, 1024, 2048})
                for (int tg : {16, 32})
                    for (int rb : {128, 256})
                    {
                        if ((split != 0 && rb != 128) || (split == 3 && nb > 2) || (split == 4 && (nb > 2 || e->refine_steps == 0)))
                            continue;
                        dr_launch_cfg c = saved;
                        c.rollout_split = split;
                        c.noise_batch = nb;
                        c.reduce_chunk = ch;
                        c.reduce_tgroup = tg;
                        c.rollout_block = rb;
                        c.rollout_grid = 0;
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
        rc = dr_engine_set_launch(e, &c);
        if (rc)
            break;
        rc = e->run(1, false); // warm-up and graph build
        if (rc)
            break;
        float ms = 0;
        if (hipEventRecord(a, e->stream) != hipSuccess)
            rc = fail(DR_EHIP, "autotune: hipEventRecord failed");
        if (!rc)
            rc = e->run(trials, false);
        if (!rc && (hipEventRecord(b, e->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                    hipEventElapsedTime(&ms, a, b) != hipSuccess))
            rc = fail(DR_EHIP, "autotune: event timing failed");
        if (rc)
            break;
        ms /= (float)trials;
        if (ms < best_t)
        {
            best_t = ms;
            best_cfg = c;
        }
    }
    // Then the steps per graph replay for the best shape: consecutive replays
    // of one graph leave a gap on the device that the steps inside a graph do
    // not have (runs shorter than a graph use one-step graphs).  A longer graph
    // is kept only if it is at least 2 % faster per step.
    for (int spg : {2, 4, 8, 16})
    {
        if (rc)
            break;
        dr_launch_cfg c = best_cfg;
        c.steps_per_graph = spg;
        rc = dr_engine_set_launch(e, &c);
        if (rc)
            break;
        rc = e->run(spg, false); // warm-up and graph build
        if (rc)
            break;
        const int n_run = ((std::max(trials, 2 * spg) + spg - 1) / spg) * spg;
        float ms = 0;
        if (hipEventRecord(a, e->stream) != hipSuccess)
            rc = fail(DR_EHIP, "autotune: hipEventRecord failed");
        if (!rc)
            rc = e->run(n_run, false);
        if (!rc && (hipEventRecord(b, e->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                    hipEventElapsedTime(&ms, a, b) != hipSuccess))
            rc = fail(DR_EHIP, "autotune: event timing failed");
        if (rc)
            break;
        ms /= (float)n_run;
        if (ms < 0.98f * best_t)
        {
            best_t = ms;
            best_cfg = c;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc)
    {
        dr_engine_set_launch(e, &saved);
        return rc;
    }
    rc = dr_engine_set_launch(e, &best_cfg);
    if (rc)
        return rc;
    HIPCHK(hipMemcpy(e->d_nominal.p, nom.data(), nom.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_dyn.p, dyn.data(), dyn.size() * sizeof(InstDyn), hipMemcpyHostToDevice));
    e->noise_ready = false; // the step counters are back where they were
    if (best)
        *best = best_cfg;
    if (best_ms)
        *best_ms = best_t;
    return DR_OK;
}

// ---------------------------------------------------------------------------
// single controller

struct dr_mppi
{
    dr_engine* e = nullptr;
    int max_opp = 0;
    int T = 0;
};

int dr_mppi_create(const dr_quad_params* params, const dr_cost_cfg* cost, const dr_track* track, const dr_mppi_cfg* cfg, uint64_t seed,
                   dr_mppi** out)
{
    if (!out || !cfg)
        return fail(DR_EINVAL, "NULL argument");
    *out = nullptr;
    if (cfg->max_opponents < 0 || cfg->max_opponents > DR_MAX_WORLD_DRONES - 1)
        return fail(DR_EINVAL, "max_opponents must be in [0, 7]");
    dr_instance in{};
    in.n_samples = cfg->n_samples;
    in.n_timesteps = cfg->n_timesteps;
    in.track = 0;
    in.drone = 0;
    in.seed = seed;
    in.lambda = cfg->lambda;
    for (int d = 0; d < 4; ++d)
    {
        in.sigma[d] = cfg->sigma[d];
        in.shift_fill[d] = cfg->shift_fill[d];
    }
    in.adapt_lambda = cfg->adapt_lambda;
    const int nd = 1 + cfg->max_opponents;
    std::vector<int32_t> world(nd, 0);
    dr_engine* e = nullptr;
    int rc = dr_engine_create(params, cost, track, 1, &in, 1, world.data(), nd, &e);
    if (rc)
        return rc;
    dr_mppi* h = new (std::nothrow) dr_mppi();
    if (!h)
    {
        dr_engine_destroy(e);
        return fail(DR_ENOMEM, "host allocation failed");
    }
    h->e = e;
    h->max_opp = cfg->max_opponents;
    h->T = cfg->n_timesteps;
    *out = h;
    return DR_OK;
}

void dr_mppi_destroy(dr_mppi* h)
{
    if (h)
    {
        dr_engine_destroy(h->e);
        delete h;
    }
}

int dr_mppi_step(dr_mppi* h, const dr_drone_state* state, const dr_opponent* opponents, int n_opponents, float* out_action)
{
    if (!h || !state || !out_action)
        return fail(DR_EINVAL, "NULL argument");
    if (n_opponents != h->max_opp || (n_opponents > 0 && !opponents))
        return fail(DR_EINVAL, "mppi_step: n_opponents must equal max_opponents given at creation");
    std::vector<dr_drone_state> w(1 + h->max_opp);
    w[0] = *state;
    for (int j = 0; j < h->max_opp; ++j)
    {
        std::memset(&w[1 + j], 0, sizeof(dr_drone_state));
        w[1 + j].q[0] = 1.0f;
        for (int d = 0; d < 3; ++d)
        {
            w[1 + j].p[d] = opponents[j].p[d];
            w[1 + j].v[d] = opponents[j].v[d];
        }
    }
    int rc = dr_engine_set_drones(h->e, w.data(), (int)w.size());
    if (rc)
        return rc;
    return dr_engine_step(h->e, 1, 0, out_action);
}

int dr_mppi_get_nominal(dr_mppi* h, float* nominal, int n_floats)
{
    if (!h)
        return fail(DR_EINVAL, "NULL handle");
    return dr_engine_get_nominal(h->e, 0, nominal, n_floats);
}

int dr_mppi_set_nominal(dr_mppi* h, const float* nominal, int n_floats)
{
    if (!h)
        return fail(DR_EINVAL, "NULL handle");
    return dr_engine_set_nominal(h->e, 0, nominal, n_floats);
}

int dr_mppi_get_diag(dr_mppi* h, dr_diag* diag)
{
    if (!h)
        return fail(DR_EINVAL, "NULL handle");
    return dr_engine_get_diag(h->e, diag, 1);
}

dr_engine* dr_mppi_engine(dr_mppi* h) { return h ? h->e : nullptr; }

} // extern "C"
