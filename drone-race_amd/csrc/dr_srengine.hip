// Host side of the StratRace2D plain-MPPI engine: descriptor tables, device
// buffers, the per-step launch sequence, hipGraph capture and the dr_sr_*
// C ABI of include/dr_api.h.
//
// Reference host orchestration this replaces (dense, one committed model):
//   MPPIController::getControl  src/mppi_controller.cu:250-466
//   SimulationEngine::dynStep   src/engine.cpp:286-312
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "dr_host.h"
#include "dr_srkernels.h"

namespace dr {

hipError_t launch_wavg(int chunk, int n_wtiles, hipStream_t st, const InstDev* insts, InstDyn* dyn, const WTile* wtiles,
                       const float* costs, const float4* noise, const float* tmin, double* part, double* den);
hipError_t launch_sr_rollout(int block, int grid, hipStream_t st, const InstDev* insts, const InstDyn* dyn, const SrInstDev* sinst,
                             const RTile* tiles, int n_tiles, const dr_sr_env* env, const dr_sr_cost& KC, const dr_sr_state* states,
                             const float* nominal, float* costs, float4* noise, float* tmin);
hipError_t launch_sr_update(int n_inst, hipStream_t st, const InstDev* insts, InstDyn* dyn, const SrInstDev* sinst, const double* part,
                            const double* den, float* nominal, float* actions, float max_accel);
hipError_t launch_sr_advance(int n_inst, hipStream_t st, const InstDev* insts, const InstDyn* dyn, const SrInstDev* sinst,
                             const dr_sr_env* env, dr_sr_state* states, const float* actions, int* events);

} // namespace dr

using namespace dr;

struct dr_sr_engine
{
    int device = 0;
    hipStream_t stream = nullptr;
    dr_sr_env env{};
    dr_sr_cost cost{};
    std::vector<dr_sr_instance> inst;
    dr_launch_cfg launch{};
    std::vector<InstDev> idev;
    std::vector<SrInstDev> sdev;
    std::vector<RTile> rtiles;
    std::vector<WTile> wtiles;
    long long noise_total = 0, part_total = 0;
    int cost_total = 0, nom_total = 0, den_total = 0;

    DevBuf<InstDev> d_inst;
    DevBuf<InstDyn> d_dyn;
    DevBuf<SrInstDev> d_sinst;
    DevBuf<RTile> d_rtiles;
    DevBuf<WTile> d_wtiles;
    DevBuf<dr_sr_env> d_env;
    DevBuf<dr_sr_state> d_states;
    DevBuf<float> d_nominal, d_costs, d_tmin, d_actions;
    DevBuf<float4> d_noise;
    DevBuf<double> d_part, d_den;
    DevBuf<int> d_events;

    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    int graph_steps[2] = {0, 0};
    bool timing = false;
    std::vector<hipEvent_t> events;
    int timed_steps = 0;
    float last_ms[4] = {0, 0, 0, 0};

    ~dr_sr_engine()
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
    int build_layout();
    int enqueue_step(bool advance, int timed_index);
    int run(int n_steps, bool advance);
};

int dr_sr_engine::build_layout()
{
    const int n = (int)inst.size();
    const int rb = launch.rollout_block, chunk = launch.reduce_chunk;
    idev.assign(n, InstDev{});
    sdev.assign(n, SrInstDev{});
    noise_total = part_total = 0;
    cost_total = nom_total = den_total = 0;
    for (int i = 0; i < n; ++i)
    {
        const dr_sr_instance& s = inst[i];
        InstDev& I = idev[i];
        I.K = s.n_samples;
        I.T = s.n_timesteps;
        I.key0 = (uint32_t)(s.seed & 0xffffffffu);
        I.key1 = (uint32_t)(s.seed >> 32);
        I.noise_off = noise_total;
        noise_total += (long long)I.K * I.T;
        I.cost_off = cost_total;
        cost_total += I.K;
        I.nom_off = nom_total;
        nom_total += I.T * 2;
        I.n_chunks = (I.K + chunk - 1) / chunk;
        I.part_off = part_total;
        part_total += (long long)I.n_chunks * I.T * 4;
        I.den_off = den_total;
        den_total += I.n_chunks;
        SrInstDev& S = sdev[i];
        S.theta = s.true_theta;
        S.sigma = s.sigma;
        S.fill[0] = s.shift_fill[0];
        S.fill[1] = s.shift_fill[1];
        S.adapt = s.adapt_lambda;
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return inst[a].n_timesteps > inst[b].n_timesteps; });
    rtiles.clear();
    wtiles.clear();
    for (int oi = 0; oi < n; ++oi)
    {
        const int i = order[oi];
        InstDev& I = idev[i];
        I.rtile_first = (int)rtiles.size();
        for (int k0 = 0; k0 < I.K; k0 += rb)
            rtiles.push_back(RTile{i, k0});
        I.rtile_count = (int)rtiles.size() - I.rtile_first;
        for (int c = 0; c < I.n_chunks; ++c)
            for (int t0 = 0; t0 < I.T; t0 += launch.reduce_tgroup)
                wtiles.push_back(WTile{i, c, t0, std::min(I.T, t0 + launch.reduce_tgroup)});
    }
    HIPCHK(d_inst.alloc(n));
    HIPCHK(hipMemcpy(d_inst.p, idev.data(), n * sizeof(InstDev), hipMemcpyHostToDevice));
    HIPCHK(d_sinst.alloc(n));
    HIPCHK(hipMemcpy(d_sinst.p, sdev.data(), n * sizeof(SrInstDev), hipMemcpyHostToDevice));
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

int dr_sr_engine::enqueue_step(bool advance, int timed_index)
{
    const int n = (int)inst.size();
    hipEvent_t* ev = (timed_index >= 0) ? &events[(size_t)timed_index * 4] : nullptr;
    if (ev)
        HIPCHK(hipEventRecord(ev[0], stream));
    HIPCHK(launch_sr_rollout(launch.rollout_block, launch.rollout_grid, stream, d_inst.p, d_dyn.p, d_sinst.p, d_rtiles.p, (int)rtiles.size(),
                             d_env.p, cost, d_states.p, d_nominal.p, d_costs.p, d_noise.p, d_tmin.p));
    if (ev)
        HIPCHK(hipEventRecord(ev[1], stream));
    HIPCHK(launch_wavg(launch.reduce_chunk, (int)wtiles.size(), stream, d_inst.p, d_dyn.p, d_wtiles.p, d_costs.p, d_noise.p, d_tmin.p, d_part.p,
                       d_den.p));
    if (ev)
        HIPCHK(hipEventRecord(ev[2], stream));
    HIPCHK(launch_sr_update(n, stream, d_inst.p, d_dyn.p, d_sinst.p, d_part.p, d_den.p, d_nominal.p, d_actions.p, env.max_accel[0]));
    if (advance)
        HIPCHK(launch_sr_advance(n, stream, d_inst.p, d_dyn.p, d_sinst.p, d_env.p, d_states.p, d_actions.p, d_events.p));
    if (ev)
        HIPCHK(hipEventRecord(ev[3], stream));
    return DR_OK;
}

int dr_sr_engine::run(int n_steps, bool advance)
{
    if (n_steps < 0)
        return host_fail(DR_EINVAL, "n_steps < 0");
    timed_steps = 0;
    if (timing)
    {
        while (events.size() < (size_t)n_steps * 4)
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

namespace {

int check_sr(dr_sr_engine* e, int i)
{
    if (!e)
        return host_fail(DR_EINVAL, "NULL handle");
    if (i < 0 || i >= (int)e->inst.size())
        return host_fail(DR_EINVAL, "instance index out of range");
    return DR_OK;
}

template <class T>
int sr_copy_out(dr_sr_engine* e, T* dst, const T* src, size_t n)
{
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

template <class T>
int sr_copy_in(dr_sr_engine* e, T* dst, const T* src, size_t n)
{
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

bool finite_pos(float x) { return std::isfinite(x) && x > 0.0f; }

} // namespace

namespace dr {

int check_sr_env(const dr_sr_env& env)
{
    if (!finite_pos(env.dt) || !finite_pos(env.drone_radius) || !finite_pos(env.track_length) || !finite_pos(env.track_width))
        return host_fail(DR_EINVAL, "dt, drone_radius, track_length and track_width must be finite and > 0");
    for (int a = 0; a < 2; ++a)
        if (!finite_pos(env.max_speed[a]) || !finite_pos(env.max_accel[a]))
            return host_fail(DR_EINVAL, "max_speed and max_accel must be finite and > 0");
    if (env.n_models < 1 || env.n_models > DR_SR_MAX_MODELS)
        return host_fail(DR_EINVAL, "n_models must be in [1, 4]");
    for (int i = 0; i < DR_SR_TRACK_SAMPLES; ++i)
        if (!std::isfinite(env.p_grid[i][0]) || !std::isfinite(env.p_grid[i][1]) || !std::isfinite(env.kappa_grid[i]) ||
            !(env.t_grid[i][0] * env.t_grid[i][0] + env.t_grid[i][1] * env.t_grid[i][1] > 0.0f))
            return host_fail(DR_EINVAL, "track grids must be finite with nonzero tangents");
    return DR_OK;
}

} // namespace dr

extern "C" {

int dr_sr_create(const dr_sr_env* env, const dr_sr_cost* cost, const dr_sr_instance* instances, int n_instances, dr_sr_engine** out)
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
        const dr_sr_instance& s = instances[i];
        const std::string w = "instance " + std::to_string(i) + ": ";
        if (s.n_samples < 1 || s.n_samples > (1 << 24))
            return host_fail(DR_EINVAL, w + "n_samples must be in [1, 2^24]");
        if (s.n_timesteps < 1 || s.n_timesteps > DR_MAX_T)
            return host_fail(DR_EINVAL, w + "n_timesteps must be in [1, 256]");
        if (s.true_theta < 0 || s.true_theta >= env->n_models)
            return host_fail(DR_EINVAL, w + "true_theta out of range");
        if (!(s.lambda > 0.0f))
            return host_fail(DR_EINVAL, w + "lambda must be > 0");
        if (!(s.sigma >= 0.0f) || !std::isfinite(s.sigma))
            return host_fail(DR_EINVAL, w + "sigma must be finite and >= 0");
    }

    dr_sr_engine* e = new (std::nothrow) dr_sr_engine();
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
    e->env = *env;
    e->cost = *cost;
    e->inst.assign(instances, instances + n_instances);
    e->launch = dr_launch_cfg{128, 0, 1024, 1, 1, 1, 16, 0};
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
    if (!chk(e->d_dyn.alloc(n_instances), "alloc dyn") || !chk(e->d_env.alloc(1), "alloc env") ||
        !chk(e->d_states.alloc(n_instances), "alloc states") || !chk(e->d_nominal.alloc(e->nom_total), "alloc nominal") ||
        !chk(e->d_costs.alloc(e->cost_total), "alloc costs") || !chk(e->d_noise.alloc(e->noise_total), "alloc noise") ||
        !chk(e->d_actions.alloc((size_t)n_instances * 2), "alloc actions") || !chk(e->d_events.alloc(n_instances), "alloc events"))
        return bail(DR_ENOMEM);
    std::vector<InstDyn> dyn(n_instances);
    for (int i = 0; i < n_instances; ++i)
    {
        dyn[i] = InstDyn{};
        dyn[i].lambda = instances[i].lambda;
    }
    std::vector<float> nom(e->nom_total, 0.0f);
    std::vector<dr_sr_state> st(n_instances);
    std::memset(st.data(), 0, st.size() * sizeof(dr_sr_state));
    std::vector<int> ev(n_instances, 0);
    if (!chk(hipMemcpy(e->d_dyn.p, dyn.data(), dyn.size() * sizeof(InstDyn), hipMemcpyHostToDevice), "copy dyn") ||
        !chk(hipMemcpy(e->d_env.p, env, sizeof(dr_sr_env), hipMemcpyHostToDevice), "copy env") ||
        !chk(hipMemcpy(e->d_states.p, st.data(), st.size() * sizeof(dr_sr_state), hipMemcpyHostToDevice), "copy states") ||
        !chk(hipMemcpy(e->d_nominal.p, nom.data(), nom.size() * sizeof(float), hipMemcpyHostToDevice), "copy nominal") ||
        !chk(hipMemcpy(e->d_events.p, ev.data(), ev.size() * sizeof(int), hipMemcpyHostToDevice), "copy events"))
        return bail(DR_EHIP);
    *out = e;
    return DR_OK;
}

void dr_sr_destroy(dr_sr_engine* e)
{
    if (e)
    {
        (void)hipStreamSynchronize(e->stream);
        delete e;
    }
}

int dr_sr_set_states(dr_sr_engine* e, const dr_sr_state* states, int n)
{
    if (!e || !states || n != (int)e->inst.size())
        return host_fail(DR_EINVAL, "set_states: bad handle, pointer or count");
    if (int rc = sr_copy_in(e, e->d_states.p, states, (size_t)n))
        return rc;
    std::vector<int> ev(n, 0);
    return sr_copy_in(e, e->d_events.p, ev.data(), ev.size());
}

int dr_sr_get_states(dr_sr_engine* e, dr_sr_state* states, int n)
{
    if (!e || !states || n != (int)e->inst.size())
        return host_fail(DR_EINVAL, "get_states: bad handle, pointer or count");
    return sr_copy_out(e, states, e->d_states.p, (size_t)n);
}

int dr_sr_set_nominal(dr_sr_engine* e, int instance, const float* nominal, int n_floats)
{
    if (int rc = check_sr(e, instance))
        return rc;
    const InstDev& I = e->idev[instance];
    if (!nominal || n_floats != I.T * 2)
        return host_fail(DR_EINVAL, "set_nominal: expected T*2 floats");
    return sr_copy_in(e, e->d_nominal.p + I.nom_off, nominal, (size_t)n_floats);
}

int dr_sr_get_nominal(dr_sr_engine* e, int instance, float* nominal, int n_floats)
{
    if (int rc = check_sr(e, instance))
        return rc;
    const InstDev& I = e->idev[instance];
    if (!nominal || n_floats != I.T * 2)
        return host_fail(DR_EINVAL, "get_nominal: expected T*2 floats");
    return sr_copy_out(e, nominal, e->d_nominal.p + I.nom_off, (size_t)n_floats);
}

int dr_sr_step(dr_sr_engine* e, int n_steps, int advance, float* actions)
{
    if (!e)
        return host_fail(DR_EINVAL, "NULL handle");
    if (int rc = e->run(n_steps, advance != 0))
        return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->timed_steps > 0)
    {
        float acc[4] = {0, 0, 0, 0};
        for (int s = 0; s < e->timed_steps; ++s)
        {
            hipEvent_t* ev = &e->events[(size_t)s * 4];
            for (int q = 0; q < 3; ++q)
            {
                float v;
                HIPCHK(hipEventElapsedTime(&v, ev[q], ev[q + 1]));
                acc[q] += v;
            }
            float t;
            HIPCHK(hipEventElapsedTime(&t, ev[0], ev[3]));
            acc[3] += t;
        }
        std::memcpy(e->last_ms, acc, sizeof(acc));
        e->timed_steps = 0;
    }
    if (actions)
        return sr_copy_out(e, actions, e->d_actions.p, e->inst.size() * 2);
    return DR_OK;
}

int dr_sr_get_events(dr_sr_engine* e, int32_t* events, int n)
{
    if (!e || !events || n != (int)e->inst.size())
        return host_fail(DR_EINVAL, "get_events: bad handle, pointer or count");
    return sr_copy_out(e, events, e->d_events.p, (size_t)n);
}

int dr_sr_get_costs(dr_sr_engine* e, int instance, float* costs, int n)
{
    if (int rc = check_sr(e, instance))
        return rc;
    const InstDev& I = e->idev[instance];
    if (!costs || n != I.K)
        return host_fail(DR_EINVAL, "get_costs: expected K floats");
    return sr_copy_out(e, costs, e->d_costs.p + I.cost_off, (size_t)n);
}

int dr_sr_get_noise(dr_sr_engine* e, int instance, float* noise, size_t n)
{
    if (int rc = check_sr(e, instance))
        return rc;
    const InstDev& I = e->idev[instance];
    if (!noise || n != (size_t)I.K * I.T * 2)
        return host_fail(DR_EINVAL, "get_noise: expected K*T*2 floats");
    std::vector<float4> rows((size_t)I.K * I.T);
    if (int rc = sr_copy_out(e, rows.data(), e->d_noise.p + I.noise_off, rows.size()))
        return rc;
    for (size_t i = 0; i < rows.size(); ++i)
    {
        noise[2 * i] = rows[i].x;
        noise[2 * i + 1] = rows[i].y;
    }
    return DR_OK;
}

int dr_sr_get_diag(dr_sr_engine* e, dr_diag* diag, int n)
{
    if (!e || !diag || n != (int)e->inst.size())
        return host_fail(DR_EINVAL, "get_diag: bad handle, pointer or count");
    std::vector<InstDyn> dyn(n);
    if (int rc = sr_copy_out(e, dyn.data(), e->d_dyn.p, dyn.size()))
        return rc;
    for (int i = 0; i < n; ++i)
    {
        std::memset(&diag[i], 0, sizeof(dr_diag));
        diag[i].min_cost = dyn[i].min_cost;
        diag[i].nu = dyn[i].nu;
        diag[i].lambda = dyn[i].lambda;
        diag[i].step = dyn[i].step;
    }
    return DR_OK;
}

int dr_sr_set_launch(dr_sr_engine* e, const dr_launch_cfg* cfg)
{
    if (!e || !cfg)
        return host_fail(DR_EINVAL, "NULL argument");
    if (cfg->rollout_block != 64 && cfg->rollout_block != 128 && cfg->rollout_block != 256)
        return host_fail(DR_EINVAL, "rollout_block must be 64, 128 or 256");
    if (cfg->rollout_grid < 0)
        return host_fail(DR_EINVAL, "rollout_grid must be >= 0");
    if (cfg->reduce_chunk < 64 || cfg->reduce_chunk > WAVG_MAX_CHUNK || cfg->reduce_chunk % 64 != 0)
        return host_fail(DR_EINVAL, "reduce_chunk must be a multiple of 64 in [64, 4096]");
    if (cfg->reduce_tgroup < 1 || cfg->reduce_tgroup > WAVG_BLOCK)
        return host_fail(DR_EINVAL, "reduce_tgroup must be in [1, 256]");
    if (cfg->steps_per_graph < 1)
        return host_fail(DR_EINVAL, "steps_per_graph must be >= 1");
    HIPCHK(hipStreamSynchronize(e->stream));
    e->launch = *cfg;
    return e->build_layout();
}

// Times the candidate launch shapes on the engine's own workload (trials
// steps each after one warm-up step, no world step) and keeps the fastest;
// the nominals and per-instance state the timing steps change are restored.
// Candidates: rollout block x weighted-average chunk x timestep group.
int dr_sr_autotune(dr_sr_engine* e, int trials, dr_launch_cfg* best, float* best_ms)
{
    if (!e || trials < 1)
        return host_fail(DR_EINVAL, "autotune: bad handle or trials");
    HIPCHK(hipStreamSynchronize(e->stream));
    std::vector<float> nom(e->nom_total);
    std::vector<InstDyn> dyn(e->inst.size());
    HIPCHK(hipMemcpy(nom.data(), e->d_nominal.p, nom.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(dyn.data(), e->d_dyn.p, dyn.size() * sizeof(InstDyn), hipMemcpyDeviceToHost));
    const dr_launch_cfg saved = e->launch;
    dr_launch_cfg best_cfg = saved;
    float best_t = 1e30f;
    std::vector<dr_launch_cfg> cands;
    for (int rb : {64, 128, 256})
        for (int ch : {512, 1024, 2048, 4096})
            for (int tg : {16, 32})
            {
                dr_launch_cfg c = saved;
                c.rollout_block = rb;
                c.rollout_grid = 0;
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
        rc = dr_sr_set_launch(e, &c);
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
        dr_sr_set_launch(e, &saved);
        e->drop_graphs();
        return rc;
    }
    best_cfg.steps_per_graph = saved.steps_per_graph;
    if ((rc = dr_sr_set_launch(e, &best_cfg)))
        return rc;
    e->drop_graphs();
    HIPCHK(hipMemcpy(e->d_nominal.p, nom.data(), nom.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_dyn.p, dyn.data(), dyn.size() * sizeof(InstDyn), hipMemcpyHostToDevice));
    if (best)
        *best = best_cfg;
    if (best_ms)
        *best_ms = best_t;
    return DR_OK;
}

int dr_sr_get_launch(dr_sr_engine* e, dr_launch_cfg* cfg)
{
    if (!e || !cfg)
        return host_fail(DR_EINVAL, "NULL argument");
    *cfg = e->launch;
    return DR_OK;
}

int dr_sr_set_timing(dr_sr_engine* e, int on)
{
    if (!e)
        return host_fail(DR_EINVAL, "NULL handle");
    e->timing = on != 0;
    return DR_OK;
}

int dr_sr_last_timing(dr_sr_engine* e, float* ms, int n)
{
    if (!e || !ms || n != 4)
        return host_fail(DR_EINVAL, "last_timing: expected 4 floats");
    std::memcpy(ms, e->last_ms, sizeof(e->last_ms));
    return DR_OK;
}

} // extern "C"
