// Host side of the branching spline MPPI engine on StratRace2D, HiddenObs and DroneRace: the natural
// cubic spline basis, descriptor tables, device buffers, the per-step
// launch sequence, hipGraph capture and the dr_bm_* C ABI.
//
// Reference host orchestration this replaces:
//   MPPIController::buildSplineMatrix      src/mppi_controller.cu:67-141
//   MPPIController::getControl (no verification) src/mppi_controller.cu:250-466
//   SimulationEngine::dynStep              src/engine.cpp:286-312
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "dr_bmkernels.h"
#include "dr_cp.h"
#include "dr_host.h"

namespace dr {

hipError_t launch_bm_rollout(int env_kind, int max_m, int max_tm, int grid, hipStream_t st, const BmInstDev* insts, const BmDyn* dyn,
                             const BmTile* tiles, int n_tiles, const void* env, const dr_sr_cost& KC, const void* states,
                             const float* knots, const float* basis, float2* noise, float* ctrue, int* used, int* btime, int* rlen);
hipError_t launch_bm_reduce(int n_inst, int max_nbk, int n_rtiles, hipStream_t st, const BmInstDev* insts, BmDyn* dyn, const BmRTile* rt,
                            const float* ctrue, float* costs, const int* used, const int* btime, float* pmin, float* mins,
                            const float2* noise, double* part, hipEvent_t mid);
hipError_t launch_bm_update(int n_inst, hipStream_t st, const BmInstDev* insts, BmDyn* dyn, const double* part, const float* basis,
                            float* knots, float* dnew, float* nu, int* fail, float max_accel);
hipError_t launch_bm_verify(int env_kind, int n_tiles, hipStream_t st, const BmInstDev* insts, const BmDyn* dyn, const BmVTile* tiles,
                            const void* env, const void* states, const float* dnew, const float* dense, int* fail);
hipError_t launch_bm_certify(int n_inst, hipStream_t st, const BmInstDev* insts, BmDyn* dyn, const int* verdict, const float* dnew,
                             float* dense, float* actions);
hipError_t launch_bm_advance(int env_kind, int n_inst, hipStream_t st, const BmInstDev* insts, BmDyn* dyn, const void* env, void* states,
                             const float* actions, int* events);

// Natural cubic spline basis B[t][i] = spline through e_i at the knots,
// evaluated at t = 0..T-1 with endpoint hold (mppi_controller.cu:67-141, the
// same FP32 operations: tridiagonal system by the Thomas algorithm, then the
// piecewise cubic).
std::vector<float> spline_basis(const int* tau, int M, int T)
{
    std::vector<float> B((size_t)T * M, 0.0f);
    std::vector<float> h(M - 1);
    for (int j = 0; j < M - 1; ++j)
        h[j] = (float)(tau[j + 1] - tau[j]);
    for (int basis = 0; basis < M; ++basis)
    {
        std::vector<float> y(M, 0.0f);
        y[basis] = 1.0f;
        std::vector<float> m(M, 0.0f);
        if (M > 2)
        {
            const int N = M - 2;
            std::vector<float> a(N), b(N), c(N), d(N);
            for (int i = 0; i < N; ++i)
            {
                const int j = i + 1;
                a[i] = h[j - 1];
                b[i] = 2.0f * (h[j - 1] + h[j]);
                c[i] = h[j];
                d[i] = 6.0f * ((y[j + 1] - y[j]) / h[j] - (y[j] - y[j - 1]) / h[j - 1]);
            }
            for (int i = 1; i < N; ++i)
            {
                const float w = a[i] / b[i - 1];
                b[i] = b[i] - w * c[i - 1];
                d[i] = d[i] - w * d[i - 1];
            }
            std::vector<float> x(N);
            x[N - 1] = d[N - 1] / b[N - 1];
            for (int i = N - 2; i >= 0; --i)
                x[i] = (d[i] - c[i] * x[i + 1]) / b[i];
            for (int i = 0; i < N; ++i)
                m[i + 1] = x[i];
        }
        int seg = 0;
        for (int t = 0; t < T; ++t)
        {
            float val;
            if (t <= tau[0])
                val = y[0];
            else if (t >= tau[M - 1])
                val = y[M - 1];
            else
            {
                while (!(tau[seg] <= t && t <= tau[seg + 1]))
                    ++seg;
                const float x0 = (float)tau[seg], x1 = (float)tau[seg + 1];
                const float hh = x1 - x0;
                const float A = (x1 - (float)t) / hh;
                const float BB = ((float)t - x0) / hh;
                val = m[seg] * (A * A * A - A) * (hh * hh) / 6.0f + m[seg + 1] * (BB * BB * BB - BB) * (hh * hh) / 6.0f + y[seg] * A +
                      y[seg + 1] * BB;
            }
            B[(size_t)t * M + basis] = val;
        }
    }
    return B;
}

} // namespace dr

using namespace dr;

// Host side of the plan verification of one instance: the Clopper-Pearson
// quantiles (memoised per failure count; n and beta are fixed) and the
// outcome of the last step (computeCertifiedLoss / computeEpsilon,
// mppi_controller.cu:468-514).
struct BmVerif
{
    std::unordered_map<uint32_t, double> lo_half, up_half, up_full;
    double certified = -1.0, epsilon = 0.0, eps_partial = 0.0;
    uint32_t fail_total = 0;
    int use_new = 1;
};

constexpr int BM_NEV = 6; // timing events per step

struct dr_bm_engine
{
    int device = 0;
    hipStream_t stream = nullptr;
    int env_kind = BM_ENV_SR;
    int n_models = 0;
    float max_accel = 0.0f;
    size_t state_size = 0; // bytes of one instance's world state
    dr_sr_cost cost{};
    std::vector<dr_bm_instance> inst;
    std::vector<BmInstDev> idev;
    std::vector<BmTile> tiles;
    std::vector<BmRTile> rtiles;
    std::vector<BmVTile> vtiles;
    std::vector<BmVerif> verif;
    bool any_verif = false;
    int max_m = 0, max_tm = 0, max_nbk = 0;
    long long noise_total = 0, part_total = 0;
    int ct_total = 0, cost_total = 0, min_total = 0, pmin_total = 0, knot_total = 0, dense_total = 0, B_total = 0, nu_total = 0,
        fail_total = 0;

    DevBuf<BmInstDev> d_inst;
    DevBuf<BmDyn> d_dyn;
    DevBuf<BmTile> d_tiles;
    DevBuf<BmRTile> d_rt;
    DevBuf<BmVTile> d_vt;
    DevBuf<uint8_t> d_env;    // dr_sr_env or dr_ho_env
    DevBuf<uint8_t> d_states; // dr_sr_state or dr_ho_state per instance
    DevBuf<float> d_knots, d_dense, d_dnew, d_B, d_ctrue, d_costs, d_mins, d_pmin, d_nu, d_actions;
    DevBuf<int> d_used, d_btime, d_rlen, d_events, d_fail, d_verdict;
    DevBuf<float2> d_noise;
    DevBuf<double> d_part;

    // [advance][part]: part 0 = sample .. verification (or the whole step
    // without verification), part 1 = plan choice + world step
    hipGraphExec_t graph_exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    bool timing = false;
    bool record_lengths = false; // rollout lengths into d_rlen (a separate kernel instantiation)
    int rollout_grid = 0;        // rollout workgroups (0: one per tile; else a persistent grid over the tiles)
    std::vector<hipEvent_t> events;
    int timed_steps = 0;
    float last_ms[BM_NEV] = {0, 0, 0, 0, 0, 0};
    std::vector<int> h_fail, h_verdict;

    ~dr_bm_engine()
    {
        for (auto& row : graph_exec)
            for (auto& g : row)
                if (g)
                    (void)hipGraphExecDestroy(g);
        for (auto ev : events)
            (void)hipEventDestroy(ev);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
    int enqueue_a(hipEvent_t* ev);
    int enqueue_b(bool advance, hipEvent_t* ev);
    int certify_host();
    int graph(int gi, int part, bool advance);
    int run(int n_steps, bool advance);
};

// sample, rollout, reduce, update, verification
int dr_bm_engine::enqueue_a(hipEvent_t* ev)
{
    const int n = (int)inst.size();
    if (ev)
        HIPCHK(hipEventRecord(ev[0], stream));
    HIPCHK(launch_bm_rollout(env_kind, max_m, max_tm, rollout_grid, stream, d_inst.p, d_dyn.p, d_tiles.p, (int)tiles.size(), d_env.p, cost, d_states.p,
                             d_knots.p, d_B.p, d_noise.p, d_ctrue.p, d_used.p, d_btime.p, record_lengths ? d_rlen.p : nullptr));
    if (ev)
        HIPCHK(hipEventRecord(ev[1], stream));
    HIPCHK(launch_bm_reduce(n, max_nbk, (int)rtiles.size(), stream, d_inst.p, d_dyn.p, d_rt.p, d_ctrue.p, d_costs.p, d_used.p, d_btime.p,
                            d_pmin.p, d_mins.p, d_noise.p, d_part.p, ev ? ev[2] : nullptr));
    if (ev)
        HIPCHK(hipEventRecord(ev[3], stream));
    HIPCHK(launch_bm_update(n, stream, d_inst.p, d_dyn.p, d_part.p, d_B.p, d_knots.p, d_dnew.p, d_nu.p, d_fail.p, max_accel));
    HIPCHK(launch_bm_verify(env_kind, (int)vtiles.size(), stream, d_inst.p, d_dyn.p, d_vt.p, d_env.p, d_states.p, d_dnew.p, d_dense.p, d_fail.p));
    if (ev)
        HIPCHK(hipEventRecord(ev[4], stream));
    return DR_OK;
}

// plan choice, action, shift, world step
int dr_bm_engine::enqueue_b(bool advance, hipEvent_t* ev)
{
    const int n = (int)inst.size();
    HIPCHK(launch_bm_certify(n, stream, d_inst.p, d_dyn.p, d_verdict.p, d_dnew.p, d_dense.p, d_actions.p));
    if (advance)
        HIPCHK(launch_bm_advance(env_kind, n, stream, d_inst.p, d_dyn.p, d_env.p, d_states.p, d_actions.p, d_events.p));
    if (ev)
        HIPCHK(hipEventRecord(ev[5], stream));
    return DR_OK;
}

// computeCertifiedLoss + computeEpsilon (mppi_controller.cu:468-514) from
// the failure counts of this step; writes the verdicts to the device.
int dr_bm_engine::certify_host()
{
    HIPCHK(hipMemcpyAsync(h_fail.data(), d_fail.p, h_fail.size() * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (size_t i = 0; i < inst.size(); ++i)
    {
        const BmInstDev& I = idev[i];
        BmVerif& V = verif[i];
        if (I.n_verif <= 0)
        {
            h_verdict[i] = 1;
            continue;
        }
        const dr_bm_instance& S = inst[i];
        const unsigned n = (unsigned)I.n_verif;
        const double half = (double)S.beta / 2.0;
        auto memo = [](std::unordered_map<uint32_t, double>& m, uint32_t k, double v_fn(unsigned, unsigned, double), unsigned nn,
                       double a) {
            auto it = m.find(k);
            if (it != m.end())
                return it->second;
            const double v = v_fn(k, nn, a);
            m.emplace(k, v);
            return v;
        };
        const int* fnew = h_fail.data() + I.fail_off;
        const int* fold = fnew + I.T;
        double certified = -1.0;
        uint32_t totNew = 0, totOld = 0;
        for (int t = 0; t < I.T; ++t)
        {
            totNew += (uint32_t)fnew[t];
            totOld += (uint32_t)fold[t];
            const double e1 = memo(V.up_half, totNew, cp::cp_upper, n, half);
            const double e2 = memo(V.lo_half, totOld, cp::cp_lower, n, half);
            if (e1 - e2 > certified)
                certified = e1 - e2;
        }
        V.certified = certified;
        V.use_new = certified < (double)S.max_verif_eps ? 1 : 0;
        V.fail_total = V.use_new ? totNew : totOld;
        V.eps_partial = memo(V.up_full, V.fail_total, cp::cp_upper, n, (double)S.beta);
        V.epsilon = V.eps_partial + (double)S.verif_horizon * (double)S.max_verif_eps;
        h_verdict[i] = V.use_new;
    }
    HIPCHK(hipMemcpyAsync(d_verdict.p, h_verdict.data(), h_verdict.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    return DR_OK;
}

int dr_bm_engine::graph(int gi, int part, bool advance)
{
    if (!graph_exec[gi][part])
    {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = DR_OK;
        if (part == 0)
        {
            rc = enqueue_a(nullptr);
            if (!rc && !any_verif)
                rc = enqueue_b(advance, nullptr);
        }
        else
            rc = enqueue_b(advance, nullptr);
        hipError_t ce = hipStreamEndCapture(stream, &g);
        if (rc)
        {
            if (g)
                (void)hipGraphDestroy(g);
            return rc;
        }
        if (ce != hipSuccess)
            return host_fail(DR_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        hipError_t ie = hipGraphInstantiate(&graph_exec[gi][part], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess)
            return host_fail(DR_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    HIPCHK(hipGraphLaunch(graph_exec[gi][part], stream));
    return DR_OK;
}

// Without verification one control step is one graph replay; with it, the
// step pauses after the verification for the host's Clopper-Pearson bounds
// (as the reference does) and resumes with the plan choice.
int dr_bm_engine::run(int n_steps, bool advance)
{
    if (n_steps < 0)
        return host_fail(DR_EINVAL, "n_steps < 0");
    timed_steps = 0;
    const int gi = advance ? 1 : 0;
    if (timing)
        while (events.size() < (size_t)n_steps * BM_NEV)
        {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            events.push_back(ev);
        }
    for (int s = 0; s < n_steps; ++s)
    {
        hipEvent_t* ev = timing ? &events[(size_t)s * BM_NEV] : nullptr;
        if (timing)
        {
            if (int rc = enqueue_a(ev))
                return rc;
        }
        else if (int rc = graph(gi, 0, advance))
            return rc;
        if (any_verif)
            if (int rc = certify_host())
                return rc;
        if (timing)
        {
            if (int rc = enqueue_b(advance, ev))
                return rc;
        }
        else if (any_verif)
            if (int rc = graph(gi, 1, advance))
                return rc;
    }
    if (timing)
        timed_steps = n_steps;
    return DR_OK;
}

namespace dr {

// HiddenObs configuration: finite geometry, positive step, radius and
// limits, nonnegative noise levels, consecutive gates apart (getAdvance
// divides by their distance).
int check_ho_env(const dr_ho_env& env)
{
    const float* f = &env.arena_min[0];
    const float* end = reinterpret_cast<const float*>(&env.n_win_laps);
    for (; f < end; ++f)
        if (!std::isfinite(*f))
            return host_fail(DR_EINVAL, "HiddenObs configuration values must be finite");
    for (f = &env.gate_centers[0][0]; f < &env.annulus_radius[0] + 2; ++f)
        if (!std::isfinite(*f))
            return host_fail(DR_EINVAL, "HiddenObs geometry must be finite");
    if (!(env.dt > 0.0f) || !(env.drone_radius > 0.0f) || !(env.max_speed > 0.0f) || !(env.max_accel > 0.0f))
        return host_fail(DR_EINVAL, "dt, drone_radius, max_speed and max_accel must be > 0");
    if (env.pos_noise < 0.0f || env.speed_noise < 0.0f || env.action_noise < 0.0f)
        return host_fail(DR_EINVAL, "noise levels must be >= 0");
    if (env.n_win_laps < 1)
        return host_fail(DR_EINVAL, "n_win_laps must be >= 1");
    for (int g = 0; g < 3; ++g)
    {
        const float* a = env.gate_centers[g];
        const float* b = env.gate_centers[(g + 1) % 3];
        if (a[0] == b[0] && a[1] == b[1])
            return host_fail(DR_EINVAL, "consecutive gate centres must differ");
    }
    return DR_OK;
}

// DroneRace configuration: counts within the compiled maxima, finite
// values, positive step, radius and limits, nonnegative noise levels,
// racelines of the opponent models that exist, consecutive gates apart.
int check_drace_env(const dr_drace_env& e)
{
    if (e.n_gates < 2 || e.n_gates > DR_DRACE_MAX_GATES || e.n_obstacles < 0 || e.n_obstacles > DR_DRACE_MAX_OBSTACLES ||
        e.n_round_obstacles < 0 || e.n_round_obstacles > DR_DRACE_MAX_OBSTACLES || e.n_racelines < 1 || e.n_racelines > DR_DRACE_MAX_RACELINES ||
        e.n_models < 1 || e.n_models > DR_SR_MAX_MODELS || (e.i_mppi != 0 && e.i_mppi != 1))
        return host_fail(DR_EINVAL, "DroneRace counts out of range (gates 2..8, obstacles 0..4, racelines 1..4, models 1..4, i_mppi 0/1)");
    const float* f = &e.arena_min[0];
    for (; f < reinterpret_cast<const float*>(&e.n_win_laps); ++f)
        if (!std::isfinite(*f))
            return host_fail(DR_EINVAL, "DroneRace configuration values must be finite");
    if (!(e.dt > 0.0f) || !(e.drone_radius > 0.0f) || e.pos_noise < 0.0f || e.speed_noise < 0.0f || e.action_noise < 0.0f)
        return host_fail(DR_EINVAL, "dt and drone_radius must be > 0, noise levels >= 0");
    for (int a = 0; a < 2; ++a)
        if (!(e.max_speed[a] > 0.0f) || !(e.max_accel[a] > 0.0f))
            return host_fail(DR_EINVAL, "max_speed and max_accel must be > 0");
    if (e.n_win_laps < 1 || !std::isfinite(e.target_distance))
        return host_fail(DR_EINVAL, "n_win_laps must be >= 1 and target_distance finite");
    for (int g = 0; g < e.n_gates; ++g)
    {
        const float* a = e.gate_centers[g];
        const float* b = e.gate_centers[(g + 1) % e.n_gates];
        if (!std::isfinite(a[0]) || !std::isfinite(a[1]) || !std::isfinite(e.gate_vectors[g][0]) || !std::isfinite(e.gate_vectors[g][1]) ||
            !std::isfinite(e.gate_radius[g]))
            return host_fail(DR_EINVAL, "gates must be finite");
        if (a[0] == b[0] && a[1] == b[1])
            return host_fail(DR_EINVAL, "consecutive gate centres must differ");
    }
    for (int o = 0; o < e.n_obstacles; ++o)
        for (int j = 0; j < 4; ++j)
            if (!std::isfinite(e.obstacles[o][j]))
                return host_fail(DR_EINVAL, "obstacles must be finite");
    for (int o = 0; o < e.n_round_obstacles; ++o)
        if (!std::isfinite(e.round_centers[o][0]) || !std::isfinite(e.round_centers[o][1]) || !std::isfinite(e.round_radius[o]))
            return host_fail(DR_EINVAL, "round obstacles must be finite");
    for (int m = 0; m < e.n_models; ++m)
    {
        const dr_drace_pid& p = e.opp_pid[m];
        if (p.raceline_index < 0 || p.raceline_index >= e.n_racelines)
            return host_fail(DR_EINVAL, "opponent model raceline_index must name one of the racelines");
        if (!std::isfinite(p.kp) || !std::isfinite(p.kd) || !std::isfinite(p.repulsion_factor) || !std::isfinite(p.repulsion_power) ||
            !std::isfinite(p.repulsion_dist_fact) || !(p.action_noise >= 0.0f) || !std::isfinite(p.action_noise))
            return host_fail(DR_EINVAL, "opponent PD configurations must be finite, action_noise >= 0");
    }
    for (int r = 0; r < e.n_racelines; ++r)
        for (int i = 0; i < DR_DRACE_SAMPLES; ++i)
            if (!std::isfinite(e.track_points[r][i][0]) || !std::isfinite(e.track_points[r][i][1]))
                return host_fail(DR_EINVAL, "track points must be finite");
    return DR_OK;
}

} // namespace dr

namespace {

int check_bm(dr_bm_engine* e, int i)
{
    if (!e)
        return host_fail(DR_EINVAL, "NULL handle");
    if (i < 0 || i >= (int)e->inst.size())
        return host_fail(DR_EINVAL, "instance index out of range");
    return DR_OK;
}

template <class T>
int bm_out(dr_bm_engine* e, T* dst, const T* src, size_t n)
{
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

template <class T>
int bm_in(dr_bm_engine* e, T* dst, const T* src, size_t n)
{
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DR_OK;
}

int get_dyn(dr_bm_engine* e, std::vector<BmDyn>& dyn)
{
    dyn.resize(e->inst.size());
    return bm_out(e, dyn.data(), e->d_dyn.p, dyn.size());
}

} // namespace

extern "C" {

// Shared by dr_bm_create and dr_bmho_create: env points at a validated
// environment configuration of env_size bytes with n_models opponent models.
static int bm_create(int env_kind, const void* env, size_t env_size, size_t state_size, int n_models, float max_accel, const dr_sr_cost* cost,
                     const dr_bm_instance* instances, int n_instances, dr_bm_engine** out)
{
    for (int i = 0; i < n_instances; ++i)
    {
        const dr_bm_instance& s = instances[i];
        const std::string w = "instance " + std::to_string(i) + ": ";
        if (s.n_samples < 1 || s.n_samples > (1 << 22))
            return host_fail(DR_EINVAL, w + "n_samples must be in [1, 2^22]");
        if (s.n_timesteps < 2 || s.n_timesteps > DR_MAX_T)
            return host_fail(DR_EINVAL, w + "n_timesteps must be in [2, 256]");
        if (s.n_knots < 2 || s.n_knots > DR_MAX_KNOTS || s.n_knots * s.n_timesteps > BM_MAX_TM)
            return host_fail(DR_EINVAL, w + "n_knots must be in [2, 60] with n_knots * n_timesteps <= 8192");
        if (s.knots[0] < 0 || s.knots[s.n_knots - 1] > s.n_timesteps - 1)
            return host_fail(DR_EINVAL, w + "knots must lie in [0, T - 1]");
        for (int m = 1; m < s.n_knots; ++m)
            if (s.knots[m] <= s.knots[m - 1])
                return host_fail(DR_EINVAL, w + "knots must be strictly increasing");
        if (!(s.lambda > 0.0f) || !(s.sigma >= 0.0f) || !std::isfinite(s.sigma) || !std::isfinite(s.min_confidence))
            return host_fail(DR_EINVAL, w + "lambda must be > 0, sigma finite and >= 0, min_confidence finite");
        if (s.true_theta < 0 || s.true_theta >= n_models)
            return host_fail(DR_EINVAL, w + "true_theta out of range");
        for (int th = 0; th < n_models; ++th)
            if (!(s.init_belief[th] >= 0.0f) || !std::isfinite(s.init_belief[th]))
                return host_fail(DR_EINVAL, w + "belief entries must be finite and >= 0");
        if (s.n_verif_samples < 0 || s.n_verif_samples > (1 << 22))
            return host_fail(DR_EINVAL, w + "n_verif_samples must be in [0, 2^22]");
        if (s.n_verif_samples > 0 &&
            (s.verif_horizon < 1 || s.verif_horizon > s.n_timesteps || !(s.beta > 0.0f && s.beta < 1.0f) || !std::isfinite(s.max_verif_eps)))
            return host_fail(DR_EINVAL, w + "verification needs verif_horizon in [1, T], beta in (0, 1) and a finite max_verif_eps");
    }
    dr_bm_engine* e = new (std::nothrow) dr_bm_engine();
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
    e->env_kind = env_kind;
    e->n_models = n_models;
    e->max_accel = max_accel;
    e->state_size = state_size;
    e->cost = *cost;
    e->inst.assign(instances, instances + n_instances);
    const int n = n_instances, NT = n_models, NB = NT + 1;
    e->idev.assign(n, BmInstDev{});
    std::vector<float> Ball;
    for (int i = 0; i < n; ++i)
    {
        const dr_bm_instance& s = instances[i];
        BmInstDev& I = e->idev[i];
        I.K = s.n_samples;
        I.T = s.n_timesteps;
        I.M = s.n_knots;
        I.NB = NB;
        I.NT = NT;
        for (int m = 0; m < DR_MAX_KNOTS; ++m)
            I.knots[m] = m < s.n_knots ? s.knots[m] : 0;
        I.key0 = (uint32_t)(s.seed & 0xffffffffu);
        I.key1 = (uint32_t)(s.seed >> 32);
        I.sigma = s.sigma;
        I.min_conf = s.min_confidence;
        I.true_theta = s.true_theta;
        I.adapt = s.adapt_lambda;
        I.noise_off = e->noise_total;
        e->noise_total += (long long)NB * I.M * I.K;
        I.ct_off = e->ct_total;
        e->ct_total += NT * I.K;
        I.cost_off = e->cost_total;
        e->cost_total += NB * I.K;
        I.min_off = e->min_total;
        e->min_total += NB * I.M;
        I.knot_off = e->knot_total;
        e->knot_total += NB * I.M * 2;
        I.dense_off = e->dense_total;
        e->dense_total += NB * I.T * 2;
        I.B_off = e->B_total;
        e->B_total += I.T * I.M;
        const std::vector<float> B = spline_basis(s.knots, I.M, I.T);
        Ball.insert(Ball.end(), B.begin(), B.end());
        I.n_chunks = (I.K + BM_CHUNK - 1) / BM_CHUNK;
        I.part_off = e->part_total;
        e->part_total += (long long)NB * I.M * I.n_chunks * 3;
        I.pmin_off = e->pmin_total;
        e->pmin_total += NB * I.M * I.n_chunks;
        I.nu_off = e->nu_total;
        e->nu_total += NB;
        I.n_verif = s.n_verif_samples;
        I.verif_h = s.n_verif_samples > 0 ? s.verif_horizon : 0;
        I.fail_off = e->fail_total;
        e->fail_total += 2 * I.T;
        if (I.n_verif > 0)
        {
            e->any_verif = true;
            for (int set = 0; set < 2; ++set)
                for (int s0 = 0; s0 < I.n_verif; s0 += BM_VBLOCK)
                    e->vtiles.push_back(BmVTile{i, set, s0, 0});
        }
        e->max_m = std::max(e->max_m, I.M);
        e->max_tm = std::max(e->max_tm, I.T * I.M);
        e->max_nbk = std::max(e->max_nbk, NB * I.K);
        for (int k0 = 0; k0 < I.K; k0 += BM_BLOCK)
            for (int th = 0; th < NT; ++th)
                e->tiles.push_back(BmTile{i, k0, th, 0});
        for (int b = 0; b < NB; ++b)
            for (int m = 0; m < I.M; ++m)
                for (int c = 0; c < I.n_chunks; ++c)
                    e->rtiles.push_back(BmRTile{i, b, m, c});
    }
    for (const float v : Ball)
        if (!std::isfinite(v))
            return bail(host_fail(DR_EINVAL, "spline basis not finite (check the knots)"));
    auto chk = [&](hipError_t x, const char* what) {
        if (x != hipSuccess)
        {
            host_fail(x == hipErrorOutOfMemory ? DR_ENOMEM : DR_EHIP, std::string(what) + ": " + hipGetErrorString(x));
            return false;
        }
        return true;
    };
    if (!chk(e->d_inst.alloc(n), "alloc inst") || !chk(e->d_dyn.alloc(n), "alloc dyn") || !chk(e->d_tiles.alloc(e->tiles.size()), "alloc tiles") ||
        !chk(e->d_rt.alloc(e->rtiles.size()), "alloc rt") || !chk(e->d_pmin.alloc(e->pmin_total), "alloc pmin") ||
        !chk(e->d_vt.alloc(e->vtiles.size()), "alloc vtiles") || !chk(e->d_dnew.alloc(e->dense_total), "alloc dnew") ||
        !chk(e->d_fail.alloc(e->fail_total), "alloc fail") || !chk(e->d_verdict.alloc(n), "alloc verdict") ||
        !chk(e->d_env.alloc(env_size), "alloc env") || !chk(e->d_states.alloc(n * state_size), "alloc states") || !chk(e->d_knots.alloc(e->knot_total), "alloc knots") ||
        !chk(e->d_dense.alloc(e->dense_total), "alloc dense") || !chk(e->d_B.alloc(e->B_total), "alloc B") ||
        !chk(e->d_ctrue.alloc(e->ct_total), "alloc ctrue") || !chk(e->d_costs.alloc(e->cost_total), "alloc costs") ||
        !chk(e->d_mins.alloc(e->min_total), "alloc mins") || !chk(e->d_nu.alloc(e->nu_total), "alloc nu") ||
        !chk(e->d_actions.alloc((size_t)n * 2), "alloc actions") || !chk(e->d_used.alloc(e->ct_total), "alloc used") ||
        !chk(e->d_btime.alloc(e->ct_total), "alloc btime") || !chk(e->d_rlen.alloc(e->ct_total), "alloc rlen") || !chk(e->d_events.alloc(n), "alloc events") ||
        !chk(e->d_noise.alloc(e->noise_total), "alloc noise") || !chk(e->d_part.alloc(e->part_total), "alloc part"))
        return bail(DR_ENOMEM);
    std::vector<BmDyn> dyn(n);
    for (int i = 0; i < n; ++i)
    {
        std::memset(&dyn[i], 0, sizeof(BmDyn));
        dyn[i].lambda = instances[i].lambda;
        dyn[i].use_new = 1;
        for (int th = 0; th < NT; ++th)
            dyn[i].belief[th] = instances[i].init_belief[th];
    }
    std::vector<float> zk(e->knot_total, 0.0f), zd(e->dense_total, 0.0f);
    std::vector<uint8_t> st(n * state_size, 0);
    std::vector<int> ev(n, 0);
    if (!chk(hipMemcpy(e->d_inst.p, e->idev.data(), n * sizeof(BmInstDev), hipMemcpyHostToDevice), "copy inst") ||
        !chk(hipMemcpy(e->d_dyn.p, dyn.data(), n * sizeof(BmDyn), hipMemcpyHostToDevice), "copy dyn") ||
        !chk(hipMemcpy(e->d_tiles.p, e->tiles.data(), e->tiles.size() * sizeof(BmTile), hipMemcpyHostToDevice), "copy tiles") ||
        !chk(hipMemcpy(e->d_rt.p, e->rtiles.data(), e->rtiles.size() * sizeof(BmRTile), hipMemcpyHostToDevice), "copy rt") ||
        (!e->vtiles.empty() &&
         !chk(hipMemcpy(e->d_vt.p, e->vtiles.data(), e->vtiles.size() * sizeof(BmVTile), hipMemcpyHostToDevice), "copy vtiles")) ||
        !chk(hipMemcpy(e->d_fail.p, std::vector<int>(e->fail_total, 0).data(), e->fail_total * sizeof(int), hipMemcpyHostToDevice),
             "copy fail") ||
        !chk(hipMemcpy(e->d_verdict.p, std::vector<int>(n, 1).data(), n * sizeof(int), hipMemcpyHostToDevice), "copy verdict") ||
        !chk(hipMemcpy(e->d_env.p, env, env_size, hipMemcpyHostToDevice), "copy env") ||
        !chk(hipMemcpy(e->d_states.p, st.data(), st.size(), hipMemcpyHostToDevice), "copy states") ||
        !chk(hipMemcpy(e->d_knots.p, zk.data(), zk.size() * sizeof(float), hipMemcpyHostToDevice), "copy knots") ||
        !chk(hipMemcpy(e->d_dense.p, zd.data(), zd.size() * sizeof(float), hipMemcpyHostToDevice), "copy dense") ||
        !chk(hipMemcpy(e->d_dnew.p, zd.data(), zd.size() * sizeof(float), hipMemcpyHostToDevice), "copy dnew") ||
        !chk(hipMemcpy(e->d_B.p, Ball.data(), Ball.size() * sizeof(float), hipMemcpyHostToDevice), "copy B") ||
        !chk(hipMemcpy(e->d_events.p, ev.data(), n * sizeof(int), hipMemcpyHostToDevice), "copy events"))
        return bail(DR_EHIP);
    e->verif.assign(n, BmVerif{});
    e->h_fail.assign(e->fail_total, 0);
    e->h_verdict.assign(n, 1);
    *out = e;
    return DR_OK;
}

int dr_bm_create(const dr_sr_env* env, const dr_sr_cost* cost, const dr_bm_instance* instances, int n_instances, dr_bm_engine** out)
{
    if (!out)
        return host_fail(DR_EINVAL, "out is NULL");
    *out = nullptr;
    if (!env || !cost || !instances)
        return host_fail(DR_EINVAL, "NULL argument");
    if (int rc = check_sr_env(*env))
        return rc;
    return bm_create(BM_ENV_SR, env, sizeof(dr_sr_env), sizeof(dr_sr_state), env->n_models, env->max_accel[0], cost, instances, n_instances,
                     out);
}

int dr_bmho_create(const dr_ho_env* env, const dr_sr_cost* cost, const dr_bm_instance* instances, int n_instances, dr_bm_engine** out)
{
    if (!out)
        return host_fail(DR_EINVAL, "out is NULL");
    *out = nullptr;
    if (!env || !cost || !instances)
        return host_fail(DR_EINVAL, "NULL argument");
    if (int rc = check_ho_env(*env))
        return rc;
    return bm_create(BM_ENV_HO, env, sizeof(dr_ho_env), sizeof(dr_ho_state), 2, env->max_accel, cost, instances, n_instances, out);
}

void dr_bm_destroy(dr_bm_engine* e)
{
    if (e)
    {
        (void)hipStreamSynchronize(e->stream);
        delete e;
    }
}

static int bm_set_states(dr_bm_engine* e, int kind, const void* states, int n)
{
    if (!e || !states || n != (int)e->inst.size())
        return host_fail(DR_EINVAL, "set_states: bad handle, pointer or count");
    if (e->env_kind != kind)
        return host_fail(DR_EINVAL, "set_states: state type does not match the engine's environment");
    if (int rc = bm_in(e, e->d_states.p, static_cast<const uint8_t*>(states), (size_t)n * e->state_size))
        return rc;
    std::vector<int> ev(n, 0);
    return bm_in(e, e->d_events.p, ev.data(), ev.size());
}

static int bm_get_states(dr_bm_engine* e, int kind, void* states, int n)
{
    if (!e || !states || n != (int)e->inst.size())
        return host_fail(DR_EINVAL, "get_states: bad handle, pointer or count");
    if (e->env_kind != kind)
        return host_fail(DR_EINVAL, "get_states: state type does not match the engine's environment");
    return bm_out(e, static_cast<uint8_t*>(states), e->d_states.p, (size_t)n * e->state_size);
}

int dr_bmdrace_create(const dr_drace_env* env, const dr_sr_cost* cost, const dr_bm_instance* instances, int n_instances, dr_bm_engine** out)
{
    if (!out)
        return host_fail(DR_EINVAL, "out is NULL");
    *out = nullptr;
    if (!env || !cost || !instances)
        return host_fail(DR_EINVAL, "NULL argument");
    if (int rc = check_drace_env(*env))
        return rc;
    return bm_create(BM_ENV_DRACE, env, sizeof(dr_drace_env), sizeof(dr_drace_state), env->n_models, env->max_accel[env->i_mppi], cost,
                     instances, n_instances, out);
}

int dr_bm_set_states(dr_bm_engine* e, const dr_sr_state* states, int n) { return bm_set_states(e, BM_ENV_SR, states, n); }

int dr_bm_get_states(dr_bm_engine* e, dr_sr_state* states, int n) { return bm_get_states(e, BM_ENV_SR, states, n); }

int dr_bmho_set_states(dr_bm_engine* e, const dr_ho_state* states, int n) { return bm_set_states(e, BM_ENV_HO, states, n); }

int dr_bmho_get_states(dr_bm_engine* e, dr_ho_state* states, int n) { return bm_get_states(e, BM_ENV_HO, states, n); }

int dr_bmdrace_set_states(dr_bm_engine* e, const dr_drace_state* states, int n) { return bm_set_states(e, BM_ENV_DRACE, states, n); }

int dr_bmdrace_get_states(dr_bm_engine* e, dr_drace_state* states, int n) { return bm_get_states(e, BM_ENV_DRACE, states, n); }

int dr_bm_set_belief(dr_bm_engine* e, int instance, const float* belief, int n_models)
{
    if (int rc = check_bm(e, instance))
        return rc;
    if (!belief || n_models != e->n_models)
        return host_fail(DR_EINVAL, "set_belief: expected n_models values");
    for (int th = 0; th < n_models; ++th)
        if (!(belief[th] >= 0.0f) || !std::isfinite(belief[th]))
            return host_fail(DR_EINVAL, "set_belief: entries must be finite and >= 0");
    float b[DR_SR_MAX_MODELS] = {0};
    std::memcpy(b, belief, n_models * sizeof(float));
    return bm_in(e, e->d_dyn.p[instance].belief, b, (size_t)DR_SR_MAX_MODELS);
}

int dr_bm_get_belief(dr_bm_engine* e, int instance, float* belief, int n_models)
{
    if (int rc = check_bm(e, instance))
        return rc;
    if (!belief || n_models != e->n_models)
        return host_fail(DR_EINVAL, "get_belief: expected n_models values");
    std::vector<BmDyn> dyn;
    if (int rc = get_dyn(e, dyn))
        return rc;
    std::memcpy(belief, dyn[instance].belief, n_models * sizeof(float));
    return DR_OK;
}

int dr_bm_set_knots(dr_bm_engine* e, int instance, const float* knots, int n_floats)
{
    if (int rc = check_bm(e, instance))
        return rc;
    const BmInstDev& I = e->idev[instance];
    if (!knots || n_floats != I.NB * I.M * 2)
        return host_fail(DR_EINVAL, "set_knots: expected (n_models + 1) * M * 2 floats");
    return bm_in(e, e->d_knots.p + I.knot_off, knots, (size_t)n_floats);
}

int dr_bm_get_knots(dr_bm_engine* e, int instance, float* knots, int n_floats)
{
    if (int rc = check_bm(e, instance))
        return rc;
    const BmInstDev& I = e->idev[instance];
    if (!knots || n_floats != I.NB * I.M * 2)
        return host_fail(DR_EINVAL, "get_knots: expected (n_models + 1) * M * 2 floats");
    return bm_out(e, knots, e->d_knots.p + I.knot_off, (size_t)n_floats);
}

int dr_bm_get_nominal(dr_bm_engine* e, int instance, float* nominal, int n_floats)
{
    if (int rc = check_bm(e, instance))
        return rc;
    const BmInstDev& I = e->idev[instance];
    if (!nominal || n_floats != I.NB * I.T * 2)
        return host_fail(DR_EINVAL, "get_nominal: expected (n_models + 1) * T * 2 floats");
    return bm_out(e, nominal, e->d_dense.p + I.dense_off, (size_t)n_floats);
}

int dr_bm_get_basis(dr_bm_engine* e, int instance, float* B, int n_floats)
{
    if (int rc = check_bm(e, instance))
        return rc;
    const BmInstDev& I = e->idev[instance];
    if (!B || n_floats != I.T * I.M)
        return host_fail(DR_EINVAL, "get_basis: expected T * M floats");
    return bm_out(e, B, e->d_B.p + I.B_off, (size_t)n_floats);
}

int dr_bm_step(dr_bm_engine* e, int n_steps, int advance, float* actions)
{
    if (!e)
        return host_fail(DR_EINVAL, "NULL handle");
    if (int rc = e->run(n_steps, advance != 0))
        return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->timed_steps > 0)
    {
        float acc[BM_NEV] = {0, 0, 0, 0, 0, 0};
        for (int s = 0; s < e->timed_steps; ++s)
        {
            hipEvent_t* ev = &e->events[(size_t)s * BM_NEV];
            for (int q = 0; q < BM_NEV - 1; ++q)
            {
                float v;
                HIPCHK(hipEventElapsedTime(&v, ev[q], ev[q + 1]));
                acc[q] += v;
            }
            float t;
            HIPCHK(hipEventElapsedTime(&t, ev[0], ev[BM_NEV - 1]));
            acc[BM_NEV - 1] += t;
        }
        std::memcpy(e->last_ms, acc, sizeof(acc));
        e->timed_steps = 0;
    }
    if (actions)
        return bm_out(e, actions, e->d_actions.p, e->inst.size() * 2);
    return DR_OK;
}

int dr_bm_get_events(dr_bm_engine* e, int32_t* events, int n)
{
    if (!e || !events || n != (int)e->inst.size())
        return host_fail(DR_EINVAL, "get_events: bad handle, pointer or count");
    return bm_out(e, events, e->d_events.p, (size_t)n);
}

int dr_bm_get_noise(dr_bm_engine* e, int instance, float* noise, size_t n)
{
    if (int rc = check_bm(e, instance))
        return rc;
    const BmInstDev& I = e->idev[instance];
    if (!noise || n != (size_t)I.NB * I.M * I.K * 2)
        return host_fail(DR_EINVAL, "get_noise: expected (n_models + 1) * M * K * 2 floats");
    return bm_out(e, reinterpret_cast<float2*>(noise), e->d_noise.p + I.noise_off, (size_t)I.NB * I.M * I.K);
}

int dr_bm_get_costs_true(dr_bm_engine* e, int instance, float* costs, int n)
{
    if (int rc = check_bm(e, instance))
        return rc;
    const BmInstDev& I = e->idev[instance];
    if (!costs || n != I.NT * I.K)
        return host_fail(DR_EINVAL, "get_costs_true: expected n_models * K floats");
    return bm_out(e, costs, e->d_ctrue.p + I.ct_off, (size_t)n);
}

int dr_bm_get_branches(dr_bm_engine* e, int instance, int32_t* used, int32_t* time, int n)
{
    if (int rc = check_bm(e, instance))
        return rc;
    const BmInstDev& I = e->idev[instance];
    if (!used || !time || n != I.NT * I.K)
        return host_fail(DR_EINVAL, "get_branches: expected n_models * K values each");
    if (int rc = bm_out(e, used, e->d_used.p + I.ct_off, (size_t)n))
        return rc;
    return bm_out(e, time, e->d_btime.p + I.ct_off, (size_t)n);
}

int dr_bm_get_rollout_lengths(dr_bm_engine* e, int instance, int32_t* lengths, int n)
{
    if (int rc = check_bm(e, instance))
        return rc;
    if (!e->record_lengths)
        return host_fail(DR_EINVAL, "get_rollout_lengths: recording is off (dr_bm_set_record_lengths)");
    const BmInstDev& I = e->idev[instance];
    if (!lengths || n != I.NT * I.K)
        return host_fail(DR_EINVAL, "get_rollout_lengths: expected n_models * K values");
    return bm_out(e, lengths, e->d_rlen.p + I.ct_off, (size_t)n);
}

int dr_bm_get_costs(dr_bm_engine* e, int instance, float* costs, int n)
{
    if (int rc = check_bm(e, instance))
        return rc;
    const BmInstDev& I = e->idev[instance];
    if (!costs || n != I.NB * I.K)
        return host_fail(DR_EINVAL, "get_costs: expected (n_models + 1) * K floats");
    return bm_out(e, costs, e->d_costs.p + I.cost_off, (size_t)n);
}

int dr_bm_get_min_costs(dr_bm_engine* e, int instance, float* mins, int n)
{
    if (int rc = check_bm(e, instance))
        return rc;
    const BmInstDev& I = e->idev[instance];
    if (!mins || n != I.NB * I.M)
        return host_fail(DR_EINVAL, "get_min_costs: expected (n_models + 1) * M floats");
    return bm_out(e, mins, e->d_mins.p + I.min_off, (size_t)n);
}

int dr_bm_get_nu(dr_bm_engine* e, int instance, float* nu, int n)
{
    if (int rc = check_bm(e, instance))
        return rc;
    const BmInstDev& I = e->idev[instance];
    if (!nu || n != I.NB)
        return host_fail(DR_EINVAL, "get_nu: expected n_models + 1 floats");
    return bm_out(e, nu, e->d_nu.p + I.nu_off, (size_t)n);
}

int dr_bm_get_diag(dr_bm_engine* e, dr_bm_diag* diag, int n)
{
    if (!e || !diag || n != (int)e->inst.size())
        return host_fail(DR_EINVAL, "get_diag: bad handle, pointer or count");
    std::vector<BmDyn> dyn;
    if (int rc = get_dyn(e, dyn))
        return rc;
    std::vector<float> nu(e->nu_total), mins(e->min_total);
    if (int rc = bm_out(e, nu.data(), e->d_nu.p, nu.size()))
        return rc;
    if (int rc = bm_out(e, mins.data(), e->d_mins.p, mins.size()))
        return rc;
    for (int i = 0; i < n; ++i)
    {
        const BmInstDev& I = e->idev[i];
        std::memset(&diag[i], 0, sizeof(dr_bm_diag));
        diag[i].lambda = dyn[i].lambda;
        diag[i].step = dyn[i].step;
        diag[i].branch = dyn[i].branch;
        diag[i].nu = nu[I.nu_off + dyn[i].branch];
        diag[i].min_cost = mins[I.min_off + dyn[i].branch * I.M];
        const BmVerif& V = e->verif[i];
        diag[i].use_new_plan = dyn[i].use_new;
        diag[i].fail_count = V.fail_total;
        diag[i].epsilon = V.epsilon;
        diag[i].epsilon_partial = V.eps_partial;
        diag[i].certified_loss = I.n_verif > 0 ? V.certified : 0.0;
    }
    return DR_OK;
}

int dr_bm_get_fail_counts(dr_bm_engine* e, int instance, int32_t* fail_new, int32_t* fail_old, int n)
{
    if (int rc = check_bm(e, instance))
        return rc;
    const BmInstDev& I = e->idev[instance];
    if (!fail_new || !fail_old || n != I.T)
        return host_fail(DR_EINVAL, "get_fail_counts: expected T values each");
    if (int rc = bm_out(e, fail_new, e->d_fail.p + I.fail_off, (size_t)n))
        return rc;
    return bm_out(e, fail_old, e->d_fail.p + I.fail_off + I.T, (size_t)n);
}

int dr_bm_set_timing(dr_bm_engine* e, int on)
{
    if (!e)
        return host_fail(DR_EINVAL, "NULL handle");
    e->timing = on != 0;
    return DR_OK;
}

int dr_bm_set_record_lengths(dr_bm_engine* e, int on)
{
    if (!e)
        return host_fail(DR_EINVAL, "NULL handle");
    if (e->record_lengths != (on != 0))
    {
        // the captured graphs hold the other rollout instantiation
        HIPCHK(hipStreamSynchronize(e->stream));
        for (auto& row : e->graph_exec)
            for (auto& g : row)
                if (g)
                {
                    (void)hipGraphExecDestroy(g);
                    g = nullptr;
                }
        if (on)
        {
            std::vector<int> z(e->d_rlen.n, 0);
            HIPCHK(hipMemcpy(e->d_rlen.p, z.data(), z.size() * sizeof(int), hipMemcpyHostToDevice));
        }
    }
    e->record_lengths = on != 0;
    return DR_OK;
}

static void bm_drop_graphs(dr_bm_engine* e)
{
    for (auto& row : e->graph_exec)
        for (auto& g : row)
            if (g)
            {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
}

int dr_bm_set_launch(dr_bm_engine* e, const dr_launch_cfg* cfg)
{
    if (!e || !cfg)
        return host_fail(DR_EINVAL, "NULL argument");
    if (cfg->rollout_grid < 0)
        return host_fail(DR_EINVAL, "rollout_grid must be >= 0");
    HIPCHK(hipStreamSynchronize(e->stream));
    if (cfg->rollout_grid != e->rollout_grid)
        bm_drop_graphs(e);
    e->rollout_grid = cfg->rollout_grid;
    return DR_OK;
}

int dr_bm_get_launch(dr_bm_engine* e, dr_launch_cfg* cfg)
{
    if (!e || !cfg)
        return host_fail(DR_EINVAL, "NULL argument");
    *cfg = dr_launch_cfg{};
    cfg->rollout_block = BM_BLOCK;
    cfg->rollout_grid = e->rollout_grid;
    cfg->reduce_chunk = BM_CHUNK;
    cfg->use_graph = 1;
    cfg->steps_per_graph = 1;
    return DR_OK;
}

// Times the rollout grid candidates (one workgroup per tile, or a persistent
// grid of 1, 2, 4 or 8 workgroups per CU) on the engine's own workload,
// trials steps each after one warm-up step, no world step, and keeps the
// fastest; the knots, plans and per-instance state the timing steps change
// are restored.
int dr_bm_autotune(dr_bm_engine* e, int trials, dr_launch_cfg* best, float* best_ms)
{
    if (!e || trials < 1)
        return host_fail(DR_EINVAL, "autotune: bad handle or trials");
    HIPCHK(hipStreamSynchronize(e->stream));
    std::vector<float> knots(e->knot_total), dense(e->dense_total);
    std::vector<BmDyn> dyn(e->inst.size());
    HIPCHK(hipMemcpy(knots.data(), e->d_knots.p, knots.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(dense.data(), e->d_dense.p, dense.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(dyn.data(), e->d_dyn.p, dyn.size() * sizeof(BmDyn), hipMemcpyDeviceToHost));
    int n_cu = 256;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, e->device) == hipSuccess)
            n_cu = prop.multiProcessorCount;
    }
    const int saved = e->rollout_grid;
    int best_g = saved;
    float best_t = 1e30f;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    int rc = DR_OK;
    for (int g : {0, n_cu, 2 * n_cu, 4 * n_cu, 8 * n_cu})
    {
        if (g >= (int)e->tiles.size())
            continue;
        dr_launch_cfg c{};
        c.rollout_grid = g;
        if ((rc = dr_bm_set_launch(e, &c)))
            break;
        if ((rc = e->run(1, false)))
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
            best_g = g;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    dr_launch_cfg c{};
    c.rollout_grid = rc ? saved : best_g;
    dr_bm_set_launch(e, &c);
    if (rc)
        return rc;
    HIPCHK(hipMemcpy(e->d_knots.p, knots.data(), knots.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_dense.p, dense.data(), dense.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_dyn.p, dyn.data(), dyn.size() * sizeof(BmDyn), hipMemcpyHostToDevice));
    if (best)
        dr_bm_get_launch(e, best);
    if (best_ms)
        *best_ms = best_t;
    return DR_OK;
}

int dr_bm_last_timing(dr_bm_engine* e, float* ms, int n)
{
    if (!e || !ms || n != BM_NEV)
        return host_fail(DR_EINVAL, "last_timing: expected 6 floats");
    std::memcpy(ms, e->last_ms, sizeof(e->last_ms));
    return DR_OK;
}

} // extern "C"
