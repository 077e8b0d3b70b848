// ORACLE — test infrastructure only (see oracle.h).  C entry points loaded
// with ctypes by tests/, the smoke check and bench.py's CPU-baseline leg.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "oracle.h"

using namespace orc;

thread_local OpTally orc::g_tally;

namespace {

template <class S>
void state_to_arr(const St<S>& s, double* a)
{
    for (int i = 0; i < 3; ++i)
        a[i] = (double)s.p[i];
    for (int i = 0; i < 4; ++i)
        a[3 + i] = (double)s.q[i];
    for (int i = 0; i < 3; ++i)
        a[7 + i] = (double)s.v[i];
    for (int i = 0; i < 3; ++i)
        a[10 + i] = (double)s.w[i];
}

template <class S>
St<S> arr_to_state(const double* a, int gate, int laps)
{
    St<S> s;
    for (int i = 0; i < 3; ++i)
        s.p[i] = (S)a[i];
    for (int i = 0; i < 4; ++i)
        s.q[i] = (S)a[3 + i];
    for (int i = 0; i < 3; ++i)
        s.v[i] = (S)a[7 + i];
    for (int i = 0; i < 3; ++i)
        s.w[i] = (S)a[10 + i];
    s.gate = gate;
    s.laps = laps;
    return s;
}

// One sample's rollout (src/kernels.cu:100-155): for t < T with early stop
// on a terminal state, control = clamp(nominal[t] + eps[t]), Euler step, gate
// update, running cost * decay (decay *= 0.97), then the terminal cost.  eps
// is NULL for the noise-free rollout of the nominal.
template <class S>
S rollout_one(const dr_quad_params& P, const Model<S>& M, const dr_cost_cfg& K, const dr_track& tr, const dr_drone_state& ego, const dr_opponent* opp, int n_opp,
              const float* nominal, int T, const float* eps, long long eps_stride)
{
    St<S> s = load_state<S>(ego);
    S acc = 0, decay = 1;
    const S margin = (S)K.gate_traversal_margin;
    const CostK<S> Q(P, K);
    for (int t = 0; t < T; ++t)
    {
        S uin[4], u[4];
        for (int d = 0; d < 4; ++d)
        {
            const float e = eps ? eps[(long long)t * eps_stride + d] : 0.0f;
            if (sizeof(S) == 4)
                uin[d] = (S)(nominal[t * 4 + d] + e);
            else
                uin[d] = (S)nominal[t * 4 + d] + (S)e;
        }
        clamp_u(M, uin, u);
        const S p0[3] = {s.p[0], s.p[1], s.p[2]};
        euler(M, s, u);
        gate_update(tr, margin, p0, s);
        int term = 0;
        const S tau = (S)(M.dt * (S)(t + 1));
        const S c = running_cost(M, Q, K, tr, s, opp, n_opp, tau, &term);
        acc = fm(c, decay, acc);
        decay = decay * (S)0.97f;
        if (term)
            break;
    }
    return acc + (-(S)K.final_adv_weight * progress(tr, s));
}

template <class S>
void rollout_all(const dr_quad_params& P, const dr_cost_cfg& K, const dr_track& tr, const dr_drone_state& ego, const dr_opponent* opp,
                 int n_opp, const float* nominal, int n_samples, int T, uint32_t key0, uint32_t key1, uint32_t step, const float sigma[4],
                 double* costs, float* noise_out, double* traj_cost, int n_threads)
{
    const Model<S> M(P);
    if (traj_cost)
        *traj_cost = (double)rollout_one<S>(P, M, K, tr, ego, opp, n_opp, nominal, T, nullptr, 0);
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : 1)
    {
        std::vector<float> eps((size_t)T * 4);
#pragma omp for schedule(static)
        for (int k = 0; k < n_samples; ++k)
        {
            for (int t = 0; t < T; ++t)
            {
                float z[4];
                normals4(key0, key1, (uint32_t)k, (uint32_t)t, step, 0u, z);
                for (int d = 0; d < 4; ++d)
                    eps[(size_t)t * 4 + d] = sigma[d] * z[d];
                if (noise_out)
                    for (int d = 0; d < 4; ++d)
                        noise_out[((size_t)k * T + t) * 4 + d] = eps[(size_t)t * 4 + d];
            }
            costs[k] = (double)rollout_one<S>(P, M, K, tr, ego, opp, n_opp, nominal, T, eps.data(), 4);
        }
    }
}

// Plain MPPI update (kernels.cu:307-408 with one branch and coefficient 1,
// clamp kernels.cu:434-456, action mppi_controller.cu:392-397, shift
// mppi_controller.cu:430-436).  FP32 mode: weights in FP32 with exp32, sums
// in FP64, increment rounded to FP32 and added in FP32.  FP64 mode: all FP64.
template <class S>
void update(const dr_quad_params& P, const double* costs, const float* noise, const float* nominal, int n_samples, int T, float lambda,
            const float fill[4], double* updated, double* action, double* shifted, double* min_cost, double* nu, double* weights)
{
    const Model<S> M(P);
    S mnc = (S)INFINITY;
    for (int k = 0; k < n_samples; ++k)
        mnc = mn(mnc, (S)costs[k]);
    std::vector<double> w(n_samples);
    double den = 0.0;
    for (int k = 0; k < n_samples; ++k)
    {
        if (sizeof(S) == 4)
            w[k] = (double)exp32(-((float)costs[k] - (float)mnc) / lambda);
        else
            w[k] = exp(-(costs[k] - (double)mnc) / (double)lambda);
        den += w[k];
        if (weights)
            weights[k] = w[k];
    }
    std::vector<S> u((size_t)T * 4);
    for (int t = 0; t < T; ++t)
        for (int d = 0; d < 4; ++d)
        {
            double num = 0.0;
            for (int k = 0; k < n_samples; ++k)
                num += w[k] * (double)noise[((size_t)k * T + t) * 4 + d];
            S v = (S)nominal[t * 4 + d];
            if (den > 1e-30)
            {
                if (sizeof(S) == 4)
                    v = (S)((float)v + (float)(num / den));
                else
                    v = v + (S)(num / den);
            }
            u[(size_t)t * 4 + d] = v;
        }
    for (int t = 0; t < T; ++t)
    {
        S o[4];
        clamp_u(M, &u[(size_t)t * 4], o);
        for (int d = 0; d < 4; ++d)
            u[(size_t)t * 4 + d] = o[d];
    }
    for (int i = 0; i < T * 4; ++i)
        updated[i] = (double)u[i];
    for (int d = 0; d < 4; ++d)
        action[d] = (double)u[d];
    for (int t = 0; t < T; ++t)
        for (int d = 0; d < 4; ++d)
            shifted[t * 4 + d] = (t + 1 < T) ? (double)u[(size_t)(t + 1) * 4 + d] : (double)fill[d];
    *min_cost = (double)mnc;
    *nu = sizeof(S) == 4 ? (double)(float)den : den;
}

} // namespace

extern "C" {

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    W4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
    const W4 o = philox(c, key[0], key[1]);
    memcpy(out, o.v, sizeof(o.v));
}

void or_normals4(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, float out[4])
{
    normals4(k0, k1, c0, c1, c2, c3, out);
}

float or_log32(float x) { return log32(x); }
float or_exp32(float x) { return exp32(x); }
void or_sincos_turns(uint32_t k, float* s, float* c) { sincos_turns(k, s, c); }

// One clamped Euler step + gate update on a 13-vector (p, q, v, w).
void or_quad_step(int f64, const dr_quad_params* P, const dr_track* tr, float margin, double* x, int* gate, int* laps, const double* u_in)
{
    if (f64)
    {
        const Model<double> M(*P);
        St<double> s = arr_to_state<double>(x, *gate, *laps);
        double u[4];
        clamp_u(M, u_in, u);
        const double p0[3] = {s.p[0], s.p[1], s.p[2]};
        euler(M, s, u);
        if (tr)
            gate_update(*tr, (double)margin, p0, s);
        state_to_arr(s, x);
        *gate = s.gate;
        *laps = s.laps;
    }
    else
    {
        const Model<float> M(*P);
        St<float> s = arr_to_state<float>(x, *gate, *laps);
        float ui[4] = {(float)u_in[0], (float)u_in[1], (float)u_in[2], (float)u_in[3]}, u[4];
        clamp_u(M, ui, u);
        const float p0[3] = {s.p[0], s.p[1], s.p[2]};
        euler(M, s, u);
        if (tr)
            gate_update(*tr, margin, p0, s);
        state_to_arr(s, x);
        *gate = s.gate;
        *laps = s.laps;
    }
}

double or_progress(int f64, const dr_track* tr, const double* x, int gate, int laps)
{
    if (f64)
        return progress(*tr, arr_to_state<double>(x, gate, laps));
    return (double)progress(*tr, arr_to_state<float>(x, gate, laps));
}

double or_running_cost(int f64, const dr_quad_params* P, const dr_cost_cfg* K, const dr_track* tr, const double* x, int gate, int laps,
                       const dr_opponent* opp, int n_opp, double tau, int* term)
{
    if (f64)
        return running_cost(Model<double>(*P), CostK<double>(*P, *K), *K, *tr, arr_to_state<double>(x, gate, laps), opp, n_opp, tau,
                            term);
    return (double)running_cost(Model<float>(*P), CostK<float>(*P, *K), *K, *tr, arr_to_state<float>(x, gate, laps), opp, n_opp,
                                (float)tau, term);
}

// K sampled rollouts of one controller.  costs[K] (FP32 values widened in
// FP32 mode); noise_out [k][t][4] (may be NULL); traj_cost = noise-free
// rollout of the nominal (may be NULL).
int or_rollout(int f64, const dr_quad_params* P, const dr_cost_cfg* K, const dr_track* tr, const dr_drone_state* ego, const dr_opponent* opp,
               int n_opp, const float* nominal, int n_samples, int T, uint32_t key0, uint32_t key1, uint32_t step, const float* sigma,
               double* costs, float* noise_out, double* traj_cost, int n_threads)
{
    if (!P || !K || !tr || !ego || !nominal || !costs || !sigma || n_samples < 1 || T < 1 || T > DR_MAX_T || n_opp < 0 ||
        (n_opp > 0 && !opp))
        return 1;
    if (f64)
        rollout_all<double>(*P, *K, *tr, *ego, opp, n_opp, nominal, n_samples, T, key0, key1, step, sigma, costs, noise_out, traj_cost,
                            n_threads);
    else
        rollout_all<float>(*P, *K, *tr, *ego, opp, n_opp, nominal, n_samples, T, key0, key1, step, sigma, costs, noise_out, traj_cost,
                           n_threads);
    return 0;
}

int or_mppi_update(int f64, const dr_quad_params* P, const double* costs, const float* noise, const float* nominal, int n_samples, int T,
                   float lambda, const float* fill, double* updated, double* action, double* shifted, double* min_cost, double* nu,
                   double* weights)
{
    if (!P || !costs || !noise || !nominal || !fill || !updated || !action || !shifted || !min_cost || !nu || n_samples < 1 || T < 1)
        return 1;
    if (f64)
        update<double>(*P, costs, noise, nominal, n_samples, T, lambda, fill, updated, action, shifted, min_cost, nu, weights);
    else
        update<float>(*P, costs, noise, nominal, n_samples, T, lambda, fill, updated, action, shifted, min_cost, nu, weights);
    return 0;
}

// Closed-loop run of a batch on the CPU in FP32 (the CPU baseline and the
// multi-step parity check): per control step, every instance rolls out K
// samples against constant-velocity opponents snapshotted at the step start,
// updates, clamps, takes its action and shifts; then every controlled drone
// is integrated with its action.  nominal holds sum_i T_i*4 floats in
// instance order; steps and lambdas are per instance and updated in place.
int or_world_run(const dr_quad_params* P, const dr_cost_cfg* K, const dr_track* tracks, const dr_instance* inst, int n_inst,
                 const int32_t* drone_world, dr_drone_state* drones, int n_drones, float* nominal, uint32_t* steps, float* lambdas,
                 int n_steps, int n_threads, float* actions_out, float* traj_costs_out)
{
    if (!P || !K || !tracks || !inst || !drone_world || !drones || !nominal || !steps || !lambdas || n_inst < 1 || n_steps < 0)
        return 1;
    const Model<float> M(*P);
    std::vector<int> nom_off(n_inst);
    int off = 0;
    for (int i = 0; i < n_inst; ++i)
    {
        nom_off[i] = off;
        off += inst[i].n_timesteps * 4;
    }
    std::vector<float> act((size_t)n_inst * 4);
    for (int step = 0; step < n_steps; ++step)
    {
        std::vector<dr_drone_state> snap(drones, drones + n_drones);
        for (int i = 0; i < n_inst; ++i)
        {
            const dr_instance& I = inst[i];
            std::vector<dr_opponent> opp;
            for (int d = 0; d < n_drones; ++d)
                if (d != I.drone && drone_world[d] == drone_world[I.drone])
                {
                    dr_opponent o;
                    for (int c = 0; c < 3; ++c)
                    {
                        o.p[c] = snap[d].p[c];
                        o.v[c] = snap[d].v[c];
                    }
                    opp.push_back(o);
                }
            const int Kn = I.n_samples, T = I.n_timesteps;
            std::vector<double> costs(Kn);
            std::vector<float> noise((size_t)T * Kn * 4);
            double traj = 0.0;
            rollout_all<float>(*P, *K, tracks[I.track], snap[I.drone], opp.data(), (int)opp.size(), nominal + nom_off[i], Kn, T,
                               (uint32_t)(I.seed & 0xffffffffu), (uint32_t)(I.seed >> 32), steps[i], I.sigma, costs.data(), noise.data(),
                               &traj, n_threads);
            std::vector<double> upd((size_t)T * 4), sh((size_t)T * 4);
            double a[4], mc, nu;
            update<float>(*P, costs.data(), noise.data(), nominal + nom_off[i], Kn, T, lambdas[i], I.shift_fill, upd.data(), a, sh.data(),
                          &mc, &nu, nullptr);
            for (int j = 0; j < T * 4; ++j)
                nominal[nom_off[i] + j] = (float)sh[j];
            for (int d = 0; d < 4; ++d)
                act[(size_t)i * 4 + d] = (float)a[d];
            if (I.adapt_lambda)
            {
                const float nuf = (float)nu;
                if (nuf > 0.001f * (float)Kn)
                    lambdas[i] = lambdas[i] * 0.9f;
                else if (nuf < 0.0005f * (float)Kn)
                    lambdas[i] = lambdas[i] * 1.2f;
            }
            if (traj_costs_out)
                traj_costs_out[i] = (float)traj;
            steps[i] = steps[i] + 1u;
        }
        for (int d = 0; d < n_drones; ++d)
        {
            int ctrl = -1;
            for (int i = 0; i < n_inst; ++i)
                if (inst[i].drone == d)
                    ctrl = i;
            if (ctrl < 0)
            {
                for (int c = 0; c < 3; ++c)
                    drones[d].p[c] = fmaf(M.dt, drones[d].v[c], drones[d].p[c]);
                continue;
            }
            St<float> s = load_state<float>(drones[d]);
            float u[4];
            clamp_u(M, &act[(size_t)ctrl * 4], u);
            const float p0[3] = {s.p[0], s.p[1], s.p[2]};
            euler(M, s, u);
            gate_update(tracks[inst[ctrl].track], K->gate_traversal_margin, p0, s);
            for (int c = 0; c < 3; ++c)
            {
                drones[d].p[c] = s.p[c];
                drones[d].v[c] = s.v[c];
                drones[d].w[c] = s.w[c];
            }
            for (int c = 0; c < 4; ++c)
                drones[d].q[c] = s.q[c];
            drones[d].gate = s.gate;
            drones[d].laps = s.laps;
        }
    }
    if (actions_out)
        memcpy(actions_out, act.data(), act.size() * sizeof(float));
    return 0;
}

// Algorithmic FP work of the rollouts of n_samples samples (samples
// 0..n_samples-1 of the given key/step): flops spent in dynamics, gate logic
// and costs (op-counting scalar, data-dependent branches included) and the
// number of active sample-steps.  The per-step noise (Box-Muller + scaling,
// a fixed count read off the code) is not included here.
int or_count_rollout_ops(const dr_quad_params* P, const dr_cost_cfg* K, const dr_track* tr, const dr_drone_state* ego,
                         const dr_opponent* opp, int n_opp, const float* nominal, int n_samples, int T, uint32_t key0, uint32_t key1,
                         uint32_t step, const float* sigma, long long* flops_out)
{
    if (!P || !K || !tr || !ego || !nominal || !sigma || !flops_out || n_samples < 1 || T < 1)
        return 1;
    const Model<Cn> M(*P);
    std::vector<float> eps((size_t)T * 4);
    g_tally.flops = 0;
    for (int k = 0; k < n_samples; ++k)
    {
        for (int t = 0; t < T; ++t)
        {
            float z[4];
            normals4(key0, key1, (uint32_t)k, (uint32_t)t, step, 0u, z);
            for (int d = 0; d < 4; ++d)
                eps[(size_t)t * 4 + d] = sigma[d] * z[d];
        }
        (void)rollout_one<Cn>(*P, M, *K, *tr, *ego, opp, n_opp, nominal, T, eps.data(), 4);
    }
    *flops_out = g_tally.flops;
    return 0;
}

} // extern "C"
