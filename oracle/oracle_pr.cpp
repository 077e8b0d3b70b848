// ORACLE — test infrastructure only (see oracle.h).  CPU restatement of one
// PR-MPPI control step of one controller for the quadrotor race, templated
// on the scalar: float repeats the engine's FP32 operation sequence, double
// is the FP64 oracle.
//
// Reference logic restated (pablo-gr-rey/drone-race):
//   common-random-number noise (T, N, dim)     src/prmppi_kernels.cu:8-20
//   model draw from the belief                 src/prmppi_kernels.cu:23-51
//   nominal/robust rollouts per (model, s)     src/prmppi_kernels.cu:55-130
//   model average + safety penalty, max safety src/prmppi_kernels.cu:134-166
//   three minima                               src/prmppi_kernels.cu:169-199
//   three weighted averages sum w(u+eps)/sum w src/prmppi_kernels.cu:204-282
//   candidate costs, safety of the chosen plan src/prmppi_kernels.cu:285-393
//   choice, action, temperature adaptation     src/prmppi_controller.cu:227-346
// Opponent models are the quadrotor's: model b moves every opponent j at
// constant velocity v_j + vel_shift[b][j].  The shift of both plans at the
// start of a step (prmppi_controller.cu:163-175) is the caller's: nom and
// rob passed here are the plans already shifted.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "oracle.h"

using namespace orc;

namespace {

void model_opponents(const dr_opponent* opp, int n_opp, const dr_pr_instance& I, int b, dr_opponent* out)
{
    for (int j = 0; j < n_opp; ++j)
    {
        out[j] = opp[j];
        for (int i = 0; i < 3; ++i)
            out[j].v[i] = opp[j].v[i] + I.vel_shift[b][j][i];
    }
}

// One rollout of plan U (+ eps when not NULL) under one opponent model
// (prmppi_kernels.cu:98-127, :315-340): running cost * decay, max of the
// safety over the visited states, early stop, terminal cost.
template <class S, class UT>
void pr_rollout(const Model<S>& M, const CostK<S>& Q, const dr_cost_cfg& K, const dr_track& tr, const dr_drone_state& ego,
                const dr_opponent* opp, int n_opp, const UT* U, const float* eps, int T, S* cost, S* safe)
{
    St<S> s = load_state<S>(ego);
    S acc = 0, decay = 1, smax = -(S)INFINITY;
    const S margin = (S)K.gate_traversal_margin;
    for (int t = 0; t < T; ++t)
    {
        S uin[4], u[4];
        for (int d = 0; d < 4; ++d)
        {
            const float e = eps ? eps[t * 4 + d] : 0.0f;
            if (sizeof(S) == 4)
                uin[d] = (S)((float)U[t * 4 + d] + e);
            else
                uin[d] = (S)U[t * 4 + d] + (S)e;
        }
        clamp_u(M, uin, u);
        const S p0[3] = {s.p[0], s.p[1], s.p[2]};
        euler(M, s, u);
        gate_update(tr, margin, p0, s);
        int term = 0;
        S sf;
        const S tau = (S)(M.dt * (S)(t + 1));
        const S c = pr_running_cost(M, Q, K, tr, s, opp, n_opp, tau, &sf, &term);
        acc = fm(c, decay, acc);
        smax = mx(smax, sf);
        decay = decay * (S)0.97f;
        if (term)
            break;
    }
    *cost = acc + (-(S)K.final_adv_weight * progress(tr, s));
    *safe = smax;
}

// Model average of c[0..M): sampled models uniformly (sum, then / P as
// prmppi_kernels.cu:152-160), branches weighted by the belief.
template <class S>
S model_avg(const dr_pr_instance& I, const float* belief, const S* c, int M)
{
    S acc = 0;
    if (I.sampled)
    {
        for (int m = 0; m < M; ++m)
            acc = acc + c[m];
        acc = acc / (S)M;
    }
    else
        for (int m = 0; m < M; ++m)
            acc = fm((S)belief[m], c[m], acc);
    return acc;
}

int n_models(const dr_pr_instance& I) { return I.sampled ? I.n_models : I.n_branches; }

template <class S>
void pr_step(const dr_quad_params& P, const dr_cost_cfg& K, const dr_track& tr, const dr_drone_state& ego, const dr_opponent* opp,
             int n_opp, const dr_pr_instance& I, const float* belief, uint32_t step, const float lam[3], const float* nom, const float* rob,
             int n_threads, float* noise_out, double* pair_out, double* agg_out, int* models_out, double* cand_out, double* res_out)
{
    const int N = I.n_samples, T = I.n_timesteps, M = n_models(I);
    const uint32_t key0 = (uint32_t)(I.seed & 0xffffffffu), key1 = (uint32_t)(I.seed >> 32);
    const Model<S> Md(P);
    const CostK<S> Q(P, K);
    std::vector<int> br(M);
    for (int m = 0; m < M; ++m)
        br[m] = I.sampled ? sample_branch(key0, key1, (uint32_t)m, step, belief, I.n_branches) : m;
    if (models_out)
        for (int m = 0; m < M; ++m)
            models_out[m] = br[m];
    const int no = n_opp > 0 ? n_opp : 1;
    std::vector<dr_opponent> mopp((size_t)M * no);
    for (int m = 0; m < M; ++m)
        model_opponents(opp, n_opp, I, br[m], &mopp[(size_t)m * no]);

    std::vector<float> eps_all((size_t)N * T * 4);
    std::vector<S> fn(N), fr(N), sr(N);
#pragma omp parallel for schedule(static) num_threads(n_threads > 0 ? n_threads : 1)
    for (int k = 0; k < N; ++k)
    {
        float* eps = &eps_all[(size_t)k * T * 4];
        for (int t = 0; t < T; ++t)
        {
            float z[4];
            normals4(key0, key1, (uint32_t)k, (uint32_t)t, step, 0u, z);
            for (int d = 0; d < 4; ++d)
                eps[t * 4 + d] = I.sigma[d] * z[d];
        }
        S c[2][DR_MAX_MODELS], sf[2][DR_MAX_MODELS];
        for (int plan = 0; plan < 2; ++plan)
            for (int m = 0; m < M; ++m)
            {
                pr_rollout<S, float>(Md, Q, K, tr, ego, &mopp[(size_t)m * no], n_opp, plan ? rob : nom, eps, T, &c[plan][m], &sf[plan][m]);
                if (pair_out)
                {
                    double* o = pair_out + (((size_t)plan * M + m) * N + k) * 2;
                    o[0] = (double)c[plan][m];
                    o[1] = (double)sf[plan][m];
                }
            }
        S full[2], smax[2];
        for (int plan = 0; plan < 2; ++plan)
        {
            const S avg = model_avg<S>(I, belief, c[plan], M);
            S mxs = -(S)INFINITY;
            for (int m = 0; m < M; ++m)
                mxs = mx(mxs, sf[plan][m]);
            full[plan] = mxs > (S)0 ? avg + (S)I.safety_weight : avg;
            smax[plan] = mxs;
        }
        fn[k] = full[0];
        fr[k] = full[1];
        sr[k] = smax[1];
        if (agg_out)
        {
            agg_out[(size_t)k * 4 + 0] = (double)full[0];
            agg_out[(size_t)k * 4 + 1] = (double)smax[0];
            agg_out[(size_t)k * 4 + 2] = (double)full[1];
            agg_out[(size_t)k * 4 + 3] = (double)smax[1];
        }
    }
    if (noise_out)
        memcpy(noise_out, eps_all.data(), eps_all.size() * sizeof(float));

    const std::vector<S>* cs[3] = {&fn, &fr, &sr};
    for (int j = 0; j < 3; ++j)
    {
        const std::vector<S>& c = *cs[j];
        S mnc = (S)INFINITY;
        for (int k = 0; k < N; ++k)
            mnc = mn(mnc, c[k]);
        const float* base = j == 0 ? nom : rob;
        std::vector<double> num((size_t)T * 4, 0.0);
        double den = 0.0;
        for (int k = 0; k < N; ++k)
        {
            double w;
            if (sizeof(S) == 4)
                w = isfinite((double)mnc) ? (double)exp32(-((float)c[k] - (float)mnc) / lam[j]) : 0.0;
            else
                w = isfinite((double)mnc) ? exp(-((double)c[k] - (double)mnc) / (double)lam[j]) : 0.0;
            const float* eps = &eps_all[(size_t)k * T * 4];
            for (int i = 0; i < T * 4; ++i)
            {
                const double u = sizeof(S) == 4 ? (double)(base[i] + eps[i]) : (double)base[i] + (double)eps[i];
                num[i] += w * u;
            }
            den += w;
        }
        for (int i = 0; i < T * 4; ++i)
        {
            double v = (double)base[i];
            if (den > 1e-30)
                v = sizeof(S) == 4 ? (double)(float)(num[i] / den) : num[i] / den;
            cand_out[(size_t)j * T * 4 + i] = v;
        }
        res_out[j] = (double)mnc;
        res_out[3 + j] = sizeof(S) == 4 ? (double)(float)den : den;
    }
}

// Candidate costs and the choice (prmppi_kernels.cu:285-393,
// prmppi_controller.cu:227-346).  cand = cand1, cand2, new robust plan.
// res: cand cost 1, 2, use_nom, reset_nom, action[4], lambda[3].
template <class S>
void pr_choose(const dr_quad_params& P, const dr_cost_cfg& K, const dr_track& tr, const dr_drone_state& ego, const dr_opponent* opp,
               int n_opp, const dr_pr_instance& I, const float* belief, const int* models, const double* cand, const float lam[3],
               const double nu[3], double* cc_out, double* cs_out, double* res_out)
{
    const int T = I.n_timesteps, M = n_models(I);
    const Model<S> Md(P);
    const CostK<S> Q(P, K);
    const int no = n_opp > 0 ? n_opp : 1;
    std::vector<dr_opponent> mopp((size_t)M * no);
    for (int m = 0; m < M; ++m)
        model_opponents(opp, n_opp, I, models[m], &mopp[(size_t)m * no]);
    std::vector<S> U((size_t)3 * T * 4);
    for (size_t i = 0; i < U.size(); ++i)
        U[i] = sizeof(S) == 4 ? (S)(float)cand[i] : (S)cand[i];
    S cc[2][DR_MAX_MODELS], sf[2][DR_MAX_MODELS];
    for (int which = 0; which < 2; ++which)
        for (int m = 0; m < M; ++m)
        {
            S c;
            pr_rollout<S, S>(Md, Q, K, tr, ego, &mopp[(size_t)m * no], n_opp, &U[(size_t)which * T * 4], nullptr, T, &c, &sf[which][m]);
            cc[which][m] = sf[which][m] > (S)0 ? c + (S)I.safety_weight : c;
            cc_out[which * M + m] = (double)cc[which][m];
            cs_out[which * M + m] = (double)sf[which][m];
        }
    const S c1 = model_avg<S>(I, belief, cc[0], M);
    const S c2 = model_avg<S>(I, belief, cc[1], M);
    const int pick = c1 <= c2 ? 0 : 1;
    int use_nom = 1;
    for (int m = 0; m < M; ++m)
        if (sf[pick][m] > (S)0)
            use_nom = 0;
    res_out[0] = (double)c1;
    res_out[1] = (double)c2;
    res_out[2] = use_nom;
    res_out[3] = c2 < c1 ? 1 : 0;
    const S* act = use_nom ? &U[(size_t)pick * T * 4] : &U[(size_t)2 * T * 4];
    for (int d = 0; d < 4; ++d)
        res_out[4 + d] = (double)act[d];
    for (int j = 0; j < 3; ++j)
    {
        float l = lam[j];
        if (I.adapt_lambda)
        {
            const float v = (float)nu[j];
            if (v > 0.001f * (float)I.n_samples && l > 0.01f)
                l = l * 0.9f;
            else if (v < 0.0005f * (float)I.n_samples && l < 100.0f)
                l = l * 1.2f;
        }
        res_out[8 + j] = (double)l;
    }
}

} // namespace

extern "C" {

// One PR-MPPI step of one controller: noise [k][t][4] (may be NULL), pair
// costs [plan][m][k][cost, safety] (may be NULL), per-sample aggregates
// [k][full nom, max safety nom, full rob, max safety rob] (may be NULL), the
// models' branches [M], candidates [cand1 | cand2 | new rob][T][4] and
// res = min[3], nu[3].  nom and rob are the already shifted plans.
int or_pr_step(int f64, const dr_quad_params* P, const dr_cost_cfg* K, const dr_track* tr, const dr_drone_state* ego, const dr_opponent* opp,
               int n_opp, const dr_pr_instance* inst, const float* belief, uint32_t step, const float* lam, const float* nom, const float* rob,
               int n_threads, float* noise_out, double* pair_out, double* agg_out, int* models_out, double* cand_out, double* res_out)
{
    if (!P || !K || !tr || !ego || !inst || !belief || !lam || !nom || !rob || !cand_out || !res_out)
        return 1;
    if (n_models(*inst) < 1 || n_models(*inst) > DR_MAX_MODELS || n_opp < 0 || n_opp > DR_MAX_WORLD_DRONES - 1)
        return 1;
    if (f64)
        pr_step<double>(*P, *K, *tr, *ego, opp, n_opp, *inst, belief, step, lam, nom, rob, n_threads, noise_out, pair_out, agg_out, models_out,
                        cand_out, res_out);
    else
        pr_step<float>(*P, *K, *tr, *ego, opp, n_opp, *inst, belief, step, lam, nom, rob, n_threads, noise_out, pair_out, agg_out, models_out,
                       cand_out, res_out);
    return 0;
}

// Candidate evaluation and choice for given candidates [3][T][4]:
// cc/cs [which][m] = cost (with the safety penalty) and max safety;
// res = cand cost 1, 2, use_nom, reset_nom, action[4], lambda after
// adaptation[3].
int or_pr_choose(int f64, const dr_quad_params* P, const dr_cost_cfg* K, const dr_track* tr, const dr_drone_state* ego,
                 const dr_opponent* opp, int n_opp, const dr_pr_instance* inst, const float* belief, const int* models, const double* cand,
                 const float* lam, const double* nu, double* cc_out, double* cs_out, double* res_out)
{
    if (!P || !K || !tr || !ego || !inst || !belief || !models || !cand || !lam || !nu || !cc_out || !cs_out || !res_out)
        return 1;
    if (n_models(*inst) < 1 || n_models(*inst) > DR_MAX_MODELS || n_opp < 0 || n_opp > DR_MAX_WORLD_DRONES - 1)
        return 1;
    if (f64)
        pr_choose<double>(*P, *K, *tr, *ego, opp, n_opp, *inst, belief, models, cand, lam, nu, cc_out, cs_out, res_out);
    else
        pr_choose<float>(*P, *K, *tr, *ego, opp, n_opp, *inst, belief, models, cand, lam, nu, cc_out, cs_out, res_out);
    return 0;
}

} // extern "C"
