// Environment policies of the branching MPPI kernels: what a rollout lane
// needs from an environment (its configuration, state, LDS staging, one
// step with belief update, the running and terminal costs), so one set of
// kernels serves StratRace2D (dr_sr.h), HiddenObs (dr_ho.h) and DroneRace
// (dr_drace.h).
#pragma once

#include "dr_drace.h"
#include "dr_ho.h"
#include "dr_sr.h"

namespace dr {

struct SrPolicy
{
    // the rollout holds its branch's knot noise in registers (true) or reads
    // it back from HBM each step (false); measured on MI355X at K = 2^16,
    // T = 80, 20 knots over 30 closed-loop steps from the start: registers
    // 1062 control steps/s (645 with plan verification), HBM 1059 (589)
    static constexpr bool REG_NOISE = true;
    using Env = dr_sr_env;
    using State = sr::St;
    using Abi = dr_sr_state;
    // track tables p, t (NS x 2), kappa (NS) and the belief constants
    static constexpr int LDS_FLOATS = sr::NS * 5 + 2 * DR_SR_MAX_MODELS;
    struct Ctx
    {
        sr::Tables tb;
        const float* bc;
    };
    // fills sh (all threads of the block; the caller synchronises)
    DR_HD static Ctx stage(float* sh, const Env& E, int tid, int nthr)
    {
        for (int i = tid; i < sr::NS * 2; i += nthr)
        {
            sh[i] = (&E.p_grid[0][0])[i];
            sh[sr::NS * 2 + i] = (&E.t_grid[0][0])[i];
        }
        for (int i = tid; i < sr::NS; i += nthr)
            sh[sr::NS * 4 + i] = E.kappa_grid[i];
        float* bc = sh + sr::NS * 5;
        if (tid < E.n_models)
            sr::belief_consts(E, tid, &bc[tid], &bc[DR_SR_MAX_MODELS + tid]);
        return Ctx{sr::Tables{sh, sh + sr::NS * 2, sh + sr::NS * 4, E.track_length}, bc};
    }
    // tables read from the configuration itself (single-lane kernels)
    DR_HD static Ctx direct(const Env& E)
    {
        return Ctx{sr::Tables{&E.p_grid[0][0], &E.t_grid[0][0], E.kappa_grid, E.track_length}, nullptr};
    }
    DR_HD static int n_models(const Env& E) { return E.n_models; }
    DR_HD static int step(const Ctx& c, const Env& E, int theta, const float ego[2], State& s, sr::NormalSeq& rng, bool lose_on_opp_win,
                          float* belief, bool update)
    {
        return sr::env_step_t<true>(c.tb, E, theta, ego, s, rng, lose_on_opp_win, belief, c.bc, update);
    }
    DR_HD static float stage_cost(const Ctx&, const Env& E, const dr_sr_cost& K, const State& s, int) { return sr::state_cost(E, K, s); }
    DR_HD static float final_cost(const Ctx& c, const Env& E, const dr_sr_cost& K, const State& s)
    {
        return sr::final_cost(E, K, s, c.tb.L);
    }
    DR_HD static State load(const Abi& a) { return sr::state_from_abi(a); }
    DR_HD static void store(const State& s, Abi& a) { sr::state_to_abi(s, a); }
};

struct HoPolicy
{
    // HiddenObs: registers 0.123 ms, HBM 0.144 ms per rollout launch at the
    // same size
    static constexpr bool REG_NOISE = true;
    using Env = dr_ho_env;
    using State = ho::St;
    using Abi = dr_ho_state;
    static constexpr int LDS_FLOATS = 1;
    struct Ctx
    {
    };
    DR_HD static Ctx stage(float*, const Env&, int, int) { return Ctx{}; }
    DR_HD static Ctx direct(const Env&) { return Ctx{}; }
    DR_HD static int n_models(const Env&) { return 2; }
    DR_HD static int step(const Ctx&, const Env& E, int theta, const float ego[2], State& s, sr::NormalSeq& rng, bool, float* belief,
                          bool update)
    {
        return ho::env_step_t<true>(E, theta, ego, s, rng, belief, update);
    }
    DR_HD static float stage_cost(const Ctx&, const Env& E, const dr_sr_cost& K, const State& s, int theta)
    {
        return ho::state_cost(E, K, s, theta);
    }
    DR_HD static float final_cost(const Ctx&, const Env& E, const dr_sr_cost& K, const State& s) { return ho::final_cost(E, K, s); }
    DR_HD static State load(const Abi& a) { return ho::state_from_abi(a); }
    DR_HD static void store(const State& s, Abi& a) { ho::state_to_abi(s, a); }
};

struct DracePolicy
{
    static constexpr bool REG_NOISE = true;
    using Env = dr_drace_env;
    using State = drace::St;
    using Abi = dr_drace_state;
    // racelines (up to 4 x 512 x 2) and the belief constants
    static constexpr int LDS_FLOATS = DR_DRACE_MAX_RACELINES * DR_DRACE_SAMPLES * 2 + 2 * DR_SR_MAX_MODELS;
    struct Ctx
    {
        drace::Lines L;
        const float* bc;
    };
    DR_HD static Ctx stage(float* sh, const Env& E, int tid, int nthr)
    {
        const int n = E.n_racelines * DR_DRACE_SAMPLES * 2;
        for (int i = tid; i < n; i += nthr)
            sh[i] = (&E.track_points[0][0][0])[i];
        float* bc = sh + DR_DRACE_MAX_RACELINES * DR_DRACE_SAMPLES * 2;
        if (tid < E.n_models)
            drace::belief_consts(E, tid, &bc[tid], &bc[DR_SR_MAX_MODELS + tid]);
        return Ctx{drace::Lines{sh}, bc};
    }
    DR_HD static Ctx direct(const Env& E) { return Ctx{drace::Lines{&E.track_points[0][0][0]}, nullptr}; }
    DR_HD static int n_models(const Env& E) { return E.n_models; }
    DR_HD static int step(const Ctx& c, const Env& E, int theta, const float ego[2], State& s, sr::NormalSeq& rng, bool lose_on_opp_win,
                          float* belief, bool update)
    {
        return drace::env_step_t<true>(c.L, E, theta, ego, s, rng, lose_on_opp_win, belief, c.bc, update);
    }
    DR_HD static float stage_cost(const Ctx&, const Env& E, const dr_sr_cost& K, const State& s, int) { return drace::state_cost(E, K, s); }
    DR_HD static float final_cost(const Ctx&, const Env& E, const dr_sr_cost& K, const State& s) { return drace::final_cost(E, K, s); }
    DR_HD static State load(const Abi& a) { return drace::state_from_abi(a); }
    DR_HD static void store(const State& s, Abi& a) { drace::state_to_abi(s, a); }
};

} // namespace dr
