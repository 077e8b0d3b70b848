// C++ host interface with the shape of the reference controllers
// (include/controllers.h:12-111 of pablo-gr-rey/drone-race), implemented
// over the C ABI of dr_api.h.  Header-only; link with libdr_amd.so.
//
//   Controller::getControl(const SimState&, float* outAction)  controllers.h:17
//   Controller::engine, Controller::h_belief                    controllers.h:19-20
//   MPPIController(EnvironmentConfig, MPPIConfig, seed, nominal) controllers.h:29
//   MPPIController::mppiConfig, h_nominal                       controllers.h:35-37
//
//   PRMPPIController(EnvironmentConfig, PRMPPIConfig, seed, nominal) controllers.h:120
//   PRMPPIController::h_nom_nominal, h_rob_nominal, resetNom, useNomPlan controllers.h:128-135
//   stratrace::MPPIController: the same MPPIController interface over the
//   reference's own StratRace2D environment (env_stratrace_defs.h:42-96).
//   stratrace::BranchingMPPIController, hiddenobs::BranchingMPPIController,
//   dronerace::BranchingMPPIController: the reference's default (branching,
//   spline) MPPIController on StratRace2D, HiddenObs (env_hiddenobs_defs.h)
//   and DroneRace (env_dronerace_defs.h).
//
// Errors from the C ABI are rethrown as std::runtime_error, as the
// reference's constructors and checks throw (mppi_controller.cu:27-29).
// MPPIController and PRMPPIController run the quadrotor gate race (13-state
// drone, constant-velocity opponents); stratrace::MPPIController runs the
// 2-D race with the reference's opponent model.
#pragma once

#include <array>
#include <cmath>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../dr_api.h"

namespace drone_race {

inline constexpr int ACTION_DIM = DR_ACTION_DIM;
inline constexpr int N_TRUE_MODELS = 1;
inline constexpr int MAX_OPPONENTS = DR_MAX_WORLD_DRONES - 1;

struct EnvironmentConfig
{
    dr_quad_params quad{};
    dr_track track{};
    dr_cost_cfg cost{};
    int maxOpponents = 0;
};

struct SimState
{
    dr_drone_state ego{};
    std::array<dr_opponent, MAX_OPPONENTS> opponents{};
    int nOpponents = 0;
};

// Field names follow MPPIConfig (include/config.h:116-154); samplingNoise is
// per control dimension here (the reference has one scalar for all).
struct MPPIConfig
{
    int nSamples = 1024;
    int nTimesteps = 50;
    float invTemperature = 1.0f; // the temperature lambda: w = exp(-(c - min)/lambda)
    std::array<float, ACTION_DIM> samplingNoise{2.0f, 0.5f, 0.5f, 0.5f};
    std::array<float, ACTION_DIM> shiftFill{0.0f, 0.0f, 0.0f, 0.0f};
    bool adaptTemperature = false;
};

class SimulationEngine;

inline void check(int rc, const char* what)
{
    if (rc != DR_OK)
        throw std::runtime_error(std::string(what) + ": " + dr_last_error());
}

// Owning handles: a constructor that throws after creating its engine releases it.
struct MppiDeleter
{
    void operator()(dr_mppi* h) const { dr_mppi_destroy(h); }
};
struct PrDeleter
{
    void operator()(dr_prengine* h) const { dr_pr_destroy(h); }
};
struct SrDeleter
{
    void operator()(dr_sr_engine* h) const { dr_sr_destroy(h); }
};
struct BmDeleter
{
    void operator()(dr_bm_engine* h) const { dr_bm_destroy(h); }
};

class Controller
{
  public:
    virtual ~Controller() = default;
    virtual void getControl(const SimState& state, float* outAction) = 0;

    SimulationEngine* engine = nullptr;
    std::array<float, N_TRUE_MODELS> h_belief{1.0f};
};

class MPPIController : public Controller
{
  public:
    MPPIController(const EnvironmentConfig& c, const MPPIConfig& mc, int seed, std::optional<std::vector<float>> nominal = std::nullopt)
        : mppiConfig(mc), envConfig(c)
    {
        dr_mppi_cfg cfg{};
        cfg.n_samples = mc.nSamples;
        cfg.n_timesteps = mc.nTimesteps;
        cfg.lambda = mc.invTemperature;
        for (int d = 0; d < ACTION_DIM; ++d)
        {
            cfg.sigma[d] = mc.samplingNoise[d];
            cfg.shift_fill[d] = mc.shiftFill[d];
        }
        cfg.adapt_lambda = mc.adaptTemperature ? 1 : 0;
        cfg.max_opponents = c.maxOpponents;
        if (nominal && (int)nominal->size() != mc.nTimesteps * ACTION_DIM)
            throw std::runtime_error("Invalid MPPI construction: expected nominal size " + std::to_string(mc.nTimesteps * ACTION_DIM) +
                                     ", got " + std::to_string(nominal->size()));
        dr_mppi* raw = nullptr;
        check(dr_mppi_create(&c.quad, &c.cost, &c.track, &cfg, (uint64_t)(uint32_t)seed, &raw), "MPPIController");
        h.reset(raw);
        h_nominal = nominal ? *nominal : std::vector<float>((size_t)mc.nTimesteps * ACTION_DIM, 0.0f);
        check(dr_mppi_set_nominal(h.get(), h_nominal.data(), (int)h_nominal.size()), "MPPIController");
    }

    MPPIController(const MPPIController&) = delete;
    MPPIController& operator=(const MPPIController&) = delete;

    void getControl(const SimState& state, float* outAction) override
    {
        check(dr_mppi_step(h.get(), &state.ego, state.opponents.data(), state.nOpponents, outAction), "getControl");
        check(dr_mppi_get_nominal(h.get(), h_nominal.data(), (int)h_nominal.size()), "getControl");
        dr_diag d{};
        check(dr_mppi_get_diag(h.get(), &d), "getControl");
        mppiConfig.invTemperature = d.lambda;
        lastMinCost = d.min_cost;
        lastNu = d.nu;
    }

    MPPIConfig mppiConfig;
    std::vector<float> h_nominal; // host mirror (T, dim), already shifted for the next step
    float lastMinCost = 0.0f;
    float lastNu = 0.0f;

  private:
    EnvironmentConfig envConfig;
    std::unique_ptr<dr_mppi, MppiDeleter> h;
};

// PR-MPPI over opponent-velocity branches (the reference's PRMPPIConfig
// fields, config.h:157-192, plus the branches of this build).  With
// sampled = true, P = ceil((1 - delta) / delta) models are drawn from the
// belief every step and averaged uniformly, as the reference does
// (python-src/utils.py:537-538, prmppi_kernels.cu:23-51, :160); otherwise
// the branches are averaged with the belief.
struct PRMPPIConfig
{
    int nSamples = 1024;
    int nTimesteps = 50;
    float invTemperature = 1.0f;
    std::array<float, ACTION_DIM> samplingNoise{2.0f, 0.5f, 0.5f, 0.5f};
    std::array<float, ACTION_DIM> shiftFill{0.0f, 0.0f, 0.0f, 0.0f};
    bool adaptTemperature = false;
    float safetyWeight = 10000.0f;
    float delta = 0.1f;
    bool sampled = false;
    std::vector<float> belief{1.0f};                           // one entry per branch
    std::vector<std::array<float, 3 * MAX_OPPONENTS>> velShift{{}}; // per branch: opponent j's shift at [3 j .. 3 j + 2]
};

class PRMPPIController : public Controller
{
  public:
    PRMPPIController(const EnvironmentConfig& c, const PRMPPIConfig& mc, int seed, std::optional<std::vector<float>> nominal = std::nullopt)
        : mppiConfig(mc), envConfig(c)
    {
        const int B = (int)mc.belief.size();
        if (B < 1 || B > DR_MAX_BRANCHES || (int)mc.velShift.size() != B)
            throw std::runtime_error("Invalid PRMPPI construction: need 1..8 branches with one velocity shift each");
        dr_pr_instance in{};
        in.n_samples = mc.nSamples;
        in.n_timesteps = mc.nTimesteps;
        in.track = 0;
        in.drone = 0;
        in.seed = (uint64_t)(uint32_t)seed;
        in.lambda = mc.invTemperature;
        for (int d = 0; d < ACTION_DIM; ++d)
        {
            in.sigma[d] = mc.samplingNoise[d];
            in.shift_fill[d] = mc.shiftFill[d];
        }
        in.adapt_lambda = mc.adaptTemperature ? 1 : 0;
        in.n_branches = B;
        in.sampled = mc.sampled ? 1 : 0;
        in.n_models = mc.sampled ? (int)std::ceil((1.0f - mc.delta) / mc.delta) : 0;
        in.safety_weight = mc.safetyWeight;
        for (int b = 0; b < B; ++b)
        {
            in.belief[b] = mc.belief[b];
            for (int j = 0; j < MAX_OPPONENTS; ++j)
                for (int k = 0; k < 3; ++k)
                    in.vel_shift[b][j][k] = mc.velShift[b][3 * j + k];
        }
        const int nd = 1 + c.maxOpponents;
        std::vector<int32_t> world(nd, 0);
        const size_t n = (size_t)mc.nTimesteps * ACTION_DIM;
        if (nominal && nominal->size() != n)
            throw std::runtime_error("Invalid PRMPPI construction: expected nominal size " + std::to_string(n) + ", got " +
                                     std::to_string(nominal->size()));
        dr_prengine* raw = nullptr;
        check(dr_pr_create(&c.quad, &c.cost, &c.track, 1, &in, 1, world.data(), nd, &raw), "PRMPPIController");
        h.reset(raw);
        h_nom_nominal = nominal ? *nominal : std::vector<float>(n, 0.0f);
        h_rob_nominal = h_nom_nominal;
        check(dr_pr_set_nominals(h.get(), 0, h_nom_nominal.data(), h_rob_nominal.data(), (int)n), "PRMPPIController");
    }

    PRMPPIController(const PRMPPIController&) = delete;
    PRMPPIController& operator=(const PRMPPIController&) = delete;

    void getControl(const SimState& state, float* outAction) override
    {
        if (state.nOpponents != envConfig.maxOpponents)
            throw std::runtime_error("getControl: nOpponents must equal maxOpponents");
        std::vector<dr_drone_state> w(1 + envConfig.maxOpponents);
        w[0] = state.ego;
        for (int j = 0; j < envConfig.maxOpponents; ++j)
        {
            w[1 + j] = dr_drone_state{};
            w[1 + j].q[0] = 1.0f;
            for (int d = 0; d < 3; ++d)
            {
                w[1 + j].p[d] = state.opponents[j].p[d];
                w[1 + j].v[d] = state.opponents[j].v[d];
            }
        }
        check(dr_pr_set_drones(h.get(), w.data(), (int)w.size()), "getControl");
        check(dr_pr_step(h.get(), 1, 0, outAction), "getControl");
        check(dr_pr_get_nominals(h.get(), 0, h_nom_nominal.data(), h_rob_nominal.data(), (int)h_nom_nominal.size()), "getControl");
        dr_pr_diag d{};
        check(dr_pr_get_diag(h.get(), &d, 1), "getControl");
        useNomPlan = d.use_nom_plan != 0;
        resetNom = d.reset_nom != 0;
        for (int j = 0; j < 3; ++j)
            temperatures[j] = d.lambda[j];
    }

    PRMPPIConfig mppiConfig;
    std::vector<float> h_nom_nominal; // (T, dim), unshifted
    std::vector<float> h_rob_nominal; // (T, dim), unshifted
    bool resetNom = false;
    bool useNomPlan = true;
    std::array<float, 3> temperatures{}; // nominal/full, robust/full, robust/safety

  private:
    EnvironmentConfig envConfig;
    std::unique_ptr<dr_prengine, PrDeleter> h;
};

namespace stratrace {

inline constexpr int ACTION_DIM = 2;

// The reference's SimState / EnvironmentConfig of the StratRace2D build
// (env_stratrace_defs.h:53-96), with the track grids held by value.
using SimState = dr_sr_state;
using EnvironmentConfig = dr_sr_env;

// MPPIConfig fields of the dense plain path (config.h:116-154).
struct MPPIConfig
{
    int nSamples = 65536;
    int nTimesteps = 70;
    float invTemperature = 10.0f;
    float samplingNoise = 1.5f;
    dr_sr_cost cost{1.3f, 50.0f, 1.7f, 1e6f, 1e6f, 1e5f, 500.0f, 0.0f};
    bool adaptTemperature = false;
};

class MPPIController
{
  public:
    // trueTheta: the opponent model the rollouts assume (a committed belief).
    MPPIController(const EnvironmentConfig& c, const MPPIConfig& mc, int seed, int trueTheta = 0,
                   std::optional<std::vector<float>> nominal = std::nullopt)
        : mppiConfig(mc)
    {
        dr_sr_instance in{};
        in.n_samples = mc.nSamples;
        in.n_timesteps = mc.nTimesteps;
        in.true_theta = trueTheta;
        in.adapt_lambda = mc.adaptTemperature ? 1 : 0;
        in.seed = (uint64_t)(uint32_t)seed;
        in.lambda = mc.invTemperature;
        in.sigma = mc.samplingNoise;
        const size_t n = (size_t)mc.nTimesteps * ACTION_DIM;
        if (nominal && nominal->size() != n)
            throw std::runtime_error("Invalid MPPI construction: expected nominal size " + std::to_string(n) + ", got " +
                                     std::to_string(nominal->size()));
        dr_sr_engine* raw = nullptr;
        check(dr_sr_create(&c, &mc.cost, &in, 1, &raw), "MPPIController");
        h.reset(raw);
        h_nominal = nominal ? *nominal : std::vector<float>(n, 0.0f);
        check(dr_sr_set_nominal(h.get(), 0, h_nominal.data(), (int)n), "MPPIController");
    }

    MPPIController(const MPPIController&) = delete;
    MPPIController& operator=(const MPPIController&) = delete;

    void getControl(const SimState& state, float* outAction)
    {
        check(dr_sr_set_states(h.get(), &state, 1), "getControl");
        check(dr_sr_step(h.get(), 1, 0, outAction), "getControl");
        check(dr_sr_get_nominal(h.get(), 0, h_nominal.data(), (int)h_nominal.size()), "getControl");
        dr_diag d{};
        check(dr_sr_get_diag(h.get(), &d, 1), "getControl");
        mppiConfig.invTemperature = d.lambda;
        lastMinCost = d.min_cost;
        lastNu = d.nu;
    }

    MPPIConfig mppiConfig;
    std::vector<float> h_nominal; // (T, 2), already shifted for the next step
    float lastMinCost = 0.0f;
    float lastNu = 0.0f;

  private:
    std::unique_ptr<dr_sr_engine, SrDeleter> h;
};

// MPPIConfig of the reference's default controller: branching over the
// opponent models with spline plans (config.h:116-154, USE_SPLINES).
// knots[0..nKnots) = 0 selects the default spacing m * floor((T - 2) / (M - 1))
// (utils.py:496-497).
struct BranchingMPPIConfig
{
    int nSamples = 65536;
    int nTimesteps = 80;
    int nKnots = 20;
    std::array<int, DR_MAX_KNOTS> knots{};
    float invTemperature = 10.0f;
    float samplingNoise = 0.5f;
    float minConfidence = 0.95f;
    dr_sr_cost cost{1.4f, 50.0f, 1.8f, 1e7f, 1e5f, 1e5f, 50.0f, 0.0f};
    // plan verification (0 samples: none, the new plan is always executed)
    int nVerifSamples = 0;
    float beta = 1e-6f;
    int verifHorizon = 40;
    float maxVerifEps = 0.01f;
};

namespace detail {

// The environment-specific calls of the branching engine: StratRace2D
// (dr_bm_*) and HiddenObs (dr_bmho_*); the rest of the dr_bm_* calls are
// shared.
struct BmStratRace
{
    using Env = dr_sr_env;
    using State = dr_sr_state;
    static int models(const Env& e) { return e.n_models; }
    static int create(const Env* e, const dr_sr_cost* k, const dr_bm_instance* in, dr_bm_engine** h) { return dr_bm_create(e, k, in, 1, h); }
    static int set_state(dr_bm_engine* h, const State* s) { return dr_bm_set_states(h, s, 1); }
};

struct BmDroneRace
{
    using Env = dr_drace_env;
    using State = dr_drace_state;
    static int models(const Env& e) { return e.n_models; }
    static int create(const Env* e, const dr_sr_cost* k, const dr_bm_instance* in, dr_bm_engine** h)
    {
        return dr_bmdrace_create(e, k, in, 1, h);
    }
    static int set_state(dr_bm_engine* h, const State* s) { return dr_bmdrace_set_states(h, s, 1); }
};

struct BmHiddenObs
{
    using Env = dr_ho_env;
    using State = dr_ho_state;
    static int models(const Env&) { return 2; }
    static int create(const Env* e, const dr_sr_cost* k, const dr_bm_instance* in, dr_bm_engine** h) { return dr_bmho_create(e, k, in, 1, h); }
    static int set_state(dr_bm_engine* h, const State* s) { return dr_bmho_set_states(h, s, 1); }
};

// MPPIController of the reference's branching build (controllers.h:23-111,
// USE_SPLINES) over the branching engine.  h_belief is the controller's
// belief, written by the caller's dynStep as in the reference
// (engine.cpp:332); h_nominal is the executed dense plans of every branch
// after the step ([nModels + 1][T][2], the chosen branch shifted).  With
// nVerifSamples > 0 the plan verification runs and failCount (per horizon
// step, of the executed plan), epsilon, epsilonPartial, certifiedLoss and
// useNewPlan are its outcome; without it failCount is all zero, the three
// bounds 0 and useNewPlan true.
template <class E, class Config>
class BranchingMPPI
{
  public:
    BranchingMPPI(const typename E::Env& c, const Config& mc, int seed, int trueTheta = 0,
                  std::optional<std::vector<float>> nominal = std::nullopt)
        : mppiConfig(mc), nModels(E::models(c))
    {
        dr_bm_instance in{};
        in.n_samples = mc.nSamples;
        in.n_timesteps = mc.nTimesteps;
        in.n_knots = mc.nKnots;
        bool given = false;
        for (int m = 0; m < mc.nKnots && m < DR_MAX_KNOTS; ++m)
            given = given || mc.knots[m] != 0;
        for (int m = 0; m < mc.nKnots && m < DR_MAX_KNOTS; ++m)
            in.knots[m] = given ? mc.knots[m] : m * ((mc.nTimesteps - 2) / (mc.nKnots > 1 ? mc.nKnots - 1 : 1));
        in.seed = (uint64_t)(uint32_t)seed;
        in.lambda = mc.invTemperature;
        in.sigma = mc.samplingNoise;
        in.min_confidence = mc.minConfidence;
        in.true_theta = trueTheta;
        in.adapt_lambda = 1; // mppi_controller.cu:413-424
        for (int th = 0; th < nModels && th < DR_SR_MAX_MODELS; ++th)
            in.init_belief[th] = 1.0f / (float)nModels;
        in.n_verif_samples = mc.nVerifSamples;
        in.verif_horizon = mc.verifHorizon;
        in.beta = mc.beta;
        in.max_verif_eps = mc.maxVerifEps;
        const size_t n = (size_t)(nModels + 1) * mc.nTimesteps * 2;
        if (nominal && nominal->size() != n)
            throw std::runtime_error("Invalid MPPI construction: expected nominal size " + std::to_string(n) + ", got " +
                                     std::to_string(nominal->size()));
        dr_bm_engine* raw = nullptr;
        check(E::create(&c, &mc.cost, &in, &raw), "MPPIController");
        h.reset(raw);
        failCount.assign((size_t)mc.nTimesteps, 0u);
        h_nominal = nominal ? *nominal : std::vector<float>(n, 0.0f);
        h_belief.fill(0.0f);
        for (int th = 0; th < nModels; ++th)
            h_belief[th] = in.init_belief[th];
    }

    BranchingMPPI(const BranchingMPPI&) = delete;
    BranchingMPPI& operator=(const BranchingMPPI&) = delete;

    void getControl(const typename E::State& state, float* outAction)
    {
        check(dr_bm_set_belief(h.get(), 0, h_belief.data(), nModels), "getControl");
        check(E::set_state(h.get(), &state), "getControl");
        check(dr_bm_step(h.get(), 1, 0, outAction), "getControl");
        check(dr_bm_get_nominal(h.get(), 0, h_nominal.data(), (int)h_nominal.size()), "getControl");
        dr_bm_diag d{};
        check(dr_bm_get_diag(h.get(), &d, 1), "getControl");
        mppiConfig.invTemperature = d.lambda;
        chosenBranch = d.branch;
        lastNu = d.nu;
        lastMinCost = d.min_cost;
        useNewPlan = d.use_new_plan != 0;
        epsilon = d.epsilon;
        epsilonPartial = d.epsilon_partial;
        certifiedLoss = d.certified_loss;
        if (mppiConfig.nVerifSamples > 0)
        {
            std::vector<int32_t> fn(failCount.size()), fo(failCount.size());
            check(dr_bm_get_fail_counts(h.get(), 0, fn.data(), fo.data(), (int)failCount.size()), "getControl");
            for (size_t t = 0; t < failCount.size(); ++t)
                failCount[t] = (unsigned)(useNewPlan ? fn[t] : fo[t]);
        }
    }

    Config mppiConfig;
    std::vector<float> h_nominal;                    // (nModels + 1, T, 2)
    std::array<float, DR_SR_MAX_MODELS> h_belief{};  // written by the caller's world step
    std::vector<unsigned> failCount;                 // (T) verification failures of the executed plan
    double epsilon = 0.0, epsilonPartial = 0.0, certifiedLoss = 0.0;
    bool useNewPlan = true;
    int chosenBranch = 0; // 0 nominal, theta + 1 committed
    float lastNu = 0.0f, lastMinCost = 0.0f;

  private:
    int nModels;
    std::unique_ptr<dr_bm_engine, BmDeleter> h;
};

} // namespace detail

// The reference's default controller on StratRace2D: branching over the
// opponent models.
using BranchingMPPIController = detail::BranchingMPPI<detail::BmStratRace, BranchingMPPIConfig>;

} // namespace stratrace

namespace hiddenobs {

// The reference's SimState / EnvironmentConfig of the HiddenObs build
// (env_hiddenobs_defs.h:48-105, without the init belief, which the
// controller holds).
using SimState = dr_ho_state;
using EnvironmentConfig = dr_ho_env;

// MPPIConfig of hiddenObsEnv with splines (python-src/main.py:811-833).
struct BranchingMPPIConfig : stratrace::BranchingMPPIConfig
{
    BranchingMPPIConfig()
    {
        samplingNoise = 0.8f;
        cost = dr_sr_cost{1.2f, 10.0f, 1.4f, 1e7f, 1e6f, 0.0f, 500.0f, 0.0f};
    }
};

// MPPIController of the HiddenObs build: branching over which hidden
// obstacle exists (two models).
using BranchingMPPIController = stratrace::detail::BranchingMPPI<stratrace::detail::BmHiddenObs, BranchingMPPIConfig>;

} // namespace hiddenobs

namespace dronerace {

// The reference's SimState / EnvironmentConfig of the DroneRace build
// (env_dronerace_defs.h:45-107; counts as fields, racelines held by value).
using SimState = dr_drace_state;
using EnvironmentConfig = dr_drace_env;
using PIDConfig = dr_drace_pid;

// MPPIConfig of tinyGateEnv with splines (python-src/main.py:140-161).
struct BranchingMPPIConfig : stratrace::BranchingMPPIConfig
{
    BranchingMPPIConfig()
    {
        nTimesteps = 60;
        nKnots = 10;
        samplingNoise = 2.0f;
        cost = dr_sr_cost{1.1f, 0.0f, 1.7f, 1e6f, 1e6f, 1000.0f, 200.0f, 0.0f};
    }
};

// MPPIController of the DroneRace build: branching over the opponent's PD
// configurations.
using BranchingMPPIController = stratrace::detail::BranchingMPPI<stratrace::detail::BmDroneRace, BranchingMPPIConfig>;

} // namespace dronerace

} // namespace drone_race
