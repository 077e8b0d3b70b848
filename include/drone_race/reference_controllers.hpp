// The controller classes the reference's SimulationEngine holds and
// downcasts (pablo-gr-rey/drone-race, include/controllers.h:12-179,
// src/engine.cpp:14-363), implemented over the C ABI of dr_api.h for the
// reference's StratRace2D build.  Header-only; link with libdr_amd.so.
//
//   Controller            controllers.h:12-21   getControl, engine, h_belief
//   MPPIController        controllers.h:23-111  the branching spline controller
//                                               (USE_SPLINES, config.cpp:18-21) over dr_bm_*
//   PRMPPIController      controllers.h:114-179 PR-MPPI with P models drawn from the
//                                               belief each step over dr_prsr_*
//
// The classes are templates over the host program's own types, so that the
// reference engine's std::unique_ptr<Controller>, its std::make_unique
// constructor calls (engine.cpp:29-47) and its dynamic_casts in sendState and
// dynStep (engine.cpp:73, 169, 300) work unchanged.  A types bundle Ty gives:
//   Ty::SimState, Ty::EnvironmentConfig     the host's state and environment (the reference's
//                                            env_stratrace_defs.h:53-96)
//   Ty::MPPIConfig, Ty::PRMPPIConfig        the host's controller configs, read by the reference's
//                                            field names (config.h:116-192)
//   Ty::Engine                               the engine class a controller points back to
//   Ty::N_TRUE_MODELS                        the size of h_belief
//   Ty::abi_env(const EnvironmentConfig&)   -> dr_sr_env (track tables by value)
//   Ty::abi_state(const SimState&)          -> dr_sr_state
//   Ty::init_belief(const EnvironmentConfig&, float* out)  the initial belief (envConfig.initBelief)
// StratRaceTypes below is such a bundle over this library's own structs.
//
// Behaviour kept from the reference: h_belief is the controller's belief,
// written by the engine's world step (engine.cpp:332) and read at every
// getControl; the temperatures adapt every step (mppi_controller.cu:413-424,
// prmppi_controller.cu:328-346); constructors throw std::runtime_error on a
// nominal of the wrong size.  Differences: with nVerifSamples = 0 the new plan
// is always executed (the reference would certify a loss of 1 and keep the
// previous plan); the reference's own host rollouts for display
// (sendState) stay in the engine.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../dr_api.h"

namespace drone_race {
namespace reference {

inline void check(int rc, const char* what)
{
    if (rc != DR_OK)
        throw std::runtime_error(std::string(what) + ": " + dr_last_error());
}

struct BmDeleter
{
    void operator()(dr_bm_engine* h) const { dr_bm_destroy(h); }
};
struct PrDeleter
{
    void operator()(dr_prengine* h) const { dr_pr_destroy(h); }
};

template <class Ty>
class Controller
{
  public:
    virtual ~Controller() = default;
    virtual void getControl(const typename Ty::SimState& state, float* outAction) = 0;

    typename Ty::Engine* engine = nullptr;
    std::array<float, Ty::N_TRUE_MODELS> h_belief{}; // updated by the engine's world step
};

// The reference's MPPIController: NT + 1 branch plans of nKnots spline knots,
// rollouts branching on the belief, masked minima, coefficient-weighted knot
// update, plan verification when nVerifSamples > 0.
template <class Ty>
class MPPIController : public Controller<Ty>
{
  public:
    using MPPIConfig = typename Ty::MPPIConfig;

    MPPIController() = default;

    // nominal: the dense plans (nBranchPlans, T, dim) shown before the first step; the knots start at zero
    MPPIController(const typename Ty::EnvironmentConfig& c, const MPPIConfig& mc, int s,
                   std::optional<std::vector<float>> nominal = std::nullopt)
        : mppiConfig(mc)
    {
        const dr_sr_env env = Ty::abi_env(c);
        n_models = env.n_models;
        if (n_models < 1 || n_models > Ty::N_TRUE_MODELS || n_models > DR_SR_MAX_MODELS)
            throw std::runtime_error("Invalid MPPI construction: the environment's model count does not fit N_TRUE_MODELS");
        const size_t n_plan = (size_t)(n_models + 1) * mc.nTimesteps * 2;
        if (nominal && nominal->size() != n_plan)
            throw std::runtime_error("Invalid MPPI construction: expected nominal size " + std::to_string(n_plan) + ", got " +
                                     std::to_string(nominal->size()));
        dr_sr_cost cost{mc.collDistFactor, mc.boundaryCost, mc.boundaryThresholdFactor, mc.outsideCost,
                        mc.winCost,        mc.oppWinCost,   mc.finalAdvWeight,          mc.finalOppAdvWeight};
        dr_bm_instance in{};
        in.n_samples = mc.nSamples;
        in.n_timesteps = mc.nTimesteps;
        in.n_knots = mc.nKnots;
        for (int m = 0; m < mc.nKnots && m < DR_MAX_KNOTS; ++m)
            in.knots[m] = mc.knots[m];
        in.seed = (uint64_t)(uint32_t)s;
        in.lambda = mc.invTemperature;
        in.sigma = mc.samplingNoise;
        in.min_confidence = mc.minConfidence;
        in.true_theta = 0; // the world is the engine's; this handle never advances it
        in.adapt_lambda = 1;
        float belief[DR_SR_MAX_MODELS] = {};
        Ty::init_belief(c, belief);
        for (int th = 0; th < n_models; ++th)
            in.init_belief[th] = belief[th];
        in.n_verif_samples = mc.nVerifSamples;
        in.verif_horizon = mc.nVerifSamples > 0 ? mc.verifHorizon : 0;
        in.beta = mc.beta;
        in.max_verif_eps = mc.maxVerifEps;
        dr_bm_engine* raw = nullptr;
        check(dr_bm_create(&env, &cost, &in, 1, &raw), "MPPIController");
        h.reset(raw);
        this->h_belief.fill(0.0f);
        for (int th = 0; th < n_models; ++th)
            this->h_belief[th] = belief[th];
        h_nominal = nominal ? *nominal : std::vector<float>(n_plan, 0.0f);
        failCountNew.assign((size_t)mc.nTimesteps, 0u);
        failCountOld.assign((size_t)mc.nTimesteps, 0u);
        failCount.assign((size_t)mc.nTimesteps, 0u);
    }

    MPPIController(const MPPIController&) = delete;
    MPPIController& operator=(const MPPIController&) = delete;

    void getControl(const typename Ty::SimState& state, float* outAction) override
    {
        if (!h)
            throw std::runtime_error("getControl: controller was default-constructed");
        const dr_sr_state st = Ty::abi_state(state);
        check(dr_bm_set_belief(h.get(), 0, this->h_belief.data(), n_models), "getControl");
        check(dr_bm_set_states(h.get(), &st, 1), "getControl");
        check(dr_bm_step(h.get(), 1, 0, outAction), "getControl");
        check(dr_bm_get_nominal(h.get(), 0, h_nominal.data(), (int)h_nominal.size()), "getControl");
        dr_bm_diag d{};
        check(dr_bm_get_diag(h.get(), &d, 1), "getControl");
        mppiConfig.invTemperature = d.lambda;
        useNewPlan = d.use_new_plan != 0;
        epsilon = d.epsilon;
        epsilonPartial = d.epsilon_partial;
        certifiedLoss = d.certified_loss;
        chosenBranch = d.branch;
        if (mppiConfig.nVerifSamples > 0)
        {
            const int T = (int)failCount.size();
            std::vector<int32_t> fn(T), fo(T);
            check(dr_bm_get_fail_counts(h.get(), 0, fn.data(), fo.data(), T), "getControl");
            for (int t = 0; t < T; ++t)
            {
                failCountNew[t] = (unsigned)fn[t];
                failCountOld[t] = (unsigned)fo[t];
                failCount[t] = useNewPlan ? failCountNew[t] : failCountOld[t];
            }
        }
    }

    // the natural cubic spline basis B (T x M), mppi_controller.cu:56-141
    std::vector<float> buildSplineMatrix()
    {
        if (!h)
            throw std::runtime_error("buildSplineMatrix: controller was default-constructed");
        std::vector<float> B((size_t)mppiConfig.nTimesteps * mppiConfig.nKnots);
        check(dr_bm_get_basis(h.get(), 0, B.data(), (int)B.size()), "buildSplineMatrix");
        return B;
    }

    MPPIConfig mppiConfig{};
    std::vector<float> h_nominal;   // (nBranchPlans, T, dim): dense plans after the last step, the chosen branch shifted
    std::vector<unsigned> failCountOld; // (T) verification failures of the previous plan
    std::vector<unsigned> failCountNew; // (T) verification failures of the new plan
    std::vector<unsigned> failCount;    // (T) those of the executed plan
    double epsilonPartial = 0.0;
    double epsilon = 0.0;
    double certifiedLoss = 0.0;
    bool useNewPlan = true;
    float min_nu = 0.0005f; // temperature rule thresholds, as a proportion of nSamples
    float max_nu = 0.001f;
    int chosenBranch = 0;   // 0 nominal, theta + 1 committed

  private:
    int n_models = 0;
    std::unique_ptr<dr_bm_engine, BmDeleter> h;
};

// The reference's PRMPPIController on StratRace2D: P = int(mc.P) opponent
// models drawn from h_belief every step and averaged uniformly; both plans,
// the candidate and safety choice and the three temperatures.
template <class Ty>
class PRMPPIController : public Controller<Ty>
{
  public:
    using PRMPPIConfig = typename Ty::PRMPPIConfig;

    PRMPPIController() = default;

    PRMPPIController(const typename Ty::EnvironmentConfig& c, const PRMPPIConfig& mc, int s,
                     std::optional<std::vector<float>> nominal = std::nullopt)
        : mppiConfig(mc)
    {
        const dr_sr_env env = Ty::abi_env(c);
        n_models = env.n_models;
        if (n_models < 1 || n_models > Ty::N_TRUE_MODELS || n_models > DR_SR_MAX_MODELS)
            throw std::runtime_error("Invalid PRMPPI construction: the environment's model count does not fit N_TRUE_MODELS");
        const size_t n = (size_t)mc.nTimesteps * 2;
        if (nominal && nominal->size() != n)
            throw std::runtime_error("Invalid PRMPPI construction: expected nominal size " + std::to_string(n) + ", got " +
                                     std::to_string(nominal->size()));
        // PRMPPIConfig carries no outside cost (config.cpp:55-84); PR-MPPI's running cost has no such term
        dr_sr_cost cost{mc.collDistFactor, mc.boundaryCost, mc.boundaryThresholdFactor, 0.0f,
                        mc.winCost,        mc.oppWinCost,   mc.finalAdvWeight,          mc.finalOppAdvWeight};
        dr_prsr_instance in{};
        in.n_samples = mc.nSamples;
        in.n_timesteps = mc.nTimesteps;
        in.sampled = 1;
        in.n_draws = (int)mc.P; // config.cpp:80 reads the float field as an int
        in.seed = (uint64_t)(uint32_t)s;
        in.lambda = mc.invTemperature;
        in.sigma = mc.samplingNoise;
        in.safety_weight = mc.safetyWeight;
        in.adapt_lambda = 1;
        float belief[DR_SR_MAX_MODELS] = {};
        Ty::init_belief(c, belief);
        for (int th = 0; th < n_models; ++th)
            in.belief[th] = belief[th];
        in.true_theta = 0; // the world is the engine's; this handle never advances it
        dr_prengine* raw = nullptr;
        check(dr_prsr_create(&env, &cost, &in, 1, &raw), "PRMPPIController");
        h.reset(raw);
        this->h_belief.fill(0.0f);
        for (int th = 0; th < n_models; ++th)
            this->h_belief[th] = belief[th];
        h_nom_nominal = nominal ? *nominal : std::vector<float>(n, 0.0f);
        h_rob_nominal = h_nom_nominal;
        check(dr_pr_set_nominals(h.get(), 0, h_nom_nominal.data(), h_rob_nominal.data(), (int)n), "PRMPPIController");
    }

    PRMPPIController(const PRMPPIController&) = delete;
    PRMPPIController& operator=(const PRMPPIController&) = delete;

    void getControl(const typename Ty::SimState& state, float* outAction) override
    {
        if (!h)
            throw std::runtime_error("getControl: controller was default-constructed");
        const dr_sr_state st = Ty::abi_state(state);
        check(dr_pr_set_belief(h.get(), 0, this->h_belief.data(), n_models), "getControl");
        check(dr_pr_set_sr_states(h.get(), &st, 1), "getControl");
        check(dr_pr_step(h.get(), 1, 0, outAction), "getControl");
        check(dr_pr_get_nominals(h.get(), 0, h_nom_nominal.data(), h_rob_nominal.data(), (int)h_nom_nominal.size()), "getControl");
        dr_pr_diag d{};
        check(dr_pr_get_diag(h.get(), &d, 1), "getControl");
        useNomPlan = d.use_nom_plan != 0;
        resetNom = d.reset_nom != 0;
        mppiConfig.invTemperature = d.lambda[0];
        for (int j = 0; j < 3; ++j)
            temperatures[j] = d.lambda[j];
    }

    PRMPPIConfig mppiConfig{};
    std::vector<float> h_nom_nominal; // (T, dim)
    std::vector<float> h_rob_nominal; // (T, dim)
    bool resetNom = false;
    bool useNomPlan = true;
    float min_nu = 0.0005f;
    float max_nu = 0.001f;
    std::array<float, 3> temperatures{}; // nominal/full, robust/full, robust/safety

  private:
    int n_models = 0;
    std::unique_ptr<dr_prengine, PrDeleter> h;
};

// A types bundle over this library's own structs, with configs that carry the
// reference's field names (config.h:116-192); a program that has the
// reference's own headers writes the same few lines over those types instead.
struct StratRaceTypes
{
    static constexpr int N_TRUE_MODELS = DR_SR_MAX_MODELS;
    class Engine; // the host's engine; only pointed to
    using SimState = dr_sr_state;
    struct EnvironmentConfig
    {
        dr_sr_env env{};
        std::array<float, DR_SR_MAX_MODELS> initBelief{};
    };
    struct MPPIConfig
    {
        int nSamples = 100;
        int nTimesteps = 20;
        int nKnots = 0;
        std::array<int, DR_MAX_KNOTS> knots{};
        float invTemperature = 10.0f;
        float samplingNoise = 0.1f;
        float gateTraversalMargin = 0.95f;
        float collDistFactor = 1.0f;
        float boundaryCost = 10.0f;
        float boundaryThresholdFactor = 1.5f;
        float outsideCost = 1000.0f;
        float winCost = 1000.0f;
        float oppWinCost = 1000.0f;
        float finalAdvWeight = 10.0f;
        float finalOppAdvWeight = 5.0f;
        float minConfidence = 0.9f;
        int nVerifSamples = 10000;
        float beta = 1e-6f;
        int verifHorizon = 40;
        float maxVerifEps = 0.01f;
    };
    struct PRMPPIConfig
    {
        int nSamples = 100;
        int nTimesteps = 20;
        float invTemperature = 10.0f;
        float samplingNoise = 0.1f;
        float gateTraversalMargin = 0.95f;
        float collDistFactor = 1.0f;
        float boundaryCost = 10.0f;
        float boundaryThresholdFactor = 1.5f;
        float winCost = 1000.0f;
        float oppWinCost = 1000.0f;
        float finalAdvWeight = 10.0f;
        float finalOppAdvWeight = 5.0f;
        float safetyWeight = 10000.0f;
        float minSafeDist = 0.0f;
        float delta = 0.1f;
        float P = 10;
    };
    static dr_sr_env abi_env(const EnvironmentConfig& c) { return c.env; }
    static dr_sr_state abi_state(const SimState& s) { return s; }
    static void init_belief(const EnvironmentConfig& c, float* out)
    {
        for (int i = 0; i < DR_SR_MAX_MODELS; ++i)
            out[i] = c.initBelief[i];
    }
};

} // namespace reference
} // namespace drone_race
