// A host engine holding its controller the way the reference's
// SimulationEngine does (src/engine.cpp:14-363): std::unique_ptr<Controller>
// built by std::visit over the controller config (engine.cpp:29-47), the
// controller pointing back to the engine, getControl once per step, and the
// controller's fields read through dynamic_cast<MPPIController*> /
// dynamic_cast<PRMPPIController*> as sendState and dynStep read them
// (engine.cpp:73-166, 169-257, 300-306).  The world itself is not simulated
// here: the states and beliefs of each step are read from the input file (the
// test records them from the Python handle's closed loop), and the engine
// writes the belief into controller->h_belief as dynStep does (engine.cpp:332).
//
// Input (text): "env <file with the raw dr_sr_env bytes>", "kind mppi|prmppi",
// "seed <int>", config lines "<MPPIConfig / PRMPPIConfig field> <value>"
// ("knots" takes nKnots integers), then per step
// "step <belief x N_TRUE_MODELS> <pos x4> <vel x4> <S x2> <latDist x2> <laps x2>".
// Output: per step the action and the fields the reference engine reads.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <variant>
#include <vector>

#include "drone_race/reference_controllers.hpp"

class SimulationEngine;

struct Types : drone_race::reference::StratRaceTypes
{
    using Engine = ::SimulationEngine;
};

using Controller = drone_race::reference::Controller<Types>;
using MPPIController = drone_race::reference::MPPIController<Types>;
using PRMPPIController = drone_race::reference::PRMPPIController<Types>;
using EnvironmentConfig = Types::EnvironmentConfig;
using MPPIConfig = Types::MPPIConfig;
using PRMPPIConfig = Types::PRMPPIConfig;
using SimState = Types::SimState;
using AnyControllerConfig = std::variant<MPPIConfig, PRMPPIConfig>;
constexpr int N_TRUE_MODELS = Types::N_TRUE_MODELS;

class SimulationEngine
{
  public:
    SimulationEngine(const EnvironmentConfig& config, const AnyControllerConfig& contConfig, int seed) : envConfig(config)
    {
        controller = std::visit(
            [&](auto&& concrete) -> std::unique_ptr<Controller>
            {
                using T = std::decay_t<decltype(concrete)>;
                if constexpr (std::is_same_v<T, MPPIConfig>)
                    return std::make_unique<MPPIController>(envConfig, concrete, seed);
                else
                    return std::make_unique<PRMPPIController>(envConfig, concrete, seed);
            },
            contConfig);
        controller->engine = this;
    }

    // getControl, then the fields sendState reads, then the world step's belief
    void step(int k, const SimState& state, const std::array<float, N_TRUE_MODELS>& belief_after)
    {
        float action[2];
        controller->getControl(state, action);
        std::printf("step %d action %.9g %.9g\n", k, action[0], action[1]);
        if (auto mppi = dynamic_cast<MPPIController*>(controller.get()))
        {
            unsigned fails = 0;
            for (unsigned f : mppi->failCount)
                fails += f;
            std::printf("step %d mppi fail_count %u epsilon %.17g epsilon_partial %.17g use_new_plan %d certified_loss %.17g "
                        "min_confidence %.9g n_timesteps %d lambda %.9g nominal0 %.9g %.9g belief %.9g %.9g back %d\n",
                        k, fails, mppi->epsilon, mppi->epsilonPartial, (int)mppi->useNewPlan, mppi->certifiedLoss,
                        mppi->mppiConfig.minConfidence, mppi->mppiConfig.nTimesteps, mppi->mppiConfig.invTemperature,
                        mppi->h_nominal[0], mppi->h_nominal[1], mppi->h_belief[0], mppi->h_belief[1], (int)(mppi->engine == this));
        }
        else if (auto pr = dynamic_cast<PRMPPIController*>(controller.get()))
        {
            std::printf("step %d prmppi use_nom_plan %d reset_nom %d n_timesteps %d safety_weight %.9g nom0 %.9g %.9g "
                        "rob0 %.9g %.9g belief %.9g %.9g back %d\n",
                        k, (int)pr->useNomPlan, (int)pr->resetNom, pr->mppiConfig.nTimesteps, pr->mppiConfig.safetyWeight,
                        pr->h_nom_nominal[0], pr->h_nom_nominal[1], pr->h_rob_nominal[0], pr->h_rob_nominal[1], pr->h_belief[0],
                        pr->h_belief[1], (int)(pr->engine == this));
        }
        else
            throw std::runtime_error("Ill-formed controller type in sendState");
        controller->h_belief = belief_after; // dynStep's belief update, engine.cpp:332
    }

    EnvironmentConfig envConfig;
    std::unique_ptr<Controller> controller;
};

int main(int argc, char** argv)
{
    if (argc < 2)
    {
        std::fprintf(stderr, "usage: %s <input>\n", argv[0]);
        return 2;
    }
    std::ifstream in(argv[1]);
    std::string line, kind;
    EnvironmentConfig env;
    MPPIConfig mc;
    PRMPPIConfig pc;
    int seed = 0;
    std::vector<std::pair<std::array<float, N_TRUE_MODELS>, SimState>> steps;
    while (std::getline(in, line))
    {
        std::istringstream ls(line);
        std::string key;
        ls >> key;
        if (key == "env")
        {
            std::string path;
            ls >> path;
            std::ifstream f(path, std::ios::binary);
            f.read(reinterpret_cast<char*>(&env.env), sizeof(dr_sr_env));
            if (f.gcount() != (std::streamsize)sizeof(dr_sr_env))
                throw std::runtime_error("short environment file");
        }
        else if (key == "belief0")
            for (int i = 0; i < N_TRUE_MODELS; ++i)
                ls >> env.initBelief[i];
        else if (key == "kind")
            ls >> kind;
        else if (key == "seed")
            ls >> seed;
        else if (key == "knots")
            for (int m = 0; m < mc.nKnots; ++m)
                ls >> mc.knots[m];
        else if (key == "step")
        {
            std::array<float, N_TRUE_MODELS> b{};
            SimState s{};
            for (int i = 0; i < N_TRUE_MODELS; ++i)
                ls >> b[i];
            for (float& v : s.pos)
                ls >> v;
            for (float& v : s.vel)
                ls >> v;
            for (float& v : s.s)
                ls >> v;
            for (float& v : s.lat_dist)
                ls >> v;
            for (int& v : s.laps)
                ls >> v;
            steps.emplace_back(b, s);
        }
        else if (!key.empty())
        {
            double v;
            ls >> v;
#define SET(cfg, f)                                                                                                                     \
    if (key == #f)                                                                                                                     \
        cfg.f = (decltype(cfg.f))v;
            SET(mc, nSamples) SET(mc, nTimesteps) SET(mc, nKnots) SET(mc, invTemperature) SET(mc, samplingNoise)
            SET(mc, collDistFactor) SET(mc, boundaryCost) SET(mc, boundaryThresholdFactor) SET(mc, outsideCost) SET(mc, winCost)
            SET(mc, oppWinCost) SET(mc, finalAdvWeight) SET(mc, finalOppAdvWeight) SET(mc, minConfidence) SET(mc, nVerifSamples)
            SET(mc, beta) SET(mc, verifHorizon) SET(mc, maxVerifEps)
            SET(pc, nSamples) SET(pc, nTimesteps) SET(pc, invTemperature) SET(pc, samplingNoise) SET(pc, collDistFactor)
            SET(pc, boundaryCost) SET(pc, boundaryThresholdFactor) SET(pc, winCost) SET(pc, oppWinCost) SET(pc, finalAdvWeight)
            SET(pc, finalOppAdvWeight) SET(pc, safetyWeight) SET(pc, delta) SET(pc, P)
#undef SET
        }
    }
    try
    {
        AnyControllerConfig cfg = kind == "prmppi" ? AnyControllerConfig(pc) : AnyControllerConfig(mc);
        SimulationEngine eng(env, cfg, seed);
        for (size_t k = 0; k < steps.size(); ++k)
        {
            // the belief the controller holds at this step was written by the previous world step
            if (k == 0)
                eng.controller->h_belief = steps[0].first;
            eng.step((int)k, steps[k].second, k + 1 < steps.size() ? steps[k + 1].first : steps[k].first);
        }
        // a nominal of the wrong size throws from the constructor, as in the reference
        try
        {
            if (kind == "prmppi")
                PRMPPIController bad(env, pc, seed, std::vector<float>(3));
            else
                MPPIController bad(env, mc, seed, std::vector<float>(3));
            std::printf("no-throw\n");
            return 1;
        }
        catch (const std::runtime_error& e)
        {
            std::printf("throws: %s\n", e.what());
        }
    }
    catch (const std::exception& e)
    {
        std::printf("error: %s\n", e.what());
        return 1;
    }
    return 0;
}
