// C++ host interface with the shape of the reference controllers
// (include/controllers.h:12-111 of pablo-gr-rey/drone-race), implemented
// over the C ABI of dr_api.h.  Header-only; link with libdr_amd.so.
//
//   Controller::getControl(const SimState&, float* outAction)  controllers.h:17
//   Controller::engine, Controller::h_belief                    controllers.h:19-20
//   MPPIController(EnvironmentConfig, MPPIConfig, seed, nominal) controllers.h:29
//   MPPIController::mppiConfig, h_nominal                       controllers.h:35-37
//
// Errors from the C ABI are rethrown as std::runtime_error, as the
// reference's constructors and checks throw (mppi_controller.cu:27-29).
// The environment is the quadrotor gate race (13-state drone, constant-
// velocity opponents); one opponent model, so the belief has one entry.
#pragma once

#include <array>
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
        check(dr_mppi_create(&c.quad, &c.cost, &c.track, &cfg, (uint64_t)(uint32_t)seed, &h), "MPPIController");
        if (nominal)
        {
            if ((int)nominal->size() != mc.nTimesteps * ACTION_DIM)
            {
                dr_mppi_destroy(h);
                throw std::runtime_error("Invalid MPPI construction: expected nominal size " +
                                         std::to_string(mc.nTimesteps * ACTION_DIM) + ", got " + std::to_string(nominal->size()));
            }
            h_nominal = *nominal;
        }
        else
            h_nominal.assign((size_t)mc.nTimesteps * ACTION_DIM, 0.0f);
        check(dr_mppi_set_nominal(h, h_nominal.data(), (int)h_nominal.size()), "MPPIController");
    }

    ~MPPIController() override { dr_mppi_destroy(h); }
    MPPIController(const MPPIController&) = delete;
    MPPIController& operator=(const MPPIController&) = delete;

    void getControl(const SimState& state, float* outAction) override
    {
        check(dr_mppi_step(h, &state.ego, state.opponents.data(), state.nOpponents, outAction), "getControl");
        check(dr_mppi_get_nominal(h, h_nominal.data(), (int)h_nominal.size()), "getControl");
        dr_diag d{};
        check(dr_mppi_get_diag(h, &d), "getControl");
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
    dr_mppi* h = nullptr;
};

} // namespace drone_race
