// Drives MPPIController (include/drone_race/controllers.hpp) the way the
// reference's SimulationEngine drives its controller (engine.cpp:314-363):
// getControl, apply the action, repeat.  Prints the actions and the final
// nominal; the GPU test compares them with the batch path.
#include <cstdio>
#include <cstdlib>

#include "drone_race/controllers.hpp"

int main(int argc, char** argv)
{
    using namespace drone_race;
    const int steps = argc > 1 ? std::atoi(argv[1]) : 3;
    EnvironmentConfig env;
    env.quad = dr_quad_params{1.0f, 9.81f, 0.1f, 20.0f, 0.0f, 30.0f, 6.0f, 0.15f, 0.02f};
    env.cost = dr_cost_cfg{0.95f, 1.0f, 10.0f, 1.5f, 1000.0f, 1000.0f, 1000.0f, 10.0f, 0.0f, 2.0f, 50.0f, 0.0f};
    env.track.n_gates = 4;
    env.track.n_win_laps = 100;
    env.track.ground_z = 0.0f;
    const float cx[4] = {10.0f, 0.0f, -10.0f, 0.0f}, cy[4] = {0.0f, 10.0f, 0.0f, -10.0f};
    const float nx[4] = {0.0f, -1.0f, 0.0f, 1.0f}, ny[4] = {1.0f, 0.0f, -1.0f, 0.0f};
    for (int g = 0; g < 4; ++g)
    {
        env.track.gates[g] = dr_gate{{cx[g], cy[g], 4.0f}, {nx[g], ny[g], 0.0f}, 1.0f, 0.0f};
    }
    env.maxOpponents = 1;
    MPPIConfig mc;
    mc.nSamples = 2048;
    mc.nTimesteps = 40;
    mc.shiftFill = {9.81f, 0.0f, 0.0f, 0.0f};
    std::vector<float> nominal((size_t)mc.nTimesteps * ACTION_DIM);
    for (int t = 0; t < mc.nTimesteps; ++t)
        nominal[(size_t)t * 4] = 9.81f;
    try
    {
        MPPIController ctrl(env, mc, 7, nominal);
        SimState s;
        s.ego = dr_drone_state{{7.0f, -7.0f, 4.0f}, {1.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, 3, 0, 0};
        s.nOpponents = 1;
        s.opponents[0] = dr_opponent{{8.0f, -6.0f, 4.0f}, {1.0f, 1.0f, 0.0f}};
        for (int i = 0; i < steps; ++i)
        {
            float a[4];
            ctrl.getControl(s, a);
            std::printf("action %d %.9g %.9g %.9g %.9g\n", i, a[0], a[1], a[2], a[3]);
        }
        std::printf("nominal0 %.9g %.9g %.9g %.9g\n", ctrl.h_nominal[0], ctrl.h_nominal[1], ctrl.h_nominal[2], ctrl.h_nominal[3]);
        std::printf("lambda %.9g nu %.9g\n", ctrl.mppiConfig.invTemperature, ctrl.lastNu);
        // wrong nominal size throws, as the reference constructor does
        try
        {
            MPPIController bad(env, mc, 7, std::vector<float>(3));
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
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
    return 0;
}
