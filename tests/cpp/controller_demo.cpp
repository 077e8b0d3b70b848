// Drives MPPIController (include/drone_race/controllers.hpp) the way the
// reference's SimulationEngine drives its controller (engine.cpp:314-363):
// getControl, apply the action, repeat.  Prints the actions and the final
// nominal; the GPU test compares them with the batch path.
#include <cmath>
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
        // PR-MPPI over two opponent-velocity branches
        PRMPPIConfig pc;
        pc.nSamples = 1024;
        pc.nTimesteps = 30;
        pc.shiftFill = {9.81f, 0.0f, 0.0f, 0.0f};
        pc.belief = {0.3f, 0.7f};
        pc.velShift.assign(2, {});
        pc.velShift[0][0] = 1.0f;
        pc.velShift[1][1] = -1.0f;
        std::vector<float> pnom((size_t)pc.nTimesteps * ACTION_DIM);
        for (int t = 0; t < pc.nTimesteps; ++t)
            pnom[(size_t)t * 4] = 9.81f;
        PRMPPIController pr(env, pc, 11, pnom);
        for (int i = 0; i < steps; ++i)
        {
            float a[4];
            pr.getControl(s, a);
            std::printf("pr_action %d %.9g %.9g %.9g %.9g use_nom %d reset %d\n", i, a[0], a[1], a[2], a[3], (int)pr.useNomPlan,
                        (int)pr.resetNom);
        }
        // StratRace2D: analytic circle of radius 6 (p, unit tangent, curvature 1/6)
        static stratrace::EnvironmentConfig se{};
        se.dt = 0.1f;
        se.drone_radius = 0.4f;
        se.max_speed[0] = 3.0f;
        se.max_speed[1] = 2.0f;
        se.max_accel[0] = 2.0f;
        se.max_accel[1] = 3.0f;
        se.track_width = 2.0f;
        se.n_win_laps = 1;
        se.kp = 10.0f;
        se.kv = 5.0f;
        se.max_opp_lat_dist_fact = 1.9f;
        se.track_length = (float)(2.0 * M_PI * 6.0);
        se.n_models = 2;
        se.opp[0] = dr_sr_opp_config{0.0f, 1.0f / (1.6f * 1.6f), 50.0f, 1.0f, 0.1f};
        se.opp[1] = dr_sr_opp_config{0.0f, 1.0f / (1.6f * 1.6f), 0.0f, 1.0f, 0.1f};
        for (int i = 0; i < DR_SR_TRACK_SAMPLES; ++i)
        {
            const double th = 2.0 * M_PI * i / DR_SR_TRACK_SAMPLES;
            se.p_grid[i][0] = (float)(6.0 * std::cos(th));
            se.p_grid[i][1] = (float)(6.0 * std::sin(th));
            se.t_grid[i][0] = (float)(-std::sin(th));
            se.t_grid[i][1] = (float)std::cos(th);
            se.kappa_grid[i] = 1.0f / 6.0f;
        }
        stratrace::MPPIConfig sc;
        sc.nSamples = 4096;
        sc.nTimesteps = 40;
        stratrace::MPPIController sr(se, sc, 5, 1);
        stratrace::SimState ss{};
        const int i0 = 26, i1 = 102;
        ss.pos[0] = se.p_grid[i0][0];
        ss.pos[1] = se.p_grid[i0][1];
        ss.pos[2] = se.p_grid[i1][0];
        ss.pos[3] = se.p_grid[i1][1];
        ss.s[0] = (float)i0 / DR_SR_TRACK_SAMPLES * se.track_length;
        ss.s[1] = (float)i1 / DR_SR_TRACK_SAMPLES * se.track_length;
        for (int i = 0; i < steps; ++i)
        {
            float a[2];
            sr.getControl(ss, a);
            std::printf("sr_action %d %.9g %.9g\n", i, a[0], a[1]);
        }
        // the reference's default controller: branching over the models with spline plans
        stratrace::BranchingMPPIConfig bc;
        bc.nSamples = 2048;
        bc.nTimesteps = 30;
        bc.nKnots = 8;
        bc.minConfidence = 0.6f;
        bc.nVerifSamples = 4096;
        bc.verifHorizon = 20;
        bc.beta = 1e-5f;
        bc.maxVerifEps = 0.01f;
        stratrace::BranchingMPPIController bm(se, bc, 9, 1);
        const float beliefs[3][2] = {{0.5f, 0.5f}, {0.3f, 0.7f}, {0.9f, 0.1f}};
        for (int i = 0; i < steps; ++i)
        {
            bm.h_belief[0] = beliefs[i % 3][0];
            bm.h_belief[1] = beliefs[i % 3][1];
            float a[2];
            bm.getControl(ss, a);
            unsigned fails = 0;
            for (unsigned f : bm.failCount)
                fails += f;
            std::printf("bm_action %d %.9g %.9g branch %d lambda %.9g use_new %d cert %.17g fails %u\n", i, a[0], a[1], bm.chosenBranch,
                        bm.mppiConfig.invTemperature, (int)bm.useNewPlan, bm.certifiedLoss, fails);
        }
        // the same controller on HiddenObs (hiddenObsEnv's course)
        hiddenobs::EnvironmentConfig he{};
        he.arena_max[0] = 13.0f;
        he.arena_max[1] = 9.5f;
        he.dt = 0.1f;
        he.drone_radius = 0.3f;
        he.max_speed = 2.0f;
        he.max_accel = 3.0f;
        he.n_win_laps = 1;
        const float gc[3][2] = {{12.0f, 7.0f}, {1.0f, 3.0f}, {5.0f, 7.0f}}, gv[3][2] = {{1, 0}, {0, 1}, {1, 0}};
        for (int g = 0; g < 3; ++g)
            for (int d = 0; d < 2; ++d)
            {
                he.gate_centers[g][d] = gc[g][d];
                he.gate_vectors[g][d] = gv[g][d];
                he.gate_radius[g] = 1.0f;
            }
        const float rect[3][4] = {{2, 0, 12, 3}, {5, 3, 12, 4.5f}, {6.5f, 6.5f, 10.5f, 7.5f}};
        for (int r = 0; r < 3; ++r)
            for (int j = 0; j < 4; ++j)
                he.rect_obstacles[r][j] = rect[r][j];
        he.hidden_centers[0][0] = 9.0f;
        he.hidden_centers[0][1] = 5.5f;
        he.hidden_centers[1][0] = 9.0f;
        he.hidden_centers[1][1] = 8.5f;
        he.hidden_radius[0] = he.hidden_radius[1] = 0.8f;
        const float lim[4] = {5, 6.5f, 10.5f, 12}, lam[4] = {3, -11, 20, 6};
        for (int j = 0; j < 4; ++j)
        {
            he.dbl_hp_limits[j] = lim[j];
            he.dbl_hp_lambdas[j] = lam[j];
        }
        he.annulus_center[0] = 5.0f;
        he.annulus_center[1] = 3.0f;
        he.annulus_radius[0] = 3.0f;
        he.annulus_radius[1] = 5.0f;
        hiddenobs::BranchingMPPIConfig hc;
        hc.nSamples = 2048;
        hc.nTimesteps = 40;
        hc.nKnots = 10;
        hiddenobs::BranchingMPPIController ho(he, hc, 13, 0);
        hiddenobs::SimState hs{};
        hs.pos[0] = 1.0f;
        hs.pos[1] = 4.0f;
        hs.vel[0] = 0.3f;
        hs.vel[1] = 1.0f;
        for (int i = 0; i < steps; ++i)
        {
            float a[2];
            ho.getControl(hs, a);
            std::printf("ho_action %d %.9g %.9g branch %d lambda %.9g\n", i, a[0], a[1], ho.chosenBranch, ho.mppiConfig.invTemperature);
        }
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
    return 0;
}
