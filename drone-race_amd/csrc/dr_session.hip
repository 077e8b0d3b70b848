// Wire protocol and simulation session of the reference's front-end for the
// StratRace2D build (host code).
//
//   byte codec                 include/protocol.h:26-142, src/protocol.cpp, python-src/protocol.py:13-116
//   environment header         src/environments/env_stratrace.cpp:127-226 (unpackEnvConfig)
//   initial state              src/environments/env_stratrace.cpp:98-118 (unpackSimState)
//   controller header          src/config.cpp:5-115 (loadControllerConfig)
//   STATE / EVENT / DONE       src/engine.cpp:55-283, env_stratrace.cpp:48-85
//   run loop                   src/engine.cpp:314-352 (SimulationEngine::run), main.cpp:9-44
//
// The controller is one of this library's engines (dr_bm_* for the MPPI
// header, dr_prsr_create + dr_pr_* for the PR-MPPI header); getControl and
// dynStep are one engine step with the world advanced on the device.  The
// display predictions of sendState run here on the host through the same
// StratRace2D step the kernels use (dr_sr.h), without noise.
//
// Differences from the reference, by design: the world's noise comes from
// the engine's Philox stream (the reference draws it from std::mt19937 with
// std::normal_distribution); with nVerifSamples = 0 no verification runs and
// the new plan is always executed (STATE frames carry failCount 0, epsilon 0,
// epsilonPartial 0, useNewPlan 1, certifiedLoss 0); nothing is printed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "dr_host.h"
#include "dr_sr.h"

using namespace dr;

namespace {

constexpr int DEFAULT_MAX_STEPS = 500; // main.cpp:11

// Little-endian frame writer (Writer, protocol.h:101-131).
struct Frame
{
    std::vector<uint8_t> b;
    void i32(int32_t v)
    {
        const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
        b.insert(b.end(), p, p + 4);
    }
    void f32(float v)
    {
        const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
        b.insert(b.end(), p, p + 4);
    }
    void arr(const float* v, size_t n)
    {
        i32((int32_t)n);
        for (size_t i = 0; i < n; ++i)
            f32(v[i]);
    }
    // integer arrays travel as float32 (pushIntArray, protocol.h:134-142)
    void iarr(const int32_t* v, size_t n)
    {
        i32((int32_t)n);
        for (size_t i = 0; i < n; ++i)
            f32((float)v[i]);
    }
};

// Bounds-checked frame reader (Reader, protocol.h:26-99); the first
// failure is kept and every later read returns zero.
struct Cursor
{
    const uint8_t* p;
    size_t n, off = 0;
    std::string err;
    Cursor(const void* data, size_t len) : p(static_cast<const uint8_t*>(data)), n(len) {}
    bool ok() const { return err.empty(); }
    void fail(const std::string& m)
    {
        if (err.empty())
            err = m;
    }
    bool take(void* dst, size_t k)
    {
        if (!ok())
            return false;
        if (!p || off + k > n)
        {
            fail("failed to read scalar from buffer (underflow)");
            return false;
        }
        std::memcpy(dst, p + off, k);
        off += k;
        return true;
    }
    int32_t i32()
    {
        int32_t v = 0;
        take(&v, 4);
        return v;
    }
    float f32()
    {
        float v = 0.0f;
        take(&v, 4);
        return v;
    }
    // An array of exactly `want` values (readArray, protocol.h:53-75).
    void arr(float* out, size_t want, const char* what)
    {
        const int32_t len = i32();
        if (!ok())
            return;
        if (len < 0 || (size_t)len != want)
        {
            fail(std::string("wrong size when reading array ") + what + ": read length " + std::to_string(len) + ", allocated length " +
                 std::to_string(want));
            return;
        }
        for (size_t i = 0; i < want; ++i)
            out[i] = f32();
    }
    // An array of at most `cap` values; returns its length.
    int arr_upto(float* out, size_t cap, const char* what)
    {
        const int32_t len = i32();
        if (!ok())
            return 0;
        if (len < 0 || (size_t)len > cap)
        {
            fail(std::string("wrong size when reading array ") + what + ": read length " + std::to_string(len) + ", allocated length " +
                 std::to_string(cap));
            return 0;
        }
        for (int32_t i = 0; i < len; ++i)
            out[i] = f32();
        return len;
    }
    void header(const char* what)
    {
        const int32_t kind = i32();
        if (ok() && kind != DR_MSG_HEADER)
            fail(std::string("expected header message type for ") + what + " (type 0) but got type " + std::to_string(kind));
    }
    void finished()
    {
        if (ok() && off != n)
            fail("header not fully parsed: " + std::to_string(n - off) + " trailing bytes");
    }
    int result() const { return ok() ? DR_OK : host_fail(DR_EINVAL, err); }
};

void push_state(Frame& f, const dr_sr_state& s)
{
    f.arr(s.pos, 4);
    f.arr(s.vel, 4);
    f.arr(s.s, 2);
    f.arr(s.lat_dist, 2);
    f.iarr(s.laps, 2);
}

sr::Tables tables(const dr_sr_env& E)
{
    return sr::Tables{&E.p_grid[0][0], &E.t_grid[0][0], E.kappa_grid, E.track_length};
}

// pushAddInfo (env_stratrace.cpp:61-70): the opponent's target lateral
// offset at the state before the step, under the world's model.
float lat_target(const dr_sr_env& E, const dr_sr_state& prev, int theta)
{
    const sr::Tables tb = tables(E);
    const sr::St s = sr::state_from_abi(prev);
    const sr::OppGeom g = sr::opp_geom(tb, E, s);
    float lat, speed;
    sr::opp_target(E, E.opp[theta], s, g, &lat, &speed);
    return lat;
}

// One display prediction (engine.cpp:96-166 with branching, :179-255
// without): from x0 under model theta, no noise, the opponent's win not a
// loss.  plans: [NB][T][2] with branching, one [T][2] plan without.
struct Prediction
{
    int32_t init_pred = 0, time = -1, pred = 0, stop_reason = DR_EVT_TRUNCATED, stop_time = -1;
    std::vector<float> full_pos, actions;
};

Prediction predict(const dr_sr_env& E, const dr_sr_state& x0, int theta, const float* plans, int T, bool branching, const float* belief,
                   float min_conf)
{
    Prediction r;
    r.full_pos.assign((size_t)T * 4, 0.0f);
    r.actions.assign((size_t)T * 2, 0.0f);
    const sr::Tables tb = tables(E);
    sr::St s = sr::state_from_abi(x0);
    float b[DR_SR_MAX_MODELS] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (branching)
    {
        for (int q = 0; q < E.n_models; ++q)
            b[q] = belief[q];
        r.pred = sr::find_confident(b, E.n_models, min_conf);
        r.time = r.pred == 0 ? -1 : 0;
    }
    r.init_pred = r.pred;
    sr::NormalSeq rng = sr::normal_seq(0u, 0u, 0u, 0u, 0u); // no draws: noise is off
    int origin = 0, t = 0;
    for (; t < T; ++t)
    {
        const float* u = branching ? plans + ((size_t)r.pred * T + (t - origin)) * 2 : plans + (size_t)t * 2;
        r.actions[t * 2] = u[0];
        r.actions[t * 2 + 1] = u[1];
        const float ego[2] = {u[0], u[1]};
        const int term = sr::env_step_t<true>(tb, E, theta, ego, s, rng, false, b, nullptr, branching, false);
        if (branching && r.pred == 0)
        {
            const int np = sr::find_confident(b, E.n_models, min_conf);
            if (np != 0)
            {
                r.pred = np;
                r.time = t + 1;
                origin = t + 1;
            }
        }
        for (int i = 0; i < 4; ++i)
            r.full_pos[t * 4 + i] = s.pos[i];
        if (term != sr::SR_NONE)
        {
            r.stop_reason = term - 1; // TerminalType -> EventType (engine.cpp:354-370)
            r.stop_time = t;
            break;
        }
    }
    for (int q = t + 1; q < T; ++q) // the stop state repeated (engine.cpp:130-131)
        for (int i = 0; i < 4; ++i)
            r.full_pos[q * 4 + i] = s.pos[i];
    return r;
}

} // namespace

struct dr_session
{
    dr_sr_env env{};
    dr_wire_env_info info{};
    dr_wire_controller cont{};
    dr_sr_state state{};
    int max_steps = DEFAULT_MAX_STEPS;
    int step = 0;
    bool started = false, finished = false;
    dr_bm_engine* bm = nullptr;
    dr_prengine* pr = nullptr;
    std::deque<std::vector<uint8_t>> queue;
    std::vector<uint8_t> current;

    ~dr_session()
    {
        if (bm)
            dr_bm_destroy(bm);
        if (pr)
            dr_pr_destroy(pr);
    }

    int NT() const { return env.n_models; }
    int T() const { return cont.n_timesteps; }

    int belief(float* b)
    {
        return bm ? dr_bm_get_belief(bm, 0, b, NT()) : dr_pr_get_belief(pr, 0, b, NT());
    }

    // sendState (engine.cpp:55-263) after the step `step` taken from prev.
    int state_frame(int stp, const float action[2], const dr_sr_state& prev)
    {
        if (!info.send_states)
            return DR_OK;
        Frame f;
        f.i32(DR_MSG_STATE);
        f.i32(stp);
        push_state(f, state);
        const float lt = lat_target(env, prev, info.true_theta);
        f.arr(&lt, 1);
        f.arr(action, 2);
        float b[DR_SR_MAX_MODELS] = {0, 0, 0, 0};
        if (int rc = belief(b))
            return rc;
        const int T = this->T(), NT = this->NT();
        if (bm)
        {
            dr_bm_diag d;
            if (int rc = dr_bm_get_diag(bm, &d, 1))
                return rc;
            f.arr(b, NT);
            f.i32((int32_t)d.fail_count);
            f.f32((float)d.epsilon);
            f.f32((float)d.epsilon_partial);
            f.i32(d.use_new_plan);
            f.f32((float)d.certified_loss);
            f.i32(NT);
            std::vector<float> nominal((size_t)(NT + 1) * T * 2);
            if (int rc = dr_bm_get_nominal(bm, 0, nominal.data(), (int)nominal.size()))
                return rc;
            for (int theta = 0; theta < NT; ++theta)
            {
                const Prediction p = predict(env, state, theta, nominal.data(), T, true, b, cont.min_confidence);
                f.iarr(&p.init_pred, 1);
                f.iarr(&p.time, 1);
                f.iarr(&p.pred, 1);
                f.arr(p.full_pos.data(), p.full_pos.size());
                f.arr(p.actions.data(), p.actions.size());
                f.i32(p.stop_reason);
                f.i32(p.stop_time);
            }
        }
        else
        {
            f.arr(prev.pos, 4);
            f.arr(b, NT);
            dr_pr_diag d;
            if (int rc = dr_pr_get_diag(pr, &d, 1))
                return rc;
            f.i32(d.use_nom_plan);
            f.i32(d.reset_nom);
            f.i32(NT + 1);
            std::vector<float> nom((size_t)T * 2), rob((size_t)T * 2);
            if (int rc = dr_pr_get_nominals(pr, 0, nom.data(), rob.data(), T * 2))
                return rc;
            for (int j = 0; j <= NT; ++j)
            {
                const int theta = std::min(j, NT - 1);
                const Prediction p = predict(env, prev, theta, j < NT ? nom.data() : rob.data(), T, false, b, 0.0f);
                f.arr(p.full_pos.data(), p.full_pos.size());
                f.arr(p.actions.data(), p.actions.size());
                f.i32(p.stop_reason);
                f.i32(p.stop_time);
            }
        }
        queue.push_back(std::move(f.b));
        return DR_OK;
    }

    void event_frame(int type)
    {
        Frame f;
        f.i32(DR_MSG_EVENT);
        f.i32(type);
        queue.push_back(std::move(f.b));
    }

    void done_frame()
    {
        Frame f;
        f.i32(DR_MSG_DONE);
        queue.push_back(std::move(f.b));
        finished = true;
    }

    // One iteration of SimulationEngine::run (engine.cpp:325-347).
    int advance()
    {
        if (!started)
        {
            started = true;
            const float zero[2] = {0.0f, 0.0f};
            return state_frame(0, zero, state);
        }
        step = step + 1;
        const dr_sr_state prev = state;
        float action[2] = {0.0f, 0.0f};
        int32_t ev = 0;
        if (bm)
        {
            if (int rc = dr_bm_step(bm, 1, 1, action))
                return rc;
            if (int rc = dr_bm_get_states(bm, &state, 1))
                return rc;
            if (int rc = dr_bm_get_events(bm, &ev, 1))
                return rc;
        }
        else
        {
            if (int rc = dr_pr_step(pr, 1, 1, action))
                return rc;
            if (int rc = dr_pr_get_sr_states(pr, &state, 1))
                return rc;
            if (int rc = dr_pr_get_events(pr, &ev, 1))
                return rc;
        }
        if (int rc = state_frame(step, action, prev))
            return rc;
        if (ev != sr::SR_NONE)
        {
            event_frame(ev - 1);
            done_frame();
        }
        else if (step >= max_steps)
        {
            event_frame(DR_EVT_TRUNCATED);
            done_frame();
        }
        return DR_OK;
    }
};

extern "C" {

int dr_wire_decode_env(const void* frame, size_t len, dr_sr_env* env, dr_wire_env_info* info)
{
    if (!frame || !env || !info)
        return host_fail(DR_EINVAL, "decode_env: NULL argument");
    dr_sr_env e;
    dr_wire_env_info w;
    std::memset(&e, 0, sizeof(e));
    std::memset(&w, 0, sizeof(w));
    Cursor c(frame, len);
    c.header("environment config");
    const int32_t envKind = c.i32();
    if (c.ok() && envKind != DR_ENV_STRATRACE)
        c.fail("environment kind mismatch: received kind " + std::to_string(envKind) + ", this engine is ENV_STRATRACE (2)");
    const int32_t nFactors = c.i32();
    if (c.ok() && nFactors != 1)
        c.fail("received " + std::to_string(nFactors) + " model factors; this engine supports 1");
    float sizes[8];
    const int nSizes = c.arr_upto(sizes, 8, "modelSizes");
    if (c.ok() && nSizes < 1)
        c.fail("modelSizes is empty");
    const int NT = c.ok() ? (int)sizes[0] : 0;
    if (c.ok() && (NT < 1 || NT > DR_SR_MAX_MODELS))
        c.fail("received model size " + std::to_string(NT) + "; this engine supports 1.." + std::to_string(DR_SR_MAX_MODELS));
    c.arr(w.arena_min, 2, "arenaMin");
    c.arr(w.arena_max, 2, "arenaMax");
    const int32_t nOpp = c.i32();
    const int32_t dim = c.i32();
    e.dt = c.f32();
    w.send_states = c.i32() != 0;
    if (c.ok() && nOpp != 1)
        c.fail("received config for " + std::to_string(nOpp) + " opponents; this engine supports 1");
    if (c.ok() && dim != 2)
        c.fail("received config for dimension " + std::to_string(dim) + "; this engine is 2-D");
    e.drone_radius = c.f32();
    e.pos_noise = c.f32();
    e.speed_noise = c.f32();
    e.action_noise = c.f32();
    c.arr(e.max_speed, 2, "maxSpeed");
    c.arr(e.max_accel, 2, "maxAccel");
    const int32_t nTrack = c.i32();
    if (c.ok() && nTrack != DR_SR_TRACK_SAMPLES)
        c.fail("received config for " + std::to_string(nTrack) + " track samples; this engine uses 512");
    e.track_width = c.f32();
    e.n_win_laps = c.i32();
    w.seed = c.i32();
    c.arr(w.init_belief, (size_t)std::max(NT, 0), "initBelief");
    for (int th = 0; th < NT && c.ok(); ++th)
    {
        c.arr(&e.opp[th].s1, 1, "s1");
        c.arr(&e.opp[th].s2, 1, "s2");
        c.arr(&e.opp[th].s3, 1, "s3");
        c.arr(&e.opp[th].speed_scale, 1, "speedScale");
        c.arr(&e.opp[th].action_noise, 1, "actionNoise");
    }
    w.true_theta = c.i32();
    e.kp = c.f32();
    e.kv = c.f32();
    e.max_opp_lat_dist_fact = c.f32();
    e.track_length = c.f32();
    c.arr(&e.p_grid[0][0], DR_SR_TRACK_SAMPLES * 2, "p_grid");
    std::vector<float> dp(DR_SR_TRACK_SAMPLES * 2);
    c.arr(dp.data(), dp.size(), "dp_grid"); // read and dropped: the step uses the unit tangents
    c.arr(&e.t_grid[0][0], DR_SR_TRACK_SAMPLES * 2, "t_grid");
    c.arr(e.kappa_grid, DR_SR_TRACK_SAMPLES, "kappa_grid");
    c.finished();
    if (int rc = c.result())
        return rc;
    e.n_models = NT;
    w.n_models = NT;
    if (w.true_theta < 0 || w.true_theta >= NT)
        return host_fail(DR_EINVAL, "trueTheta " + std::to_string(w.true_theta) + " out of range");
    if (int rc = check_sr_env(e))
        return rc;
    *env = e;
    *info = w;
    return DR_OK;
}

int dr_wire_decode_controller(const void* frame, size_t len, dr_wire_controller* cont)
{
    if (!frame || !cont)
        return host_fail(DR_EINVAL, "decode_controller: NULL argument");
    dr_wire_controller k;
    std::memset(&k, 0, sizeof(k));
    Cursor c(frame, len);
    c.header("controller config");
    k.kind = c.i32();
    if (c.ok() && k.kind == DR_CONT_MPPI)
    {
        k.n_samples = c.i32();
        k.n_timesteps = c.i32();
        k.n_knots = c.i32();
        if (c.ok() && (k.n_knots < 0 || k.n_knots > DR_MAX_KNOTS))
            c.fail("received invalid number of knots: got " + std::to_string(k.n_knots) + ", at most 60");
        float kn[DR_MAX_KNOTS];
        const int got = c.arr_upto(kn, DR_MAX_KNOTS, "knots");
        if (c.ok() && got != k.n_knots)
            c.fail("received invalid number of knots: got " + std::to_string(got) + " elements, but nKnots=" + std::to_string(k.n_knots));
        for (int m = 0; m < got; ++m)
            k.knots[m] = (int32_t)kn[m];
        k.use_splines = c.i32();
        if (c.ok() && k.use_splines != 1)
            c.fail("received useSplines=" + std::to_string(k.use_splines) + "; this engine runs spline plans (USE_SPLINES)");
        k.lambda = c.f32();
        k.sigma = c.f32();
        k.gate_traversal_margin = c.f32();
        k.cost.coll_dist_factor = c.f32();
        k.cost.boundary_cost = c.f32();
        k.cost.boundary_threshold_factor = c.f32();
        k.cost.outside_cost = c.f32();
        k.cost.win_cost = c.f32();
        k.cost.opp_win_cost = c.f32();
        k.cost.final_adv_weight = c.f32();
        k.cost.final_opp_adv_weight = c.f32();
        k.min_confidence = c.f32();
        k.n_verif_samples = c.i32();
        k.beta = c.f32();
        k.verif_horizon = c.i32();
        k.max_verif_eps = c.f32();
    }
    else if (c.ok() && k.kind == DR_CONT_PRMPPI)
    {
        k.n_samples = c.i32();
        k.n_timesteps = c.i32();
        k.lambda = c.f32();
        k.sigma = c.f32();
        k.gate_traversal_margin = c.f32();
        k.cost.coll_dist_factor = c.f32();
        k.cost.boundary_cost = c.f32();
        k.cost.boundary_threshold_factor = c.f32();
        k.cost.win_cost = c.f32();
        k.cost.opp_win_cost = c.f32();
        k.cost.final_adv_weight = c.f32();
        k.cost.final_opp_adv_weight = c.f32();
        k.safety_weight = c.f32();
        k.min_safe_dist = c.f32();
        k.delta = c.f32();
        k.P = c.i32();
    }
    else if (c.ok())
        c.fail("received invalid controller kind " + std::to_string(k.kind));
    c.finished();
    if (int rc = c.result())
        return rc;
    *cont = k;
    return DR_OK;
}

int dr_wire_decode_state(const void* frame, size_t len, dr_sr_state* state)
{
    if (!frame || !state)
        return host_fail(DR_EINVAL, "decode_state: NULL argument");
    dr_sr_state s;
    std::memset(&s, 0, sizeof(s));
    Cursor c(frame, len);
    c.header("initial state");
    c.arr(s.pos, 4, "pos");
    c.arr(s.vel, 4, "vel");
    c.arr(s.s, 2, "S");
    c.arr(s.lat_dist, 2, "latDist");
    float laps[2] = {0.0f, 0.0f};
    c.arr(laps, 2, "laps");
    c.finished();
    if (int rc = c.result())
        return rc;
    s.laps[0] = (int32_t)laps[0];
    s.laps[1] = (int32_t)laps[1];
    *state = s;
    return DR_OK;
}

int dr_session_create(const void* env_frame, size_t env_len, const void* cont_frame, size_t cont_len, const void* state_frame,
                      size_t state_len, int32_t max_steps, dr_session** out)
{
    if (!out)
        return host_fail(DR_EINVAL, "out is NULL");
    *out = nullptr;
    dr_session* s = new (std::nothrow) dr_session();
    if (!s)
        return host_fail(DR_ENOMEM, "host allocation failed");
    auto bail = [&](int rc) {
        delete s;
        return rc;
    };
    if (int rc = dr_wire_decode_env(env_frame, env_len, &s->env, &s->info))
        return bail(rc);
    if (int rc = dr_wire_decode_controller(cont_frame, cont_len, &s->cont))
        return bail(rc);
    if (int rc = dr_wire_decode_state(state_frame, state_len, &s->state))
        return bail(rc);
    s->max_steps = max_steps > 0 ? max_steps : DEFAULT_MAX_STEPS;
    const dr_wire_controller& k = s->cont;
    const uint64_t seed = (uint64_t)(uint32_t)s->info.seed;
    if (k.kind == DR_CONT_MPPI)
    {
        dr_bm_instance ins;
        std::memset(&ins, 0, sizeof(ins));
        ins.n_samples = k.n_samples;
        ins.n_timesteps = k.n_timesteps;
        ins.n_knots = k.n_knots;
        std::memcpy(ins.knots, k.knots, sizeof(ins.knots));
        ins.seed = seed;
        ins.lambda = k.lambda;
        ins.sigma = k.sigma;
        ins.min_confidence = k.min_confidence;
        ins.true_theta = s->info.true_theta;
        ins.adapt_lambda = 1; // mppi_controller.cu:413-424
        std::memcpy(ins.init_belief, s->info.init_belief, sizeof(ins.init_belief));
        ins.n_verif_samples = k.n_verif_samples;
        // a horizon beyond T would read past the plans (kernels.cu:525-531 indexes the
        // nominal at local time < verifHorizon); it is capped at T
        ins.verif_horizon = std::min(k.verif_horizon, k.n_timesteps);
        ins.beta = k.beta;
        ins.max_verif_eps = k.max_verif_eps;
        if (int rc = dr_bm_create(&s->env, &k.cost, &ins, 1, &s->bm))
            return bail(rc);
        if (int rc = dr_bm_set_states(s->bm, &s->state, 1))
            return bail(rc);
    }
    else
    {
        if (k.P < 1)
            return bail(host_fail(DR_EINVAL, "PR-MPPI header: P must be >= 1"));
        dr_prsr_instance ins;
        std::memset(&ins, 0, sizeof(ins));
        ins.n_samples = k.n_samples;
        ins.n_timesteps = k.n_timesteps;
        ins.sampled = 1; // P models drawn from the belief (prmppi_kernels.cu:23-51)
        ins.n_draws = k.P;
        ins.seed = seed;
        ins.lambda = k.lambda;
        ins.sigma = k.sigma;
        ins.safety_weight = k.safety_weight;
        ins.adapt_lambda = 1; // prmppi_controller.cu:328-346
        std::memcpy(ins.belief, s->info.init_belief, sizeof(ins.belief));
        ins.true_theta = s->info.true_theta;
        if (int rc = dr_prsr_create(&s->env, &k.cost, &ins, 1, &s->pr))
            return bail(rc);
        if (int rc = dr_pr_set_sr_states(s->pr, &s->state, 1))
            return bail(rc);
    }
    *out = s;
    return DR_OK;
}

void dr_session_destroy(dr_session* s)
{
    delete s;
}

int dr_session_next(dr_session* s, const uint8_t** frame, size_t* len, int32_t* msg_type)
{
    if (!s || !frame || !len || !msg_type)
        return host_fail(DR_EINVAL, "session_next: NULL argument");
    while (s->queue.empty() && !s->finished)
        if (int rc = s->advance())
            return rc;
    if (s->queue.empty())
    {
        *frame = nullptr;
        *len = 0;
        *msg_type = -1;
        return DR_OK;
    }
    s->current = std::move(s->queue.front());
    s->queue.pop_front();
    *frame = s->current.data();
    *len = s->current.size();
    int32_t t;
    std::memcpy(&t, s->current.data(), 4);
    *msg_type = t;
    return DR_OK;
}

} // extern "C"
