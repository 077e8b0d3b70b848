"""The C++ controller shim (include/drone_race/controllers.hpp): it compiles
and links against the engine (CPU), and on the GPU its getControl loop gives
the same actions as the Python handle over the same C ABI."""
import importlib.util
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "tests", "cpp", "controller_demo")


def _build_module():
    spec = importlib.util.spec_from_file_location("drone_race_amd_build", os.path.join(ROOT, "drone-race_amd", "build.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_shim_compiles_and_links(tmp_path):
    out = _build_module().build_demo(str(tmp_path / "demo"))
    assert os.path.exists(out)


@pytest.mark.gpu
def test_shim_matches_python_handle(pkg):
    exe = DEMO if os.path.exists(DEMO) else _build_module().build_demo()
    res = subprocess.run([exe, "3"], check=True, capture_output=True, text=True, timeout=120).stdout.split("\n")
    acts = np.array([[float(x) for x in l.split()[2:]] for l in res if l.startswith("action")], dtype=np.float32)
    assert acts.shape == (3, 4)
    assert any(l.startswith("throws:") for l in res)
    abi = pkg.abi
    P = abi.QuadParams(1.0, 9.81, 0.1, 20.0, 0.0, 30.0, 6.0, 0.15, 0.02)
    K = abi.CostCfg(0.95, 1.0, 10.0, 1.5, 1000.0, 1000.0, 1000.0, 10.0, 0.0, 2.0, 50.0, 0.0)
    tr = abi.Track()
    tr.n_gates, tr.n_win_laps, tr.ground_z = 4, 100, 0.0
    for g, (cx, cy, nx, ny) in enumerate([(10, 0, 0, 1), (0, 10, -1, 0), (-10, 0, 0, -1), (0, -10, 1, 0)]):
        tr.gates[g].center[:] = [cx, cy, 4.0]
        tr.gates[g].normal[:] = [nx, ny, 0.0]
        tr.gates[g].radius = 1.0
    h = pkg.MPPI(P, K, tr, 2048, 40, 7, fill=(9.81, 0, 0, 0), max_opponents=1)
    nom = np.zeros((40, 4), np.float32)
    nom[:, 0] = 9.81
    h.set_nominal(nom)
    s = abi.DroneState()
    s.p[:] = [7.0, -7.0, 4.0]
    s.q[:] = [1.0, 0.0, 0.0, 0.0]
    s.gate = 3
    o = abi.Opponent()
    o.p[:] = [8.0, -6.0, 4.0]
    o.v[:] = [1.0, 1.0, 0.0]
    py = np.stack([h.step(s, [o]) for _ in range(3)])
    h.close()
    assert np.array_equal(acts, py)
    # PR-MPPI shim: same steps through the Python PR handle (opponent as an uncontrolled drone)
    pr_acts = np.array([[float(x) for x in l.split()[2:6]] for l in res if l.startswith("pr_action")], dtype=np.float32)
    sc = pkg.scenarios
    shift = np.zeros((2, 1, 3), np.float32)
    shift[0, 0, 0] = 1.0
    shift[1, 0, 1] = -1.0
    ins = sc.pr_instance(1024, 30, 0, 0, 11, [0.3, 0.7], shift, lam=1.0, sigma=(2.0, 0.5, 0.5, 0.5), fill=(9.81, 0, 0, 0))
    drones = (abi.DroneState * 2)()
    drones[0] = s
    drones[1].p[:] = [8.0, -6.0, 4.0]
    drones[1].v[:] = [1.0, 1.0, 0.0]
    drones[1].q[:] = [1.0, 0.0, 0.0, 0.0]
    scn = sc.PrScenario("shim_pr", [tr], [ins], [0, 0], drones, 1, "")
    scn.params, scn.cost = P, K
    eng = pkg.PREngine(scn)
    pnom = np.zeros((30, 4), np.float32)
    pnom[:, 0] = 9.81
    eng.set_nominals(0, pnom, pnom)
    py_pr = []
    for _ in range(3):
        eng.set_drones(drones)
        py_pr.append(eng.step(1, advance=False)[0])
    eng.close()
    assert np.array_equal(pr_acts, np.stack(py_pr))
    # StratRace2D shim on the analytic circle
    sr_acts = np.array([[float(x) for x in l.split()[2:4]] for l in res if l.startswith("sr_action")], dtype=np.float32)
    env = abi.SrEnv()
    env.dt, env.drone_radius = 0.1, 0.4
    env.max_speed[:] = [3.0, 2.0]
    env.max_accel[:] = [2.0, 3.0]
    env.track_width, env.n_win_laps, env.kp, env.kv, env.max_opp_lat_dist_fact = 2.0, 1, 10.0, 5.0, 1.9
    env.track_length = float(np.float32(2.0 * np.pi * 6.0))
    env.n_models = 2
    env.opp[0] = abi.SrOppConfig(0.0, float(np.float32(1.0) / (np.float32(1.6) * np.float32(1.6))), 50.0, 1.0, 0.1)
    env.opp[1] = abi.SrOppConfig(0.0, float(np.float32(1.0) / (np.float32(1.6) * np.float32(1.6))), 0.0, 1.0, 0.1)
    th = 2.0 * np.pi * np.arange(512) / 512
    for i in range(512):
        env.p_grid[i][:] = [float(np.float32(6.0 * np.cos(th[i]))), float(np.float32(6.0 * np.sin(th[i])))]
        env.t_grid[i][:] = [float(np.float32(-np.sin(th[i]))), float(np.float32(np.cos(th[i])))]
        env.kappa_grid[i] = float(np.float32(1.0) / np.float32(6.0))
    st = abi.SrState()
    st.pos[:] = [env.p_grid[26][0], env.p_grid[26][1], env.p_grid[102][0], env.p_grid[102][1]]
    st.s[:] = [float(np.float32(np.float32(26) / np.float32(512) * np.float32(env.track_length))),
               float(np.float32(np.float32(102) / np.float32(512) * np.float32(env.track_length)))]
    sscn = sc.SrScenario("shim_sr", env, abi.SrCost(1.3, 50.0, 1.7, 1e6, 1e6, 1e5, 500.0, 0.0),
                         [sc.sr_instance(4096, 40, 5, theta=1, lam=10.0, sigma=1.5)], [st], 1, "")
    seng = pkg.SREngine(sscn)
    py_sr = []
    for _ in range(3):
        seng.set_states([st])
        py_sr.append(seng.step(1, advance=False)[0])
    seng.close()
    assert np.array_equal(sr_acts, np.stack(py_sr))
    # branching spline MPPI shim (the reference's default MPPIController)
    bm_lines = [l.split() for l in res if l.startswith("bm_action")]
    bm_acts = np.array([[float(x) for x in l[2:4]] for l in bm_lines], dtype=np.float32)
    ins = sc.bm_instance(2048, 30, 9, n_knots=8, lam=10.0, sigma=0.5, min_confidence=0.6, true_theta=1, adapt=1, n_verif=4096,
                         verif_horizon=20, beta=1e-5, max_verif_eps=0.01)
    beng = pkg.BMEngine(sc.SrScenario("shim_bm", env, abi.SrCost(1.4, 50.0, 1.8, 1e7, 1e5, 1e5, 50.0, 0.0), [ins], [st], 1, ""))
    py_bm, py_branch, py_verif = [], [], []
    for i, b in enumerate([[0.5, 0.5], [0.3, 0.7], [0.9, 0.1]][:3]):
        beng.set_belief(0, np.float32(b))
        beng.set_states([st])
        py_bm.append(beng.step(1, advance=False)[0])
        d = beng.diag()[0]
        py_branch.append(d.branch)
        py_verif.append((d.use_new_plan, d.certified_loss, d.fail_count))
    beng.close()
    assert np.array_equal(bm_acts, np.stack(py_bm))
    assert [int(l[5]) for l in bm_lines] == py_branch
    assert [(int(l[9]), float(l[11]), int(l[13])) for l in bm_lines] == py_verif
    # the same controller on HiddenObs
    ho_lines = [l.split() for l in res if l.startswith("ho_action")]
    ho_acts = np.array([[float(x) for x in l[2:4]] for l in ho_lines], dtype=np.float32)
    he = sc.ho_env()
    hs = abi.HoState()
    hs.pos[:] = [1.0, 4.0]
    hs.vel[:] = [0.3, 1.0]
    hins = sc.bm_instance(2048, 40, 13, n_knots=10, lam=10.0, sigma=0.8, min_confidence=0.95, true_theta=0, adapt=1)
    heng = pkg.HoBMEngine(sc.SrScenario("shim_ho", he, sc.ho_cost(), [hins], [hs], 1, ""))
    py_ho, py_hb = [], []
    for _ in range(3):
        heng.set_belief(0, np.float32([0.5, 0.5]))
        heng.set_states([hs])
        py_ho.append(heng.step(1, advance=False)[0])
        py_hb.append(heng.diag()[0].branch)
    heng.close()
    assert np.array_equal(ho_acts, np.stack(py_ho))
    assert [int(l[5]) for l in ho_lines] == py_hb


# --------------------------------------------------------------------------------------------
# The reference-engine binding (include/drone_race/reference_controllers.hpp): controllers that
# the reference's SimulationEngine holds as std::unique_ptr<Controller> and downcasts.

REF_DEMO = os.path.join(ROOT, "tests", "cpp", "reference_engine_demo")


def test_reference_binding_compiles_and_links(tmp_path):
    out = _build_module().build_reference_demo(str(tmp_path / "refdemo"))
    assert os.path.exists(out)


def _ref_input(tmp_path, kind, env, seed, cfg, knots, belief0, records):
    envf = tmp_path / "env.bin"
    envf.write_bytes(bytes(env))
    lines = [f"env {envf}", f"kind {kind}", f"seed {seed}", "belief0 " + " ".join(f"{b:.9g}" for b in belief0)]
    lines += [f"{k} {v!r}" for k, v in cfg.items()]
    if knots:
        lines.append("knots " + " ".join(str(k) for k in knots))
    for belief, st in records:
        vals = list(belief) + list(st.pos) + list(st.vel) + list(st.s) + list(st.lat_dist)
        lines.append("step " + " ".join(repr(float(v)) for v in vals) + " " + " ".join(str(int(x)) for x in st.laps))
    p = tmp_path / f"{kind}.txt"
    p.write_text("\n".join(lines) + "\n")
    return p


def _run_ref_demo(path):
    exe = REF_DEMO if os.path.exists(REF_DEMO) else _build_module().build_reference_demo()
    out = subprocess.run([exe, str(path)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout.split("\n")


def _field(line, name):
    toks = line.split()
    return toks[toks.index(name) + 1]


def _pad4(b):
    out = [0.0] * 4
    out[:len(b)] = [float(x) for x in b]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("n_verif", [0, 8192])
def test_reference_binding_mppi_matches_python_handle(pkg, tmp_path, n_verif):
    """MPPIController through std::unique_ptr<Controller> in an engine that mirrors the reference's
    construction, downcasts and belief hand-off: over a closed loop recorded from the Python
    branching-MPPI handle (without and with plan verification), the actions are bit-for-bit the
    handle's, and failCount, epsilon, epsilonPartial, useNewPlan, certifiedLoss, the temperature and
    the dense plans are the handle's diagnostics.  The controller's int seed s is the Philox key
    (s, 0), as the handle is built here."""
    sc = pkg.scenarios
    n_steps = 4
    scn = sc.bm_race(n_inst=1, n_samples=4096, n_timesteps=30, n_knots=6, n_verif=n_verif)
    ins = scn.instances[0]
    ins.seed = 1234567
    eng = pkg.BMEngine(scn)
    records, acts, diags, noms = [], [], [], []
    for _ in range(n_steps):
        records.append((_pad4(eng.belief(0)), eng.get_states()[0]))
        acts.append(eng.step(1, advance=True)[0].copy())
        diags.append(eng.diag()[0])
        noms.append(eng.nominal(0).reshape(-1)[:2].copy())
    eng.close()
    cfg = dict(nSamples=ins.n_samples, nTimesteps=ins.n_timesteps, nKnots=ins.n_knots, invTemperature=ins.lambda_,
               samplingNoise=ins.sigma, collDistFactor=scn.cost.coll_dist_factor, boundaryCost=scn.cost.boundary_cost,
               boundaryThresholdFactor=scn.cost.boundary_threshold_factor, outsideCost=scn.cost.outside_cost,
               winCost=scn.cost.win_cost, oppWinCost=scn.cost.opp_win_cost, finalAdvWeight=scn.cost.final_adv_weight,
               finalOppAdvWeight=scn.cost.final_opp_adv_weight, minConfidence=ins.min_confidence,
               nVerifSamples=ins.n_verif_samples, beta=ins.beta, verifHorizon=ins.verif_horizon, maxVerifEps=ins.max_verif_eps)
    path = _ref_input(tmp_path, "mppi", scn.env, ins.seed, cfg, list(ins.knots[:ins.n_knots]), records[0][0], records)
    out = _run_ref_demo(path)
    act_lines = [l for l in out if " action " in l]
    info = [l for l in out if " mppi " in l]
    assert len(act_lines) == n_steps and len(info) == n_steps
    got = np.array([[float(x) for x in l.split()[3:5]] for l in act_lines], np.float32)
    assert np.array_equal(got, np.array(acts, np.float32)), (got, acts)
    if n_verif == 0:
        assert np.abs(got).max() > 0.0
    for k, (l, d) in enumerate(zip(info, diags)):
        assert int(_field(l, "use_new_plan")) == int(d.use_new_plan)
        assert float(_field(l, "epsilon")) == d.epsilon and float(_field(l, "certified_loss")) == d.certified_loss
        assert float(_field(l, "epsilon_partial")) == d.epsilon_partial
        assert int(_field(l, "fail_count")) == int(d.fail_count)
        assert np.float32(_field(l, "lambda")) == np.float32(d.lambda_)
        assert int(_field(l, "back")) == 1
        toks = l.split()
        i = toks.index("nominal0")
        assert np.array_equal(np.float32(toks[i + 1:i + 3]), noms[k])
    assert any(l.startswith("throws:") for l in out)


@pytest.mark.gpu
def test_reference_binding_prmppi_matches_python_handle(pkg, tmp_path):
    """PRMPPIController through the same engine: P = 9 models drawn from h_belief each step, the
    actions bit-for-bit the Python PR-MPPI handle's over its recorded closed loop, and useNomPlan,
    resetNom and both plans its diagnostics."""
    sc = pkg.scenarios
    n_steps = 4
    scn = sc.pr_sr_race(n_inst=1, n_samples=2048, n_timesteps=20)
    ins = scn.instances[0]
    ins.seed = 7654321
    ins.adapt_lambda = 1  # the reference adapts its three temperatures every step
    eng = pkg.PRSREngine(scn)
    n_models = scn.env.n_models
    records, acts, diags, plans = [], [], [], []
    for _ in range(n_steps):
        records.append((_pad4(eng.belief(0, n_models)), eng.get_states()[0]))
        acts.append(eng.step(1, advance=True)[0].copy())
        diags.append(eng.diag()[0])
        nom, rob = eng.get_nominals(0)
        plans.append((nom.reshape(-1)[:2].copy(), rob.reshape(-1)[:2].copy()))
    eng.close()
    cfg = dict(nSamples=ins.n_samples, nTimesteps=ins.n_timesteps, invTemperature=ins.lambda_, samplingNoise=ins.sigma,
               collDistFactor=scn.cost.coll_dist_factor, boundaryCost=scn.cost.boundary_cost,
               boundaryThresholdFactor=scn.cost.boundary_threshold_factor, winCost=scn.cost.win_cost,
               oppWinCost=scn.cost.opp_win_cost, finalAdvWeight=scn.cost.final_adv_weight,
               finalOppAdvWeight=scn.cost.final_opp_adv_weight, safetyWeight=ins.safety_weight, P=float(ins.n_draws))
    path = _ref_input(tmp_path, "prmppi", scn.env, ins.seed, cfg, None, records[0][0], records)
    out = _run_ref_demo(path)
    act_lines = [l for l in out if " action " in l]
    info = [l for l in out if " prmppi " in l]
    assert len(act_lines) == n_steps and len(info) == n_steps
    got = np.array([[float(x) for x in l.split()[3:5]] for l in act_lines], np.float32)
    assert np.array_equal(got, np.array(acts, np.float32)), (got, acts)
    assert np.abs(got).max() > 0.0
    for k, (l, d) in enumerate(zip(info, diags)):
        assert int(_field(l, "use_nom_plan")) == int(d.use_nom_plan)
        assert int(_field(l, "reset_nom")) == int(d.reset_nom)
        toks = l.split()
        i, j = toks.index("nom0"), toks.index("rob0")
        assert np.array_equal(np.float32(toks[i + 1:i + 3]), plans[k][0])
        assert np.array_equal(np.float32(toks[j + 1:j + 3]), plans[k][1])
        assert int(_field(l, "back")) == 1
    assert any(l.startswith("throws:") for l in out)
