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
