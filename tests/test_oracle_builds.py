"""The timing builds of the oracle (-O3 for x86-64-v3 / x86-64-v4, used by
bench.py's CPU baseline) compute what the parity build computes: the same
sources with the same arithmetic flags, so the same bits."""
import os

import numpy as np
import pytest

import pyoracle


def _variants():
    out = [os.path.join(pyoracle.HERE, "liboracle_t3.so")]
    path, _ = pyoracle.timing_build()
    if path not in out:
        out.append(path)
    return out


@pytest.mark.parametrize("path", _variants())
def test_timing_build_matches_parity_build(pkg, oracle, path):
    sc = pkg.scenarios
    scn = sc.c1_four_drones(n_samples=96, n_timesteps=40)
    fast = pyoracle.Oracle(pkg.abi, path)
    ins = scn.instances[0]
    opp = []
    for d in range(1, 4):
        o = pkg.abi.Opponent()
        for c in range(3):
            o.p[c], o.v[c] = scn.drones[d].p[c], scn.drones[d].v[c]
        opp.append(o)
    nom = np.tile(np.array(ins.shift_fill[:], np.float32), (ins.n_timesteps, 1))
    nom[:, 1] = 0.3
    sig = np.array(ins.sigma[:], np.float32)
    for f64 in (0, 1):
        a = oracle.rollout(f64, scn.params, scn.cost, scn.tracks[0], scn.drones[0], opp, nom, 96, 40, ins.seed, 3, sig)
        b = fast.rollout(f64, scn.params, scn.cost, scn.tracks[0], scn.drones[0], opp, nom, 96, 40, ins.seed, 3, sig)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    U = (nom[None] + 0.2 * np.random.default_rng(1).standard_normal((3, 40, 4))).astype(np.float32)
    ja, ga, sa = oracle.cost_grad(0, scn.params, scn.cost, scn.tracks[0], scn.drones[0], opp, U)
    jb, gb, sb = fast.cost_grad(0, scn.params, scn.cost, scn.tracks[0], scn.drones[0], opp, U)
    assert np.array_equal(ja, jb) and np.array_equal(ga, gb) and np.array_equal(sa, sb)
    # closed loop with the adjoint refinement
    runs = []
    for o in (oracle, fast):
        drones = (pkg.abi.DroneState * 4)(*scn.drones)
        nominal = np.concatenate([np.tile(np.array(i.shift_fill[:], np.float32), (i.n_timesteps, 1)).reshape(-1)
                                  for i in scn.instances]).astype(np.float32)
        acts, _ = o.world_run(scn, drones, nominal, [0] * 4, [i.lambda_ for i in scn.instances], 3, threads=4,
                              refine=(3, 0.05, 10))
        runs.append((acts, nominal, [list(d.p) + list(d.q) for d in drones]))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert runs[0][2] == runs[1][2]
