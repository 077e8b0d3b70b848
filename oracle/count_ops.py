"""ORACLE — test infrastructure only.

Algorithmic FP32 work per rollout sample-step of the benchmark workload
(4 drones, K=16384, T=100, first control step): dynamics, gate logic and
costs counted by the oracle's op-counting scalar on the first `n` samples
of each instance, plus the per-step noise cost read off the code
(Box-Muller: log 20, sincos 21, 5 more per pair, two pairs; 4 sigma
scalings: 96 flops).  Prints the figure bench.py uses."""
import ctypes as C
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
NOISE_FLOPS_PER_STEP = 2 * (20 + 21 + 5) + 4


def package():
    name = "drone_race_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = os.path.join(ROOT, "drone-race_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"), submodule_search_locations=[d])
    m = importlib.util.module_from_spec(spec)
    sys.modules[name] = m
    spec.loader.exec_module(m)
    return m


def count(scn, n=1024):
    sys.path.insert(0, HERE)
    import pyoracle
    pkg = package()
    abi = pkg.abi
    L = pyoracle.lib(abi)
    L.or_count_rollout_ops.argtypes = [C.POINTER(abi.QuadParams), C.POINTER(abi.CostCfg), C.POINTER(abi.Track),
                                       C.POINTER(abi.DroneState), C.POINTER(abi.Opponent), C.c_int,
                                       C.POINTER(C.c_float), C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.POINTER(C.c_float), C.POINTER(C.c_longlong)]
    flops = 0
    steps = 0
    for ins in scn.instances:
        opp = []
        for d in range(len(scn.drone_world)):
            if d != ins.drone and scn.drone_world[d] == scn.drone_world[ins.drone]:
                o = abi.Opponent()
                for c in range(3):
                    o.p[c], o.v[c] = scn.drones[d].p[c], scn.drones[d].v[c]
                opp.append(o)
        oa = (abi.Opponent * max(1, len(opp)))(*opp)
        nom = np.tile(np.array(ins.shift_fill[:], np.float32), (ins.n_timesteps, 1)).reshape(-1)
        sig = np.array(ins.sigma[:], np.float32)
        out = C.c_longlong()
        rc = L.or_count_rollout_ops(C.byref(scn.params), C.byref(scn.cost), C.byref(scn.tracks[ins.track]),
                                    C.byref(scn.drones[ins.drone]), oa, len(opp),
                                    nom.ctypes.data_as(C.POINTER(C.c_float)), n, ins.n_timesteps,
                                    ins.seed & 0xFFFFFFFF, ins.seed >> 32, 0,
                                    sig.ctypes.data_as(C.POINTER(C.c_float)), C.byref(out))
        assert rc == 0
        flops += out.value
        steps += n * ins.n_timesteps
    return flops / steps + NOISE_FLOPS_PER_STEP


if __name__ == "__main__":
    pkg = package()
    print(f"{count(pkg.scenarios.c1_four_drones()):.2f} FP32 flops per sample-step (4 drones, K=16384, T=100)")
