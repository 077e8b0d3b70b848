"""drone-race MPPI / PR-MPPI control engine for AMD MI355X (gfx950).

autograd (torch.autograd.Function over the adjoint kernel) is imported on
demand: `from drone_race_amd.autograd import rollout_cost`.

The compute path is the HIP library libdr_amd.so (csrc/, C ABI in
include/dr_api.h); this package is its Python host side: ctypes bindings,
scenario builders, the engine handles and the wire protocol / simulation
session of the reference's ZMQ front-end (wire.py, serve.py).  There is no
CPU fallback.
"""
from . import abi, philox, scenarios, wire
from .abi import EngineError, load
from .engine import BMEngine, DraceBMEngine, Engine, HoBMEngine, MPPI, PREngine, PRSREngine, SREngine

__all__ = ["abi", "philox", "scenarios", "wire", "BMEngine", "DraceBMEngine", "Engine", "HoBMEngine", "MPPI", "PREngine", "PRSREngine", "SREngine", "EngineError", "load"]
