"""drone-race MPPI / PR-MPPI control engine for AMD MI355X (gfx950).

autograd (torch.autograd.Function over the adjoint kernel) is imported on
demand: `from drone_race_amd.autograd import rollout_cost`.

The compute path is the HIP library libdr_amd.so (csrc/, C ABI in
include/dr_api.h); this package is its Python host side: ctypes bindings,
scenario builders and the engine handles.  There is no CPU fallback.
"""
from . import abi, philox, scenarios
from .abi import EngineError, load
from .engine import BMEngine, Engine, MPPI, PREngine, PRSREngine, SREngine

__all__ = ["abi", "philox", "scenarios", "BMEngine", "Engine", "MPPI", "PREngine", "PRSREngine", "SREngine", "EngineError", "load"]
