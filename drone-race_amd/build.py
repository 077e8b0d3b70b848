"""Build recipe for the HIP control engine (gfx950).

Compiles drone-race_amd/csrc/*.hip with hipcc into drone-race_amd/libdr_amd.so.
Floating-point flags are part of the contract with the host oracle:
-ffp-contract=off (only explicit fmaf fuses), correctly rounded f32 division
and square root, no fast-math, packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdr_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["dr_kernels.hip", "dr_engine.hip", "dr_prkernels.hip", "dr_prengine.hip", "dr_srkernels.hip", "dr_srengine.hip",
           "dr_prsrkernels.hip", "dr_bmkernels.hip", "dr_bmengine.hip", "dr_session.hip"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _compile(src, verbose):
    obj = os.path.join(CSRC, src.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose=False, jobs=None):
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))


def build_demo(out=None):
    """Build tests/cpp/controller_demo against libdr_amd.so (C++ shim check)."""
    root = os.path.dirname(HERE)
    out = out or os.path.join(root, "tests", "cpp", "controller_demo")
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(root, "include"),
           os.path.join(root, "tests", "cpp", "controller_demo.cpp"), "-o", out,
           "-L", HERE, "-ldr_amd", "-Wl,-rpath," + HERE]
    subprocess.run(cmd, check=True)
    return out


def build_reference_demo(out=None):
    """Build tests/cpp/reference_engine_demo (the reference-engine binding check) against libdr_amd.so."""
    root = os.path.dirname(HERE)
    out = out or os.path.join(root, "tests", "cpp", "reference_engine_demo")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(root, "include"),
           os.path.join(root, "tests", "cpp", "reference_engine_demo.cpp"), "-o", out,
           "-L", HERE, "-ldr_amd", "-Wl,-rpath," + HERE]
    subprocess.run(cmd, check=True)
    return out
