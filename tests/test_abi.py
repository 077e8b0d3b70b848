"""The C ABI: every function include/dr_api.h declares is exported by the
built library and bound by the Python mirror, and every struct has the same
size and field offsets in C (gcc) and in ctypes.  No GPU call is made."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dr_api.h")

STRUCTS = {
    "dr_quad_params": "QuadParams", "dr_gate": "Gate", "dr_track": "Track", "dr_cost_cfg": "CostCfg",
    "dr_drone_state": "DroneState", "dr_opponent": "Opponent", "dr_instance": "Instance", "dr_diag": "Diag",
    "dr_launch_cfg": "LaunchCfg", "dr_mppi_cfg": "MPPICfg", "dr_pr_instance": "PrInstance", "dr_pr_diag": "PrDiag",
    "dr_sr_opp_config": "SrOppConfig", "dr_sr_env": "SrEnv", "dr_sr_state": "SrState", "dr_sr_cost": "SrCost",
    "dr_sr_instance": "SrInstance", "dr_prsr_instance": "PrsrInstance", "dr_bm_instance": "BmInstance", "dr_bm_diag": "BmDiag",
    "dr_wire_env_info": "WireEnvInfo", "dr_wire_controller": "WireController", "dr_ho_env": "HoEnv", "dr_ho_state": "HoState",
    "dr_drace_pid": "DracePid", "dr_drace_env": "DraceEnv", "dr_drace_state": "DraceState",
}
RENAMED = {"lambda_": "lambda"}


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"^[A-Za-z_][\w \*]*?\b(dr_\w+)\s*\(", src, flags=re.M))


def test_header_and_binding_agree(pkg):
    decl = declared_functions()
    assert "dr_engine_create" in decl and "dr_mppi_step" in decl
    assert decl == set(pkg.abi.SIGNATURES), (decl ^ set(pkg.abi.SIGNATURES))


def test_library_exports_every_symbol(pkg):
    lib = pkg.abi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.dr_version() == 1


def test_struct_layouts_match_c(pkg, tmp_path):
    abi = pkg.abi
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, pyname in STRUCTS.items():
        st = getattr(abi, pyname)
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in st._fields_:
            cf = RENAMED.get(f, f)
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {cf}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(l.split() for l in out if l)
    for cname, pyname in STRUCTS.items():
        st = getattr(abi, pyname)
        assert int(got[cname]) == C.sizeof(st), cname
        for f, _ in st._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(st, f).offset, (cname, f)


def test_missing_library_fails_loudly(pkg, tmp_path):
    saved = pkg.abi._lib
    pkg.abi._lib = None
    try:
        with pytest.raises(pkg.EngineError):
            pkg.abi.load(str(tmp_path / "missing.so"))
    finally:
        pkg.abi._lib = saved
