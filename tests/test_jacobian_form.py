"""The refinement's reverse pass carries its adjoint vectors through per-step
Jacobian blocks (dr_adjoint.h quad_step_jacobian / quad_step_jvp_t) instead of
the vector-Jacobian product quad_step_vjp; both linearise the same step, so
they agree to rounding on random states, controls and adjoint vectors (host
build of the device header, no GPU)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_jacobian_blocks_equal_the_vjp(tmp_path):
    out = str(tmp_path / "jacobian_check")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off", "-x", "hip",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    os.path.join(ROOT, "tests", "cpp", "jacobian_check.cpp"), "-o", out], check=True)
    r = subprocess.run([out], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
