"""The adjoint's dual-number element derivatives (nexdynipc_amd/csrc/dual.h), compiled
for the host, agree with central finite differences of the same formulas in doubles:
the quaternion update (also at a zero rotation, where the smooth limit is taken), the
rotation vector at a zero vector part, the friction gradient's Jacobians in the Newton
variables, step-start poses and mu, and the fixed / hinge joint gradients and
constraints in the Newton variables and reference rotations."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_dual_jacobians_match_finite_differences(tmp_path):
    exe = tmp_path / "dual_check"
    src = os.path.join(ROOT, "tests", "native", "dual_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++20", "-ffp-contract=off", "-o", str(exe), src], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("OK")
