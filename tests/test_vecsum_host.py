"""Eigen-order sums of dynamic-size vectors (nexdynipc_amd/csrc/vecsum.h, used by the
oracle and the device for the line-search slopes and the joint energy), compiled for the
host, agree bit for bit with the same reduction run on SSE2 packet instructions in
Eigen's sequence; known answers and exact integer sums."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_eigen_order_sums_match_sse2_packets(tmp_path):
    exe = tmp_path / "vecsum_check"
    src = os.path.join(ROOT, "tests", "native", "vecsum_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++20", "-ffp-contract=off", "-o", str(exe), src], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("OK")
