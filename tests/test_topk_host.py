"""The device top-K restatement (nexdynipc_amd/csrc/topk_select.h), compiled for the
host, reproduces the standard library's std::nth_element and std::sort permutations
exactly on tie-heavy inputs (the oracle's select_top_k_indices calls both)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_topk_select_matches_std_nth_element(tmp_path):
    exe = tmp_path / "topk_check"
    src = os.path.join(ROOT, "tests", "native", "topk_select_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++20", "-o", str(exe), src], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("OK")
