"""Writes topk_adversarial.npz: score arrays on which libstdc++'s std::nth_element and
std::sort reach their depth limits (heap select / heap sort branches), built by
McIlroy's adversary ("A killer adversary for quicksort", 1999) in
tests/native/topk_select_check.cpp. Inputs only: the expected output of every case is
the oracle's select_top_k (std::nth_element + std::sort) at test time.

Run from the repository root: python tests/golden/make_topk_adversarial.py
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "topk_check")
        txt = os.path.join(tmp, "adv.txt")
        subprocess.run(["g++", "-O2", "-std=c++20", "-o", exe,
                        os.path.join(ROOT, "tests", "native", "topk_select_check.cpp")], check=True)
        subprocess.run([exe, "--dump", txt], check=True)
        cases = [list(map(int, line.split())) for line in open(txt)]
    out = {}
    for c, row in enumerate(cases):
        n, k = row[0], row[1]
        assert len(row) == n + 2
        out[f"case{c:02d}_n{n}_k{k}"] = np.asarray(row[2:], dtype=np.uint16)
    np.savez_compressed(os.path.join(HERE, "topk_adversarial.npz"), **out)
    print(len(out), "cases")


if __name__ == "__main__":
    main()
