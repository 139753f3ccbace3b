"""The C-ABI library loads and exports every function include/ndipc.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ndipc.h")
LIB = os.path.join(ROOT, "nexdynipc_amd", "_build", "libndipc.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|void\*|const char\*)\s+(ndipc_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert len(names) >= 25
    for n in ("ndipc_step", "ndipc_build_candidates", "ndipc_narrow_phase", "ndipc_barrier_eval",
              "ndipc_ccd_toi_batch", "ndipc_pcg_solve"):
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libndipc.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (ndipc_\w+)", out))
    assert set(declared()) <= exported


def test_python_binding_covers_header():
    from nexdynipc_amd import _lib

    src = open(_lib.__file__).read()
    for n in declared():
        assert f'"{n}"' in src, n
