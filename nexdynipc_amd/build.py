"""Builds the gfx950 HIP library nexdynipc_amd/_build/libndipc.so with hipcc.

FP contraction is off on host and device (-ffp-contract=off) so every FP64 formula
rounds exactly as written; packed-FP32 instructions are disabled for all HIP code.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libndipc.so")
SOURCES = ["ctx.hip", "geom.hip", "solver.hip", "ipc.hip"]
HOST_SOURCES = ["scene_io.cpp"]  # host-only C++ (scene and state I/O)
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
              "-I", os.path.join(HERE, "..", "include")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")]


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
        if src.endswith(".cpp"):
            cmd = [CXX] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        else:
            cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-8000:]}")
        if verbose and r.stderr.strip():
            print(r.stderr[-4000:], file=sys.stderr)
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES) + len(HOST_SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES + HOST_SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-8000:]}")
    os.replace(tmp, LIB)
    return LIB


def build_runner(verbose=False):
    """tests/native/_build/run_scene: the C++ scene runner over the C ABI (links
    libndipc.so; build() first)."""
    src = os.path.join(HERE, "..", "tests", "native", "run_scene.cpp")
    out_dir = os.path.join(HERE, "..", "tests", "native", "_build")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "run_scene")
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(HERE, "..", "include"), src, "-o",
           exe + ".tmp", "-L", OUT_DIR, "-lndipc", "-Wl,-rpath,$ORIGIN/../../../nexdynipc_amd/_build"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"run_scene build failed:\n{r.stderr[-8000:]}")
    if verbose and r.stderr.strip():
        print(r.stderr[-4000:], file=sys.stderr)
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_runner(verbose=True))
