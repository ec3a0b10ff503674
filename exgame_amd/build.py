"""Build step of the product library: compiles the HIP kernels and the C ABI for
gfx950 with hipcc into exgame_amd/_build/libexgame_hip.so (in-tree, so it travels
with the source tree)."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libexgame_hip.so")

SOURCES = ["island.hip", "island_large.hip", "singles.hip", "explicit.hip", "gravity.hip", "rows.hip", "rows_traced.hip",
           "exg_api.cpp", "store.cpp", "topology.cpp", "arena.cpp", "noalloc.hip"]
VIETE = os.path.join(HERE, "viete")
GENERATED = os.path.join(CSRC, "viete_gen.h")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# -packed-fp32-ops: no packed FP32 instructions anywhere in this library.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-I", os.path.join(HERE, "..", "include")]


def _compile(src: str) -> str:
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def generate() -> str:
    """Build the tracer (host code) and write csrc/viete_gen.h: the joint laws and the
    Pre stage of csrc/physics_dev.h traced and emitted as device functions."""
    os.makedirs(OUT_DIR, exist_ok=True)
    exe = os.path.join(OUT_DIR, "viete_gen")
    cmd = [HIPCC, "-x", "hip", "--offload-host-only", "-O1", "-std=c++17", "-ffp-contract=off", "-I", CSRC,
           os.path.join(VIETE, "gen.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"the tracer did not build:\n{r.stderr[-6000:]}")
    tmp = GENERATED + ".tmp"
    r = subprocess.run([exe, tmp], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"the tracer failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, GENERATED)
    return exe


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    generate()
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
