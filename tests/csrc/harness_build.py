"""Builds the test harness libraries into tests/_build (called by __graft_entry__.build()):
  libkmath.so       — the device physics headers compiled for the host
  libmfma_probe.so  — the FP64 MFMA lane-layout probe (gfx950)
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(os.path.dirname(HERE), "_build")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def build() -> None:
    os.makedirs(OUT, exist_ok=True)
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-fPIC", "-shared",
                    "-ffp-contract=off", os.path.join(HERE, "kmath.cpp"), "-o", os.path.join(OUT, "libkmath.so")],
                   check=True)
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-fPIC", "-shared", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", os.path.join(HERE, "mfma_probe.hip"), "-o",
                    os.path.join(OUT, "libmfma_probe.so")], check=True, capture_output=True)


if __name__ == "__main__":
    build()
