"""The island kernels keep every value in registers: compiled for gfx950 they use no
scratch (private memory). Register spills of the two-islands-per-CU variant once cost
a fifth of its time and most of its HBM traffic, so a change that brings them back
fails here, on the CPU, before any GPU run."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "exgame_amd", "csrc")


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not on PATH")
def test_island_kernels_use_no_scratch(tmp_path):
    cmd = ["hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-x", "hip", "--cuda-device-only",
           "-c", os.path.join(CSRC, "island.hip"), "-o", str(tmp_path / "island.o"),
           "-Rpass-analysis=kernel-resource-usage"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    usage = {}
    name = None
    for line in out.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            usage[name] = int(m.group(1))
    kernels = {k: v for k, v in usage.items() if "island_step_kernel" in k}
    assert len(kernels) == 6, usage
    assert all(v == 0 for v in kernels.values()), kernels
