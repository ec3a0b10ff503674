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


def _resources(src: str, tmp_path) -> dict:
    """{kernel: (scratch bytes per lane, LDS bytes)} from hipcc's resource remarks."""
    cmd = ["hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-x", "hip", "--cuda-device-only",
           "-c", os.path.join(CSRC, src), "-o", str(tmp_path / (src + ".o")),
           "-Rpass-analysis=kernel-resource-usage"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    usage = {}
    name = None
    for line in out.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            usage[name] = [None, None]
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            usage[name][0] = int(m.group(1))
        m = re.search(r"LDS Size \[bytes/block\]: (\d+)", line)
        if m and name:
            usage[name][1] = int(m.group(1))
    return {k: tuple(v) for k, v in usage.items() if "island_step_kernel" in k}


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not on PATH")
def test_island_kernels_use_no_scratch(tmp_path):
    kernels = _resources("island.hip", tmp_path)
    assert len(kernels) == 8, kernels
    assert all(v[0] == 0 for v in kernels.values()), kernels


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not on PATH")
def test_large_island_kernel_fits_one_workgroup_per_cu(tmp_path):
    """The mechanism-sized kernel (up to 48 bodies, 128 joints) keeps its operator
    workspace in HBM and everything else in LDS: no scratch, and LDS within a CU's
    160 KiB."""
    kernels = _resources("island_large.hip", tmp_path)
    assert len(kernels) == 1, kernels
    (scratch, lds), = kernels.values()
    assert scratch == 0
    assert lds is not None and lds <= 160 * 1024
