"""The build-time tracer (exgame_amd/viete): the traced programs reproduce the
templates they were traced from bit for bit (its own interpreter against direct f64
evaluation of physics_dev.h, on seeded inputs including Pre's seams), and the
committed csrc/viete_gen.h is what the tracer emits today."""
import os
import shutil
import subprocess

import pytest

from exgame_amd import build as B

pytestmark = pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists(B.HIPCC), reason="no hipcc")


@pytest.fixture(scope="module")
def gen(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("viete") / "gen")
    cmd = [B.HIPCC, "-x", "hip", "--offload-host-only", "-O1", "-std=c++17", "-ffp-contract=off", "-I", B.CSRC,
           os.path.join(B.VIETE, "gen.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def test_traced_programs_reproduce_the_templates_bit_for_bit(gen):
    r = subprocess.run([gen, "--check", "400"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:]
    assert " 0 mismatches" in r.stdout
    # Pre's three EffectiveZero seams on one argument and one on its scaled copy
    # enumerate to four reachable paths; the joint laws have none
    assert "viete_pre: " in r.stdout and "4 paths" in r.stdout.split("viete_pre: ")[1]
    for law in ("simple_spring_damper", "critically_damped_warped", "perpendicular_damper_warped",
                "torsional_damper_warped"):
        assert f"viete_{law}_full: " in r.stdout


def test_committed_header_is_current(gen, tmp_path):
    out = str(tmp_path / "viete_gen.h")
    r = subprocess.run([gen, out], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    with open(out) as a, open(B.GENERATED) as b:
        assert a.read() == b.read(), "csrc/viete_gen.h is stale: run exgame_amd/build.py"
