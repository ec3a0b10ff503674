"""Mechanism-sized islands (more than 8 bodies or 16 joints, up to 48 bodies and 128
joints) on the device against the oracle: cloth grids of 16, 45 and 48 bodies, long
chains, alongside small islands in the same world. The reference solves such islands
with the same Newton loop as any other (crates/newton/src/integrator/implicit/newton.rs:334-683);
only the system is larger (up to 288 × 288).

Tolerances (f64): one step within 1e-12 relative, 20-step trajectories within 1e-9
relative; solver counters within 2 or 0.5 % (the Newton–Schulz products sum in MFMA
order, the oracle in row order, so a residual that lands on the tolerance can take
one more iteration on either side).
"""
import numpy as np
import pytest

import pyoracle
from exgame_amd import scenes
from exgame_amd.world import DeviceWorld

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def close_counters(d, o):
    d, o = d[:, :6].astype(np.int64).sum(0), o.sum(0)
    assert np.all(np.abs(d - o) <= np.maximum(2, 0.005 * o)), (d, o)


def test_cloth_grids_and_long_chains_one_step():
    scene = scenes.cloth_grids(shapes=((4, 4), (5, 9)), chains=(8, 12, 20))
    print("cloth_grids 4x4 5x9 + chains 8 12 20, one step", flush=True)
    dev = DeviceWorld(scene, tol2=1e-20)
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    dev.step(1, 1e-2, 1.0)
    orc.step(1, 1e-2, 1.0)
    a, b = dev.state(), orc.state()
    assert rel(a["pose"], b["pose"]) < 1e-12
    assert rel(a["momentum"], b["momentum"]) < 1e-12
    assert dev.fatals() == 0
    close_counters(dev.stats(), orc.stats())


def test_cloth_grids_20_step_trajectory():
    # m = 16 (38 joints) and m = 45 (118 joints), with an 8-body chain beside them
    scene = scenes.cloth_grids(shapes=((4, 4), (5, 9)), chains=(8,), seed=3)
    print("cloth_grids 4x4 5x9 + chain 8, 20 steps", flush=True)
    dev = DeviceWorld(scene, tol2=1e-20)
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    for _ in range(4):
        dev.step(5, 1e-2, 1.0)
        orc.step(5, 1e-2, 1.0)
        a, b = dev.state(), orc.state()
        assert rel(a["position"], b["position"]) < 1e-9
        assert rel(a["momentum"], b["momentum"]) < 1e-9
    close_counters(dev.stats(), orc.stats())


def test_largest_island_48_bodies_126_joints():
    scene = scenes.cloth_grids(shapes=((6, 8),), seed=5)
    assert scene.n_bodies == 48 and scene.n_joints == 126
    print("cloth_grids 6x8, 5 steps", flush=True)
    dev = DeviceWorld(scene, tol2=1e-20)
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    dev.step(5, 1e-2, 1.0)
    orc.step(5, 1e-2, 1.0)
    a, b = dev.state(), orc.state()
    assert rel(a["position"], b["position"]) < 1e-9
    assert rel(a["momentum"], b["momentum"]) < 1e-9
    close_counters(dev.stats(), orc.stats())


def test_stiff_cloth_subdivides_like_the_oracle():
    # k = 1e5 at h = 1e-2: the Newton solve fails at full step and the step is split
    # by the stiffness signal; the split sequence is the oracle's
    scene = scenes.cloth_grids(shapes=((4, 4), (3, 5)), k=1e5, seed=2)
    print("stiff cloth_grids 4x4 3x5 k=1e5, 3 steps", flush=True)
    dev = DeviceWorld(scene, tol2=1e-20)
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    dev.step(3, 1e-2, 1.0)
    orc.step(3, 1e-2, 1.0)
    o = orc.stats()
    assert o[:, 0].sum() > 3 * scene.n_islands, "the scene is meant to subdivide"
    a, b = dev.state(), orc.state()
    assert rel(a["position"], b["position"]) < 1e-9
    assert rel(a["momentum"], b["momentum"]) < 1e-9
    d = dev.stats()
    assert np.array_equal(d[:, 0], o[:, 0]), (d[:, 0], o[:, 0])   # spans taken
    close_counters(d, o)


def test_cloth_conserves_linear_momentum_without_gravity():
    # internal forces only: joint wrenches come in equal and opposite pairs, so the
    # cloth's world linear momentum is constant at every Newton iterate
    from exgame_amd import pga
    scene = scenes.cloth_grids(shapes=((5, 9),), seed=7)
    print("cloth_grids 5x9 without gravity, 10 steps", flush=True)
    dev = DeviceWorld(scene, tol2=1e-20, gravity=False)
    p0 = pga.world_momentum(scene.pose, scene.momentum)[:, :3].sum(0)
    dev.step(10, 1e-2, 1.0)
    st = dev.state()
    p1 = pga.world_momentum(st["pose"], st["momentum"])[:, :3].sum(0)
    scale = np.abs(pga.world_momentum(scene.pose, scene.momentum)[:, :3]).sum()
    assert np.linalg.norm(p1 - p0) < 1e-12 * scale
    orc = pyoracle.OracleWorld(scene, tol2=1e-20, gravity=False)
    orc.step(10, 1e-2, 1.0)
    b = orc.state()
    assert rel(st["position"], b["position"]) < 1e-9
    assert rel(st["momentum"], b["momentum"]) < 1e-9
