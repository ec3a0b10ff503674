"""Kinematic bodies on the device against the oracle. A body of mass 0 is
Inertia::Kinematic (crates/newton/src/inertia.rs:103-151): its momentum slot holds its
body-frame velocity; the implicit step keeps it out of the system and the commit
(crates/newton/src/integrator/implicit/newton.rs:806-830, 704-753) while its velocity
drives its joints; the explicit integrators leave it (symplectic_euler.rs:27-29,
explicit_euler.rs:22-24). The kernels take an island's kinematic bodies after its
dynamic ones. Tolerances as for the other parity tests: a step within 1e-12
relative, trajectories within 1e-9."""
import numpy as np
import pytest

import pyoracle
from exgame_amd import scenes
from exgame_amd.arena import Arena
from exgame_amd.world import DeviceWorld
from test_gpu_parity import ENGINES, rel, run_pair

pytestmark = pytest.mark.gpu


def with_kinematic(scene):
    """Kinematic bodies in every island class of a mixed world: every fifth free body,
    the last body of every other rig and chain, and the whole of the second chain (an
    island with no unknowns). Their tensors are zeroed (the kernels never read them)."""
    mass, tens = scene.mass.copy(), scene.inertia.copy()
    starts = scene.island_start
    sizes = np.diff(starts)
    singles = np.flatnonzero(sizes == 1)
    for i in singles[::5]:
        mass[starts[i]] = 0.0
    for size in (4, 8):
        isl = np.flatnonzero(sizes == size)
        for i in isl[::2]:
            mass[starts[i + 1] - 1] = 0.0
    first_chain = np.flatnonzero(sizes == 8)[1]
    mass[starts[first_chain]:starts[first_chain + 1]] = 0.0
    tens[mass == 0.0] = 0.0
    object.__setattr__(scene, "mass", mass)
    object.__setattr__(scene, "inertia", tens)
    return scene


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("m", [2, 4, 8])
def test_anchored_chains_first_steps_match_oracle(engine, m):
    # m = 2, 4: the 4-body class; m = 8 (14 joints): the two-islands-per-CU class
    scene = scenes.anchored_chains(n_chains=6, m=m, seed=m)
    dev, orc = run_pair(scene, 2, ns_engine=engine)
    a, b = dev.state(), orc.state()
    assert rel(a["position"], b["position"]) < 1e-12
    assert rel(a["momentum"], b["momentum"]) < 1e-11
    np.testing.assert_array_equal(dev.stats()[:, :6], orc.stats())


def test_anchors_never_move_and_drive_the_chains():
    scene = scenes.anchored_chains(n_chains=16, m=8, seed=3)
    kin = scene.mass == 0.0
    dev = DeviceWorld(scene, tol2=1e-20)
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    s0 = dev.state()
    dev.step(20, 1e-2, 1.0)
    orc.step(20, 1e-2, 1.0)
    a, b = dev.state(), orc.state()
    # the momentum slot (the prescribed velocity) bit for bit; the absolute pose to the
    # rounding of the island re-anchoring
    np.testing.assert_array_equal(a["momentum"][kin], s0["momentum"][kin])
    np.testing.assert_allclose(a["position"][kin], s0["position"][kin], rtol=0, atol=1e-12)
    np.testing.assert_allclose(a["pose"][kin][:, :4], s0["pose"][kin][:, :4], rtol=0, atol=1e-15)
    assert rel(a["position"], b["position"]) < 1e-9
    assert rel(a["momentum"], b["momentum"]) < 1e-9
    assert rel(dev.charges(), orc.front()) < 1e-9
    # the anchors' velocities reach the chains: with resting anchors they end elsewhere
    rest = DeviceWorld(scenes.anchored_chains(n_chains=16, m=8, seed=3, anchor_speed=0.0), tol2=1e-20)
    rest.step(20, 1e-2, 1.0)
    assert np.abs(rest.state()["position"][~kin] - a["position"][~kin]).max() > 1e-6


def test_kinematic_bodies_in_every_class_match_oracle():
    scene = with_kinematic(scenes.mixed_world(160, seed=9, rigs=True))
    kin = scene.mass == 0.0
    dev = DeviceWorld(scene, tol2=1e-20)
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    s0 = dev.state()
    dev.step(1, 1e-2, 1.0)
    orc.step(1, 1e-2, 1.0)
    # the same decisions (an island with no unknowns: one span, one converged iteration)
    np.testing.assert_array_equal(dev.stats()[:, :6], orc.stats())
    dev.step(2, 1e-2, 1.0)
    orc.step(2, 1e-2, 1.0)
    a, b = dev.state(), orc.state()
    assert rel(a["position"], b["position"]) < 1e-12
    assert rel(a["momentum"], b["momentum"]) < 1e-11
    np.testing.assert_array_equal(a["momentum"][kin], s0["momentum"][kin])
    np.testing.assert_allclose(a["position"][kin], s0["position"][kin], rtol=0, atol=1e-12)


@pytest.mark.parametrize("integrator", ["symplectic_euler", "explicit_euler"])
def test_explicit_integrators_leave_kinematic_bodies(integrator):
    scene = with_kinematic(scenes.mixed_world(160, seed=9, rigs=True))
    kin = scene.mass == 0.0
    dev = DeviceWorld(scene, integrator=integrator)
    orc = pyoracle.OracleWorld(scene, integrator=integrator)
    s0 = dev.state()
    dev.step(10, 1e-3, 1.0)
    orc.step(10, 1e-3, 1.0)
    a, b = dev.state(), orc.state()
    assert rel(a["position"], b["position"]) < 1e-12
    assert rel(a["momentum"], b["momentum"]) < 1e-12
    np.testing.assert_array_equal(a["momentum"][kin], s0["momentum"][kin])


def test_arena_world_with_kinematic_bodies_matches_oracle():
    scene = with_kinematic(scenes.mixed_world(160, seed=9, rigs=True))
    ar = Arena(bodies=1 << 15, islands=1 << 13, joints=1 << 14, inverses=1 << 22, tol2=1e-20)
    ar.admit(scenes.mixed_world(300, seed=2))
    wid = ar.admit(scene)
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    ar.step(3, 1e-2, 1.0)
    orc.step(3, 1e-2, 1.0)
    a, b = ar.state(wid), orc.state()
    assert rel(a["position"], b["position"]) < 1e-12
    assert rel(a["momentum"], b["momentum"]) < 1e-11


def test_mixed_precision_with_anchors_stays_within_the_gate():
    # FP32 gravity pairs (near field in FP64) and FP32 Jacobian lanes; the residual,
    # Newton–Schulz and the state stay FP64
    scene = scenes.anchored_chains(n_chains=16, m=8, seed=5)
    dev = DeviceWorld(scene, gravity_precision="f32c", jacobian_precision="f32c")
    orc = pyoracle.OracleWorld(scene)
    dev.step(20, 1e-2, 1.0)
    orc.step(20, 1e-2, 1.0)
    assert rel(dev.state()["position"], orc.state()["position"]) < 1e-9


@pytest.mark.parametrize("m", [12, 20])
def test_anchored_long_chains_on_the_mechanism_sized_kernel(m):
    # 12 and 20 bodies (22 and 38 joints): the one-workgroup-per-island kernel with its
    # system in HBM, an anchor last in each island
    scene = scenes.anchored_chains(n_chains=4, m=m, seed=m)
    kin = scene.mass == 0.0
    print(f"anchored chains m={m}", flush=True)
    dev = DeviceWorld(scene, tol2=1e-20)
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    s0 = dev.state()
    dev.step(1, 1e-2, 1.0)
    orc.step(1, 1e-2, 1.0)
    a, b = dev.state(), orc.state()
    assert rel(a["pose"], b["pose"]) < 1e-12
    assert rel(a["momentum"], b["momentum"]) < 1e-12
    dev.step(9, 1e-2, 1.0)
    orc.step(9, 1e-2, 1.0)
    a, b = dev.state(), orc.state()
    assert rel(a["position"], b["position"]) < 1e-9
    assert rel(a["momentum"], b["momentum"]) < 1e-9
    np.testing.assert_array_equal(a["momentum"][kin], s0["momentum"][kin])
    np.testing.assert_allclose(a["position"][kin], s0["position"][kin], rtol=0, atol=1e-12)
    assert dev.fatals() == 0


@pytest.mark.parametrize("m", [4, 8])
def test_kinematic_bodies_anywhere_in_their_island(m):
    # the anchor first in each island: DeviceWorld uploads the dynamic bodies first and
    # hands the state back in the scene's order; against the oracle stepping the scene
    # as it is (the reference's order map, newton.rs:806-830)
    scene = scenes.anchored_chains(n_chains=8, m=m, seed=7, kinematic_first=True)
    kin = scene.mass == 0.0
    assert kin[0] and kin[m]
    dev, orc = run_pair(scene, 3)
    a, b = dev.state(), orc.state()
    assert rel(a["position"], b["position"]) < 1e-12
    assert rel(a["momentum"], b["momentum"]) < 1e-11
    assert rel(dev.charges(), orc.front()) < 1e-12
    np.testing.assert_array_equal(a["momentum"][kin], scene.momentum[kin])


def test_arena_takes_kinematic_bodies_anywhere():
    scene = scenes.anchored_chains(n_chains=8, m=8, seed=8, kinematic_first=True)
    ar = Arena(bodies=1 << 15, islands=1 << 13, joints=1 << 14, inverses=1 << 22, tol2=1e-20)
    ar.admit(scenes.mixed_world(200, seed=3))
    wid = ar.admit(scene)
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    ar.step(3, 1e-2, 1.0)
    orc.step(3, 1e-2, 1.0)
    a, b = ar.state(wid), orc.state()
    assert rel(a["position"], b["position"]) < 1e-12
    assert rel(a["momentum"], b["momentum"]) < 1e-11


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_zero_size_anchors_in_both_gravity_precisions(precision):
    # the reference's kinematic body has effective_size(0) = 0: as a pair's higher slot its
    # core is 0 (1/s infinite; the law's min(1/d, 1/s) takes 1/d), and its zero charge
    # pulls nothing; the mixed-precision records split 1/s = ∞ into hi + lo without harm
    scene = scenes.anchored_chains(n_chains=16, m=8, seed=9)
    size = scene.size.copy()
    size[scene.mass == 0.0] = 0.0
    object.__setattr__(scene, "size", size)
    dev = DeviceWorld(scene, tol2=1e-20, gravity_precision=precision, jacobian_precision=precision)
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    dev.step(20, 1e-2, 1.0)
    orc.step(20, 1e-2, 1.0)
    a, b = dev.state(), orc.state()
    assert np.isfinite(a["pose"]).all() and dev.fatals() == 0
    assert rel(a["position"], b["position"]) < (1e-12 if precision == "f64" else 1e-9)
