"""UniformField on the device (exg_world_set_uniform_field; crates/newton/src/field.rs:16-44)
against the oracle: F = m·a and zero torque added after gravity, in every island class
and integrator; a zero field leaves every bit as without one."""
import numpy as np
import pytest

import pyoracle
from exgame_amd import scenes
from exgame_amd.world import DeviceWorld
from test_gpu_parity import rel
from test_gpu_kinematic import with_kinematic

pytestmark = pytest.mark.gpu

ACCEL = (0.3, -0.2, -1.0)


def field_pair(scene, steps, dt=1e-2, integrator="newton", tol2=1e-20):
    dev = DeviceWorld(scene, tol2=tol2, integrator=integrator, uniform_field=ACCEL)
    orc = pyoracle.OracleWorld(scene, tol2=tol2, integrator=integrator)
    orc.set_uniform(ACCEL)
    dev.step(steps, dt, 1.0)
    orc.step(steps, dt, 1.0)
    return dev.state(), orc.state()


@pytest.mark.parametrize("shape", ["mixed_kinematic", "chains", "cloth"])
def test_field_under_newton_matches_oracle(shape):
    # free bodies, rigs, chains and kinematic bodies; chains alone (the islands add the
    # gravity partial sums); cloths on the mechanism-sized kernel
    scene = {"mixed_kinematic": lambda: with_kinematic(scenes.mixed_world(160, seed=9, rigs=True)),
             "chains": lambda: scenes.jointed_chains(n_chains=64, m=8, seed=2),
             "cloth": lambda: scenes.cloth_grids(shapes=((4, 4), (3, 5)))}[shape]()
    a, b = field_pair(scene, 1 if shape == "cloth" else 3)
    assert rel(a["position"], b["position"]) < 1e-12
    assert rel(a["momentum"], b["momentum"]) < 1e-11


@pytest.mark.parametrize("integrator", ["symplectic_euler", "explicit_euler"])
def test_field_under_explicit_integrators_matches_oracle(integrator):
    a, b = field_pair(with_kinematic(scenes.mixed_world(160, seed=9, rigs=True)), 10, dt=1e-3, integrator=integrator)
    assert rel(a["position"], b["position"]) < 1e-12
    assert rel(a["momentum"], b["momentum"]) < 1e-12


def test_zero_field_keeps_every_bit():
    scene = scenes.mixed_world(160, seed=9, rigs=True)
    a = DeviceWorld(scene, tol2=1e-20)
    b = DeviceWorld(scene, tol2=1e-20, uniform_field=(0.0, 0.0, 0.0))
    a.step(3, 1e-2, 1.0)
    b.step(3, 1e-2, 1.0)
    for k in ("pose", "momentum", "origin"):
        assert np.array_equal(a.state()[k], b.state()[k])
