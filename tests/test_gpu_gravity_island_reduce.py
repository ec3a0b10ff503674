"""The islands add the gravity pass's f64 partial sums themselves (the reduce kernel
is not launched): under Newton for worlds whose islands all have 2 to 8 bodies, under
the explicit integrators for every world. They add each body's
chunks in the reduce kernel's order with its wrench arithmetic, so a world steps to the
same bits with the reduce kernel (exg_world_set_gravity_island_reduce(w, 0)) and
without it, in every on-chip class, with FP32 Jacobian lanes and under graph replay.
The parity tests of these worlds against the oracle run on this path by default."""
import numpy as np
import pytest

from exgame_amd import scenes
from exgame_amd.world import DeviceWorld

pytestmark = pytest.mark.gpu


def same_bits(scene, steps, dt=1e-2, **kw):
    a = DeviceWorld(scene, tol2=1e-20, **kw)
    b = DeviceWorld(scene, tol2=1e-20, gravity_island_reduce=False,
                    **{k: v for k, v in kw.items() if k != "graph"})
    for n in steps:
        a.step(n, dt, 1.0)
        b.step(n, dt, 1.0)
        sa, sb = a.state(), b.state()
        for k in ("pose", "momentum", "origin"):
            assert np.array_equal(sa[k], sb[k]), (n, k)
    assert np.array_equal(a.charges(), b.charges())


def test_headline_world_same_bits_with_and_without_the_reduce_kernel():
    # 4,096 bodies in 512 chains of 8: 16 chunks per tile of 64 targets
    same_bits(scenes.jointed_chains(n_chains=512, m=8, seed=0), (3, 5))


@pytest.mark.parametrize("shape", ["m2", "m4", "braced", "anchored"])
def test_every_on_chip_class_same_bits(shape):
    scene = {"m2": lambda: scenes.jointed_chains(n_chains=300, m=2, seed=2),
             "m4": lambda: scenes.jointed_chains(n_chains=150, m=4, seed=4),
             "braced": lambda: scenes.jointed_chains(n_chains=40, m=8, seed=4, braced=True),
             "anchored": lambda: scenes.anchored_chains(n_chains=60, m=8, seed=6)}[shape]()
    same_bits(scene, (2, 3))


def test_fp32_jacobian_lanes_same_bits():
    same_bits(scenes.jointed_chains(n_chains=128, m=8, seed=5), (4,), jacobian_precision="f32c")


def test_graph_replay_same_bits_as_eager_with_the_reduce_kernel():
    same_bits(scenes.jointed_chains(n_chains=128, m=8, seed=1), (4, 2, 6), graph=True)


@pytest.mark.parametrize("integrator", ["symplectic_euler", "explicit_euler", "lie_euler"])
def test_explicit_integrators_same_bits(integrator):
    # 1,024 point masses (the explicit configuration) and a mixed world with chains and
    # rigs, stepped by the explicit kernel, which adds the partial sums itself
    same_bits(scenes.plain_gravity(n=1024, seed=0), (5, 7), dt=1e-3, integrator=integrator)
    same_bits(scenes.mixed_world(300, seed=4, rigs=True), (4,), dt=1e-3, integrator=integrator, graph=True)
