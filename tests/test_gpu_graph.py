"""Graph replay of world steps: every replay must equal the eager steps bit for bit.

The graph holds two consecutive steps (one round trip of the charge double buffer);
odd step counts end with one eager step. The sequence below replays the graph many
times with host reads and eager steps in between, and changes dt (a re-capture)."""
import numpy as np
import pytest

from exgame_amd import scenes
from exgame_amd.world import DeviceWorld

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_graph_replays_equal_eager_steps(precision):
    scene = scenes.mixed_world(400, seed=13, rigs=True)   # singles, rigs and chains: every kernel class
    eager = DeviceWorld(scene, tol2=1e-20, gravity_precision=precision, jacobian_precision=precision)
    graph = DeviceWorld(scene, tol2=1e-20, gravity_precision=precision, jacobian_precision=precision, graph=True)
    for n, dt in [(5, 1e-2), (4, 1e-2), (1, 1e-2), (6, 1e-2), (3, 5e-3), (2, 1e-2)]:
        eager.step(n, dt, 1.0)
        graph.step(n, dt, 1.0)
        a, b = eager.state(), graph.state()
        for k in ("pose", "momentum", "origin"):
            assert np.array_equal(a[k], b[k]), (n, dt, k)
        assert np.array_equal(eager.charges(), graph.charges())
    np.testing.assert_array_equal(eager.stats(), graph.stats())
    t = graph.last_timing()
    assert t["gravity_ms"] == -1.0 and t["island_ms"] > 0.0


@pytest.mark.parametrize("integrator", ["symplectic_euler", "explicit_euler"])
def test_explicit_graph_replays_equal_eager_steps(integrator):
    # the explicit configuration as the bench replays it (1,024 point masses, dt = 1e-3)
    scene = scenes.plain_gravity(n=1024, seed=0)
    eager = DeviceWorld(scene, integrator=integrator)
    graph = DeviceWorld(scene, integrator=integrator, graph=True)
    for n in (2, 7, 10, 1, 4):
        eager.step(n, 1e-3, 1.0)
        graph.step(n, 1e-3, 1.0)
        a, b = eager.state(), graph.state()
        for k in ("pose", "momentum", "origin"):
            assert np.array_equal(a[k], b[k]), (n, k)
        assert np.array_equal(eager.charges(), graph.charges())


def test_arena_graph_replays_equal_eager_steps_across_admit_and_retire():
    # worlds admitted and retired between replays (the graph is re-captured for the new
    # set of worlds); every state must equal the eager arena's bit for bit
    from exgame_amd.arena import Arena
    worlds = [scenes.mixed_world(n, seed=20 + n, rigs=True) for n in (160, 240, 96)]
    arenas = [Arena(bodies=4096, islands=2048, joints=4096, inverses=1 << 20, tol2=1e-20, graph=g)
              for g in (False, True)]
    ids = [[ar.admit(worlds[0]), ar.admit(worlds[1])] for ar in arenas]
    for n, dt in [(4, 1e-2), (3, 1e-2), (2, 5e-3)]:
        for ar in arenas:
            ar.step(n, dt, 1.0)
            ar.sync()
        for w in range(len(ids[0])):
            a, b = arenas[0].state(ids[0][w]), arenas[1].state(ids[1][w])
            for k in ("pose", "momentum", "origin"):
                assert np.array_equal(a[k], b[k]), (n, dt, w, k)
    for ar, wid in zip(arenas, ids):
        ar.retire(wid[0])
        wid[0] = ar.admit(worlds[2])
        ar.step(4, 1e-2, 1.0)
        ar.sync()
    for w in range(2):
        a, b = arenas[0].state(ids[0][w]), arenas[1].state(ids[1][w])
        for k in ("pose", "momentum", "origin"):
            assert np.array_equal(a[k], b[k]), (w, k)


def test_graph_replays_with_subdividing_free_bodies():
    # free bodies in close encounters (dense ball, dt = 1e-2): some steps subdivide, so
    # those islands go through the requeue list to the eight-lane kernel inside each
    # captured step; every replay must equal the eager steps bit for bit, and equal the
    # eager steps of a world whose free bodies never take the requeue path
    scene = scenes.plain_gravity(n=600, seed=4)
    eager = DeviceWorld(scene, tol2=1e-20)
    graph = DeviceWorld(scene, tol2=1e-20, graph=True)
    for n in (4, 2, 5, 6):
        eager.step(n, 1e-2, 1.0)
        graph.step(n, 1e-2, 1.0)
        a, b = eager.state(), graph.state()
        for k in ("pose", "momentum", "origin"):
            assert np.array_equal(a[k], b[k]), (n, k)
        assert np.array_equal(eager.charges(), graph.charges())
    st = eager.stats()
    assert int(st[:, 0].max()) > 17, "the scene should subdivide some steps"
    np.testing.assert_array_equal(st, graph.stats())
