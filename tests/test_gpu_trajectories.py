"""Long trajectories and the solver's control paths on the device, against the oracle
and against the reference's own physics tests (crates/newton/src/integrator/implicit/newton.rs:923-1578,
crates/newton/src/integrator/explicit/lie_euler.rs:59-182).

Tolerances (f64): per-step states within 1e-12 relative, 100-step trajectories within
1e-9 relative; solver counters identical where the decisions are robust.
"""
import numpy as np
import pytest

import pyoracle
from exgame_amd import scenes
from exgame_amd.world import DeviceWorld
from small_scenes import free_body, spring_pair

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.mark.parametrize("engine", ["mfma", "valu_ordered"])
def test_jointed_chains_100_step_trajectory(engine):
    # the 4,096-body workload's shape (chains of 8, k = 1e3, c = 10, torsional c = 5,
    # N-body gravity, h = 1e-2, tol 1e-10) at 16 chains, 100 steps
    scene = scenes.jointed_chains(n_chains=16, m=8, seed=0)
    dev = DeviceWorld(scene, tol2=1e-20, ns_engine=engine)
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    for _ in range(4):
        dev.step(25, 1e-2, 1.0)
        orc.step(25, 1e-2, 1.0)
        a, b = dev.state(), orc.state()
        assert rel(a["position"], b["position"]) < 1e-9
        assert rel(a["momentum"], b["momentum"]) < 1e-9
    d, o = dev.stats()[:, :6].astype(np.int64).sum(0), orc.stats().sum(0)
    assert np.all(np.abs(d - o) <= np.maximum(2, 0.005 * o)), (d, o)


def test_plain_gravity_explicit_100_step_trajectory():
    # the 1,024-point-mass explicit configuration, all 100 steps (dt = 1e-3)
    scene = scenes.plain_gravity(n=1024, seed=0)
    dev = DeviceWorld(scene, integrator="symplectic_euler")
    orc = pyoracle.OracleWorld(scene, integrator="symplectic_euler")
    dev.step(100, 1e-3, 1.0)
    orc.step(100, 1e-3, 1.0)
    a, b = dev.state(), orc.state()
    assert rel(a["position"], b["position"]) < 1e-9
    assert rel(a["momentum"], b["momentum"]) < 1e-9


@pytest.mark.parametrize("integrator", ["symplectic_euler", "lie_euler", "explicit_euler"])
def test_spinning_rigid_bodies_explicit(integrator):
    # full inertia tensors, spin and drift: exercises the gyroscopic term (symplectic,
    # explicit) and the coadjoint update (Lie) per step
    rng = np.random.default_rng(5)
    n = 40
    pos = rng.uniform(-2, 2, (n, 3))
    rot = rng.normal(size=(n, 3))
    tens = np.stack([np.diag(rng.uniform(0.5, 2.0, 3)) for _ in range(n)])
    vb = np.concatenate([rng.normal(0, 0.5, (n, 3)), rng.normal(0, 3.0, (n, 3))], axis=1)
    scene = scenes._assemble(pos, rot, rng.uniform(0.5, 2.0, n), tens, vb, np.ones(n, dtype=np.int64), [], {})
    dev = DeviceWorld(scene, integrator=integrator)
    orc = pyoracle.OracleWorld(scene, integrator=integrator)
    for _ in range(3):
        dev.step(1, 1e-2, 1.0)
        orc.step(1, 1e-2, 1.0)
        a, b = dev.state(), orc.state()
        assert rel(a["momentum"], b["momentum"]) < 1e-12
        assert rel(a["pose"], b["pose"]) < 1e-12
    dev.step(97, 1e-2, 1.0)
    orc.step(97, 1e-2, 1.0)
    a, b = dev.state(), orc.state()
    assert rel(a["position"], b["position"]) < 1e-9
    assert rel(a["momentum"], b["momentum"]) < 1e-9


def test_lie_euler_world_momentum_on_device():
    # lie_euler.rs:96-106: world angular momentum to round-off at a coarse step
    scene = free_body(1.0, (2.0, 3.0, 4.0), w=(1.0, 1.5, 0.7))
    dev = DeviceWorld(scene, integrator="lie_euler", gravity=False)
    l0 = scene.momentum[0, 3:].copy()   # identity pose: body frame = world frame
    L0 = np.array([2.0 * 1.0, 3.0 * 1.5, 4.0 * 0.7])
    assert np.allclose(l0, L0)
    dev.step(2000, 0.01, 1.0)
    from exgame_amd import pga
    st = dev.state()
    wm = pga.world_momentum(st["pose"], st["momentum"])[0, 3:] + np.cross(st["origin"][0], pga.world_momentum(
        st["pose"], st["momentum"])[0, :3])
    assert np.linalg.norm(wm - L0) < 1e-12 * np.linalg.norm(L0)


K, C, REST, SPEED, SPAN = 5.0e4, 5.0, 2.0, 50.0, 0.5


def stiff_pair(drift=0.0):
    return spring_pair((1, 0, 0), (-1, 0, 0), (drift, SPEED, 0), (drift, -SPEED, 0), rest=REST, k=K, c=C)


def test_stiff_nonlinear_step_matches_a_fine_reference():
    # newton.rs:947-1030: the device's implicit step at dt = 0.002 against the oracle's
    # LieEuler at dt = 1e-4 (an independent scheme); the reference observes 0.065
    ref = pyoracle.OracleWorld(stiff_pair(), integrator="lie_euler", gravity=False)
    ref.step(int(round(SPAN / 1e-4)), 1e-4)
    fine = ref.state()["position"]
    dev = DeviceWorld(stiff_pair(), tol2=1e-8, gravity=False)
    orc = pyoracle.OracleWorld(stiff_pair(), tol2=1e-8, gravity=False)
    n = int(round(SPAN / 0.002))
    dev.step(n, 0.002, 1.0)
    orc.step(n, 0.002, 1.0)
    got = dev.state()["position"]
    assert np.all(np.isfinite(got))
    assert 0.055 < np.abs(got - fine).max() < 0.075
    assert rel(got, orc.state()["position"]) < 1e-9


@pytest.mark.parametrize("dt", [0.5, 0.25, 0.1])
def test_subdivision_takes_the_oracles_spans(dt):
    # the stiff pair subdivides (deepening by the stiffness signal, growing back when
    # aligned); the oracle-order engine must take every decision the oracle takes
    dev = DeviceWorld(stiff_pair(), tol2=1e-8, gravity=False, ns_engine="valu_ordered")
    orc = pyoracle.OracleWorld(stiff_pair(), tol2=1e-8, gravity=False)
    for _ in range(2):
        dev.step(1, dt, 1.0)
        orc.step(1, dt, 1.0)
        np.testing.assert_array_equal(dev.stats()[:, :6], orc.stats())
        assert len(orc.span_trace(0)) > 1
    a, b = dev.state(), orc.state()
    assert rel(a["position"], b["position"]) < 1e-9
    assert rel(a["momentum"], b["momentum"]) < 1e-9


def test_subdivided_step_moves_the_centre_of_mass_by_exactly_its_velocity():
    # Σ span = dt, seen in the centre of mass of a pair with an internal spring
    u = 0.7
    sc = stiff_pair(drift=u)
    com0 = sc.absolute_positions().mean(0)
    dev = DeviceWorld(sc, tol2=1e-20, gravity=False)
    dev.step(2, 0.25, 1.0)
    assert dev.stats()[0, 0] > 6
    com = dev.state()["position"].mean(0)
    assert abs(com[0] - (com0[0] + u * 0.5)) < 1e-5 and abs(com[1] - com0[1]) < 1e-5


def test_a_diverging_step_does_not_poison_later_steps():
    # newton.rs:1272-1330 on the device, step for step against the oracle
    sc = spring_pair((0.05, 0, 0), (-0.05, 0, 0), ma=1 / 18, mb=1 / 18, rest=0.1, k=100.0, c=0.5)
    dev = DeviceWorld(sc, gravity=False, ns_engine="valu_ordered")
    orc = pyoracle.OracleWorld(sc, gravity=False)
    for n, dt in ((10, 1 / 60), (1, 0.1), (60, 1 / 60)):
        dev.step(n, dt, 1.0)
        orc.step(n, dt, 1.0)
    a, b = dev.state(), orc.state()
    p = a["position"]
    assert np.all(np.isfinite(p)) and np.linalg.norm(p[0] - p[1]) < 1.0
    assert rel(p, b["position"]) < 1e-9
    np.testing.assert_array_equal(dev.stats()[:, :6], orc.stats())


def test_accuracy_improves_with_the_iteration_budget():
    # newton.rs:1381-1428 on the device
    sc = spring_pair((1.5, 0, 0), (-1.5, 0, 0), (0, 4.0, 0), (0, -4.0, 0), rest=2.0, k=500.0, c=1.0)

    def run(warm):
        dev = DeviceWorld(sc, ns_warm=warm, gravity=False)
        dev.step(60, 1 / 60, 1.0)
        orc = pyoracle.OracleWorld(sc, ns_warm=warm, gravity=False)
        orc.step(60, 1 / 60, 1.0)
        x = dev.state()["position"][0, 0]
        assert abs(x - orc.state()["position"][0, 0]) < 1e-9
        return x

    generous = run(12)
    errs = [abs(run(b) - generous) for b in (1, 2, 3, 5, 8)]
    assert all(np.isfinite(errs)) and min(errs) < 1e-4


def test_a_stale_warm_hint_is_reseeded_not_followed():
    # newton.rs:1537-1578 on the device: dt alternating 0.2 and 1/600 on k = 5e4
    sc = spring_pair((1, 0, 0), (-1, 0, 0), (0, 5.0, 0), (0, -5.0, 0), rest=2.0, k=5.0e4, c=5.0)
    dev = DeviceWorld(sc, gravity=False, ns_engine="valu_ordered")
    orc = pyoracle.OracleWorld(sc, gravity=False)
    for step in range(40):
        dt = 0.2 if step % 3 == 0 else 1 / 600
        dev.step(1, dt, 1.0)
        orc.step(1, dt, 1.0)
    a = dev.state()["position"]
    assert np.all(np.isfinite(a))
    assert rel(a, orc.state()["position"]) < 1e-9
    assert dev.stats()[0, 3] > 0


def coincident_pair():
    # two bodies at one point joined by an axial spring with zero softening: the
    # spring's direction divides by a zero distance
    import dataclasses
    sc = spring_pair((0, 0, 0), (0, 0, 0), rest=1.0, k=10.0, c=0.0)
    return dataclasses.replace(sc, joint_params=np.array([[0, 0, 0, 0, 0, 0, 1.0, 10.0, 0.0, 0.0]]))


def test_non_finite_explicit_step_is_reported_fatal():
    from exgame_amd import _lib
    dev = DeviceWorld(coincident_pair(), integrator="symplectic_euler", gravity=False)
    dev.step(1, 1e-2, 1.0)
    with pytest.raises(_lib.ExgError) as e:
        dev.sync()
    assert e.value.code == 3                      # EXG_ERR_FATAL, mark-and-continue
    assert dev.stats()[0, 6] == 1 and dev.fatals() == 1
    dev.sync()                                    # reported once
    assert not np.all(np.isfinite(dev.state()["momentum"]))


def test_non_finite_implicit_spans_are_not_committed():
    # Newton never commits a non-finite span (newton.rs:654-660): the state stays as it
    # was, every span at the depth limit is counted, and no step is fatal
    sc = coincident_pair()
    dev = DeviceWorld(sc, gravity=False)
    dev.step(1, 1e-2, 1.0)
    dev.sync()
    st = dev.stats()[0]
    assert st[6] == 0 and st[7] == 256
    assert np.array_equal(dev.state()["pose"], sc.pose)
