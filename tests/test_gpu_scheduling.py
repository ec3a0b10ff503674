"""How the work of one step is laid onto the device: the two joint mappings of the
8-body islands (the same results to rounding; an arena world keeps the mapping it was
admitted with), the joint-mapping autotuner (state put back), island classes launched
side by side (alone vs mixed stays bit-identical, test_gpu_arena.py; here a world of
several classes against the oracle), and the arena's completion marks that keep
serving ticks queued."""
import numpy as np
import pytest

import pyoracle
from exgame_amd import scenes
from exgame_amd.arena import Arena
from exgame_amd.world import DeviceWorld

from test_gpu_parity import rel

pytestmark = pytest.mark.gpu


def big_arena(**kw):
    return Arena(bodies=1 << 15, islands=1 << 13, joints=1 << 14, inverses=1 << 22, **kw)


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_joint_mappings_agree_to_rounding(precision):
    # the same arithmetic in another order of instructions (the fused multiply-adds
    # fall differently): one step within 1e-12, 15 steps within 1e-9, the same solver
    # decisions where they are robust
    scene = scenes.jointed_chains(n_chains=32, m=8, seed=4)
    print(f"32 chains x 8, {precision}, split vs flat", flush=True)
    a = DeviceWorld(scene, tol2=1e-20, jacobian_precision=precision, joint_mapping="split")
    b = DeviceWorld(scene, tol2=1e-20, jacobian_precision=precision, joint_mapping="flat")
    a.step(1, 1e-2, 1.0)
    b.step(1, 1e-2, 1.0)
    assert rel(a.state()["pose"], b.state()["pose"]) < 1e-12
    a.step(14, 1e-2, 1.0)
    b.step(14, 1e-2, 1.0)
    sa, sb = a.state(), b.state()
    assert rel(sa["position"], sb["position"]) < 1e-9 and rel(sa["momentum"], sb["momentum"]) < 1e-9
    d, o = a.stats()[:, :6].astype(np.int64).sum(0), b.stats()[:, :6].astype(np.int64).sum(0)
    assert np.all(np.abs(d - o) <= np.maximum(2, 0.005 * o)), (d, o)


def test_flat_mapping_world_in_an_arena_equals_its_single_world_run():
    # 32 bodies (one gravity chunk either way): a world admitted under the flat mapping
    # runs the flat kernel, bit-identical to the single-world path with that mapping
    scene = scenes.jointed_chains(n_chains=4, m=8, seed=2)
    ar = big_arena(tol2=1e-20)
    ar.set_joint_mapping("flat")
    wid = ar.admit(scene)
    dev = DeviceWorld(scene, tol2=1e-20, joint_mapping="flat")
    ar.step(5, 1e-2, 1.0)
    dev.step(5, 1e-2, 1.0)
    x, y = ar.state(wid), dev.state()
    assert np.array_equal(x["pose"], y["pose"]) and np.array_equal(x["momentum"], y["momentum"])


def test_joint_autotune_leaves_the_worlds_where_they_were():
    ws = [scenes.mixed_world(600, seed=31), scenes.jointed_chains(n_chains=20, m=8, seed=32)]
    print("arena of 2 worlds: autotune joints, then 6 steps against an untuned twin", flush=True)
    tuned, plain = big_arena(tol2=1e-20), big_arena(tol2=1e-20)
    ta = [tuned.admit(w) for w in ws]
    pa = [plain.admit(w) for w in ws]
    tuned.step(2, 1e-2, 1.0)
    plain.step(2, 1e-2, 1.0)
    before = [tuned.state(i) for i in ta]
    choice = tuned.autotune_joints()
    assert choice in ("split", "flat")
    for i, s in zip(ta, before):
        t = tuned.state(i)
        assert np.array_equal(t["pose"], s["pose"]) and np.array_equal(t["momentum"], s["momentum"])
    tuned.step(6, 1e-2, 1.0)
    plain.step(6, 1e-2, 1.0)
    for i, j in zip(ta, pa):
        x, y = tuned.state(i), plain.state(j)
        assert np.array_equal(x["pose"], y["pose"]) and np.array_equal(x["origin"], y["origin"])
        assert np.array_equal(tuned.stats(i), plain.stats(j))


def test_classes_side_by_side_match_the_oracle():
    # singletons, 4-body rigs and chains of 8 in one world: each class on its own
    # stream within the step
    scene = scenes.mixed_world(160, seed=41, rigs=True)
    sizes = np.diff(scene.island_start)
    assert {1, 4, 8} <= set(sizes.tolist())
    print("mixed world with rigs: world path vs oracle, 10 steps", flush=True)
    dev = DeviceWorld(scene, tol2=1e-20)
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    dev.step(10, 1e-2, 1.0)
    orc.step(10, 1e-2, 1.0)
    a, b = dev.state(), orc.state()
    assert rel(a["position"], b["position"]) < 1e-9
    assert rel(a["momentum"], b["momentum"]) < 1e-9


def test_completion_marks():
    ar = big_arena(tol2=1e-20)
    wid = ar.admit(scenes.jointed_chains(n_chains=8, m=8, seed=5))
    t0 = ar.mark()
    ar.step(3, 1e-2, 1.0)
    t1 = ar.mark()
    assert t1 == t0 + 1
    ar.mark_wait(t1)
    assert ar.mark_done(t0) and ar.mark_done(t1)
    # more marks than the ring holds: old tickets read as done
    ts = [ar.mark() for _ in range(20)]
    ar.mark_wait(ts[-1])
    assert all(ar.mark_done(t) for t in [t0, t1] + ts)
    from exgame_amd import _lib
    with pytest.raises(_lib.ExgError):
        ar.mark_wait(ts[-1] + 5)
    ar.sync()
    assert ar.state(wid)["pose"].shape[0] == 64
