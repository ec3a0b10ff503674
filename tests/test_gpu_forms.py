"""The batched and served forms at their workload sizes, in mixed precision: 512
uneven worlds (60 % free bodies, 40 % chains of 8) in one arena, and served requests
of 256–32,768 bodies. A sample of worlds must end bit-identical to their alone runs,
and worlds small enough for the oracle must stay within the mixed-precision gate
(final-state relative error ≤ 1e-9 against the f64 oracle)."""
import numpy as np
import pytest

import pyoracle
from exgame_amd.arena import Arena, footprint, inverse_doubles
from exgame_amd.batched import batch_worlds

from test_gpu_parity import rel

pytestmark = pytest.mark.gpu

MIXED = dict(tol2=1e-20, gravity_precision="f32c", jacobian_precision="f32c")


def arena_for_worlds(worlds, chunk):
    return Arena(bodies=sum(w.n_bodies for w in worlds), islands=sum(w.n_islands for w in worlds),
                 joints=sum(w.n_joints for w in worlds) + 1, inverses=sum(inverse_doubles(w) for w in worlds),
                 gravity_chunk=chunk, **MIXED)


def same_state(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("pose", "momentum", "origin"))


def test_batched_512_worlds_mixed_precision():
    worlds = batch_worlds(512, seed=0)
    assert len(worlds) == 512 and sum(w.n_bodies for w in worlds) > 300_000
    ar = arena_for_worlds(worlds, 1024)
    ids = [ar.admit(w) for w in worlds]
    assert all(i is not None for i in ids)
    steps = 6
    ar.step(steps, 1e-2, 1.0)
    ar.sync()
    sizes = np.array([w.n_bodies for w in worlds])
    sample = [int(np.argmin(sizes)), int(np.argmax(sizes)), int(np.argsort(sizes)[len(sizes) // 2])]
    for k in sample:
        alone = arena_for_worlds([worlds[k]], 1024)
        w = alone.admit(worlds[k])
        alone.step(steps, 1e-2, 1.0)
        assert same_state(ar.state(ids[k]), alone.state(w)), k
        alone.close()
    # the smallest worlds of the mix against the f64 oracle
    for k in np.argsort(sizes)[:2]:
        orc = pyoracle.OracleWorld(worlds[k], tol2=1e-20)
        orc.step(steps, 1e-2, 1.0)
        got, want = ar.state(ids[k]), orc.state()
        assert rel(got["position"], want["position"]) < 1e-9, k
        assert np.all(np.isfinite(got["momentum"]))


def test_served_requests_at_workload_sizes_mixed_precision():
    from exgame_amd.serve import Server, arena_for, synthetic_requests
    reqs = synthetic_requests(8, rate=50.0, seed=5, bodies=(256, 32768), steps=(20, 60), median_bodies=2048.0)
    sizes = [r.scene.n_bodies for r in reqs]
    assert min(sizes) >= 256 and max(sizes) <= 32768
    reqs[0].scene = __import__("exgame_amd.scenes", fromlist=["mixed_world"]).mixed_world(32768, seed=77, rigs=True)
    budget = int(0.5 * sum(footprint(r.scene, 1024) for r in reqs))
    ar = arena_for(reqs, budget, **MIXED)
    rep = Server(ar, budget, tick_steps=10, keep_results=True).run(reqs)
    assert rep["requests"] == len(reqs)
    assert rep["latency_p99_s"] >= rep["latency_p50_s"] > 0 and rep["body_steps_per_s"] > 0
    for r in (reqs[0], reqs[-1]):
        alone = arena_for([r], footprint(r.scene, 1024) + (1 << 20), **MIXED)
        w = alone.admit(r.scene)
        alone.step(r.steps, 1e-2, 1.0)
        assert same_state(alone.state(w), r.result), r.rid
        alone.close()
