"""Host logic of the serving loop (admission order, HBM budget, exact step counts,
latency percentiles) against a stand-in arena on a virtual clock — no device."""
import numpy as np

from exgame_amd import scenes
from exgame_amd.serve import Request, Server


class FakeArena:
    """Records the calls the server makes; steps cost 1 ms of virtual time."""
    gravity_chunk = 1024

    def __init__(self, clock, max_worlds):
        self.clock = clock
        self.max_worlds = max_worlds
        self.live = {}
        self.steps = {}
        self.next = 0
        self.log = []

    def admit(self, scene):
        if len(self.live) >= self.max_worlds:
            return None
        wid = self.next
        self.next += 1
        self.live[wid] = scene
        self.steps[wid] = 0
        self.log.append(("admit", wid))
        return wid

    def step(self, k, dt, warp):
        for w in self.live:
            self.steps[w] += k
        self.clock.t += 1e-3 * k
        self.log.append(("step", k, tuple(sorted(self.live))))

    def sync(self):
        self.log.append(("sync",))

    def mark(self):
        self.log.append(("mark",))
        return sum(1 for e in self.log if e[0] == "mark") - 1

    def mark_wait(self, ticket):
        self.log.append(("wait", ticket))

    def state(self, wid):
        return {"steps": self.steps[wid]}

    def retire(self, wid):
        del self.live[wid]
        self.log.append(("retire", wid))


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, s):
        self.t += s


def make_requests():
    sc = scenes.jointed_chains(n_chains=2, m=8, seed=0)
    return [Request(0, sc, 30, 0.0), Request(1, sc, 50, 0.005), Request(2, sc, 10, 0.006),
            Request(3, sc, 20, 0.5)]


def test_every_request_runs_exactly_its_steps_in_arrival_order():
    clk = Clock()
    fa = FakeArena(clk, max_worlds=2)
    reqs = make_requests()
    rep = Server(fa, hbm_budget=1 << 40, tick_steps=7, keep_results=True).run(reqs, clock=clk, sleep=clk.sleep)
    assert rep["requests"] == 4
    for r in reqs:
        assert r.result["steps"] == r.steps
        assert r.finished >= r.admitted >= r.arrival
    admits = [e[1] for e in fa.log if e[0] == "admit"]
    assert admits == [0, 1, 2, 3]           # FIFO
    # at most two worlds in flight; request 2 waits for a slot
    assert max(len(e[2]) for e in fa.log if e[0] == "step") == 2
    assert reqs[2].admitted > reqs[2].arrival
    # continuous batching: request 1 joined while request 0 was running
    first_two = [e for e in fa.log if e[0] == "step" and e[2] == (0, 1)]
    assert first_two


def test_hbm_budget_limits_admission_and_latency_report():
    clk = Clock()
    fa = FakeArena(clk, max_worlds=10)
    reqs = make_requests()
    from exgame_amd.arena import footprint
    one = footprint(reqs[0].scene)
    rep = Server(fa, hbm_budget=one, tick_steps=100).run(reqs, clock=clk, sleep=clk.sleep)
    assert max(len(e[2]) for e in fa.log if e[0] == "step") == 1   # budget holds one world
    lat = np.array([r.latency for r in reqs])
    assert abs(rep["latency_p50_s"] - np.percentile(lat, 50)) < 1e-12
    assert abs(rep["latency_p99_s"] - np.percentile(lat, 99)) < 1e-12
    assert rep["body_steps"] == sum(r.steps * r.scene.n_bodies for r in reqs)


def test_ticks_stay_queued_while_the_set_of_worlds_is_unchanged():
    # one long request: every tick is enqueued behind a completion mark, the host
    # waits for the oldest of two in flight, and a full sync comes only when the
    # world retires
    clk = Clock()
    fa = FakeArena(clk, max_worlds=4)
    sc = scenes.jointed_chains(n_chains=2, m=8, seed=0)
    rep = Server(fa, hbm_budget=1 << 40, tick_steps=10, keep_results=True).run([Request(0, sc, 95, 0.0)], clock=clk,
                                                                             sleep=clk.sleep)
    steps = [e for e in fa.log if e[0] == "step"]
    assert [e[1] for e in steps] == [10] * 9 + [5]
    syncs = [i for i, e in enumerate(fa.log) if e[0] == "sync"]
    assert len(syncs) == 1 and syncs[0] > fa.log.index(steps[-1])
    waits = [e for e in fa.log if e[0] == "wait"]
    assert [e[1] for e in waits] == list(range(len(waits)))      # oldest first
    assert len(waits) >= 8
    assert rep["body_steps"] == 95 * sc.n_bodies
