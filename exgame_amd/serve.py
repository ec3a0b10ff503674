"""Serving: trajectory-rollout requests for worlds of different sizes, answered with
continuous batching on one device.

Requests arrive over time (a Poisson process in the synthetic load). Each names a
world snapshot and a number of implicit steps. The server keeps every admitted
world resident in an `Arena`, admits waiting requests in arrival order whenever
the HBM budget and the arena's ranges allow, and advances all admitted worlds
together, a few steps per tick, so that work that arrives while others run
overlaps with them on the device. A request finishes after exactly its step count
and is retired, freeing room for the next. Because an arena world's trajectory does
not depend on its neighbours, a request's result is the same alone and in the mix.

The reference runs rollouts one Driver at a time (crates/newton/src/driver.rs);
this loop is a serving layer (Poisson arrivals, admission by HBM footprint,
continuous batching, p50/p99 latency) on top of the batched step.
"""
from __future__ import annotations

import dataclasses
import json
import math
import time

import numpy as np

from .arena import Arena, footprint, inverse_doubles


@dataclasses.dataclass
class Request:
    rid: int
    scene: object
    steps: int
    arrival: float                   # seconds after the start of the run
    admitted: float = math.nan
    finished: float = math.nan
    result: dict | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def synthetic_requests(n: int, rate: float, seed: int = 0, bodies=(256, 32768), steps=(200, 2000),
                       median_bodies: float = 2048.0):
    """`n` requests with exponential inter-arrival times (mean 1/rate), lognormal world
    sizes clipped to `bodies` (60% free bodies, 40% chains and 4-body rigs), step
    counts uniform in `steps`."""
    from . import scenes
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        nb = int(np.clip(rng.lognormal(math.log(median_bodies), math.log(16.0) / 3.0), bodies[0], bodies[1]))
        nb = max(16, (nb // 8) * 8)
        scene = scenes.mixed_world(nb, seed=seed * 100003 + rid, rigs=True)
        k = int(rng.integers(steps[0], steps[1] + 1))
        out.append(Request(rid, scene, k, t))
    return out


def arena_for(requests, hbm_budget: int, gravity_chunk: int = 1024, **solver) -> Arena:
    """An arena whose ranges hold the busiest plausible mix within `hbm_budget`."""
    per_body = max(footprint(r.scene, gravity_chunk) / r.scene.n_bodies for r in requests)
    bodies = int(hbm_budget / per_body)
    biggest = max(r.scene.n_bodies for r in requests)
    bodies = max(bodies, biggest)
    isl = max(r.scene.n_islands / r.scene.n_bodies for r in requests)
    jnt = max(r.scene.n_joints / r.scene.n_bodies for r in requests)
    inv = max(inverse_doubles(r.scene) / r.scene.n_bodies for r in requests)
    return Arena(bodies=bodies, islands=int(bodies * isl) + 64, joints=int(bodies * jnt) + 64,
                 inverses=int(bodies * inv) + 4096, gravity_chunk=gravity_chunk, **solver)


class Server:
    def __init__(self, arena: Arena, hbm_budget: int, dt: float = 1e-2, tick_steps: int = 20, keep_results=False):
        self.arena = arena
        self.budget = int(hbm_budget)
        self.dt = dt
        self.tick = int(tick_steps)
        self.keep = keep_results

    # ticks kept queued on the device while the host waits for the oldest one
    DEPTH = 2

    def run(self, requests, clock=time.perf_counter, sleep=time.sleep):
        """Serve `requests` (any order); returns a report dict. Latency is completion
        time minus arrival time on `clock`.

        Ticks are pipelined: while the set of worlds stays the same, the next tick is
        enqueued before the previous one has finished (up to DEPTH in flight, waited
        for by completion marks), so the device does not idle between ticks. The set
        changes — worlds retire when their last tick has run, waiting requests are
        admitted — only with nothing in flight, since admission rebuilds the arena's
        tables."""
        arrivals = sorted(requests, key=lambda r: r.arrival)
        queue: list[Request] = []
        active: dict[int, list] = {}       # world id → [request, steps not yet enqueued, bytes]
        inflight: list[int] = []           # completion-mark tickets, oldest first
        used = 0
        body_steps = 0
        nxt = 0
        t0 = clock()
        busy = 0.0
        busy_from = None
        done = []
        full = False          # the arena refused the queue's head; retry after a retirement

        def fits(r):
            return not full and (used + footprint(r.scene, self.arena.gravity_chunk) <= self.budget or not active)

        def drain():
            nonlocal busy, busy_from
            if inflight:
                self.arena.sync()
                inflight.clear()
            if busy_from is not None:
                busy += clock() - busy_from
                busy_from = None

        while nxt < len(arrivals) or queue or active or inflight:
            now = clock() - t0
            while nxt < len(arrivals) and arrivals[nxt].arrival <= now:
                queue.append(arrivals[nxt])
                nxt += 1
            finishing = [w for w, v in active.items() if v[1] <= 0]
            if finishing or (queue and fits(queue[0])):
                drain()
                for wid in finishing:
                    full = False
                    r, _, need = active.pop(wid)
                    if self.keep:
                        r.result = self.arena.state(wid)
                    self.arena.retire(wid)
                    used -= need
                    r.finished = clock() - t0
                    done.append(r)
                # admission in arrival order, by HBM footprint
                while queue and fits(queue[0]):
                    r = queue[0]
                    need = footprint(r.scene, self.arena.gravity_chunk)
                    wid = self.arena.admit(r.scene)
                    if wid is None:
                        if not active:
                            raise RuntimeError(f"request {r.rid} ({r.scene.n_bodies} bodies) does not fit the arena")
                        full = True
                        break
                    queue.pop(0)
                    r.admitted = clock() - t0
                    active[wid] = [r, r.steps, need]
                    used += need
            if not active:
                drain()
                if nxt < len(arrivals):
                    sleep(max(0.0, arrivals[nxt].arrival - (clock() - t0)))
                continue
            if len(inflight) < self.DEPTH:
                k = min(self.tick, min(v[1] for v in active.values()))
                self.arena.step(k, self.dt, 1.0)
                inflight.append(self.arena.mark())
                if busy_from is None:
                    busy_from = clock()
                body_steps += k * sum(v[0].scene.n_bodies for v in active.values())
                for v in active.values():
                    v[1] -= k
                continue
            self.arena.mark_wait(inflight.pop(0))
        drain()
        wall = clock() - t0
        lat = np.array([r.latency for r in done])
        return {
            "requests": len(done),
            "wall_s": wall,
            "device_busy_s": busy,
            "body_steps": int(body_steps),
            "body_steps_per_s": body_steps / wall if wall > 0 else 0.0,
            "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else math.nan,
            "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else math.nan,
            "latency_max_s": float(lat.max()) if len(lat) else math.nan,
        }


def main() -> int:
    import argparse
    ap = argparse.ArgumentParser(description="Serve synthetic rollout requests (continuous batching).")
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--rate", type=float, default=8.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--min-bodies", type=int, default=256)
    ap.add_argument("--max-bodies", type=int, default=32768)
    ap.add_argument("--median-bodies", type=float, default=2048.0)
    ap.add_argument("--min-steps", type=int, default=200)
    ap.add_argument("--max-steps", type=int, default=2000)
    ap.add_argument("--hbm-gib", type=float, default=16.0)
    ap.add_argument("--tick", type=int, default=20)
    ap.add_argument("--precision", default="f32c", choices=["f64", "f32c"])
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    reqs = synthetic_requests(args.requests, args.rate, args.seed, (args.min_bodies, args.max_bodies),
                              (args.min_steps, args.max_steps), args.median_bodies)
    budget = int(args.hbm_gib * (1 << 30))
    ar = arena_for(reqs, budget, tol2=1e-20, gravity_precision=args.precision, jacobian_precision=args.precision)
    rep = Server(ar, budget, tick_steps=args.tick).run(reqs)
    rep.update({"rate_per_s": args.rate, "precision": args.precision, "hbm_budget_gib": args.hbm_gib,
                "bodies_range": [args.min_bodies, args.max_bodies], "steps_range": [args.min_steps, args.max_steps]})
    print(json.dumps(rep))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
