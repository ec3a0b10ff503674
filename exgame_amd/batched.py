"""The batched form: many uneven worlds advanced together in one arena.

Workload (defaults): 512 worlds, body counts lognormal with median 512 and
σ = ln(16)/3, clipped to [64, 8192] and rounded to a multiple of 8; each world 60 %
free gravitating bodies and 40 % 8-body chains (k = 1e3, c = 10, torsional c = 5);
h = 1e-2, Newton tolerance 1e-10. Prints one JSON line: body-steps/s over a timed
rollout, the device time per step, and the precision used.

    python -m exgame_amd.batched [--worlds 512] [--steps 20] [--precision f64|f32c]
                                 [--jacobian f64|f32c] [--chunk N | --autotune] [--graph]

`--autotune` times the gravity pass under the chunk settings (0 = the single-world
rule, 128 … 8192) and one step under each joint mapping of the 8-body islands (from
a saved state that is put back) on the admitted worlds, then admits the worlds
afresh with the fastest of each; `--graph` replays two steps at a time from a HIP
graph.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

from . import scenes
from .arena import Arena, inverse_doubles


def batch_worlds(n_worlds: int, seed: int = 0, median: float = 512.0, lo: int = 64, hi: int = 8192):
    rng = np.random.default_rng(seed)
    sizes = np.clip(rng.lognormal(math.log(median), math.log(16.0) / 3.0, n_worlds), lo, hi)
    sizes = np.maximum(16, (sizes.astype(np.int64) // 8) * 8)
    return [scenes.mixed_world(int(n), seed=seed * 7919 + w) for w, n in enumerate(sizes)]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--worlds", type=int, default=512)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="f64", choices=["f64", "f32c"], help="gravity arithmetic")
    ap.add_argument("--jacobian", default=None, choices=["f64", "f32c"], help="Jacobian arithmetic (default: as gravity)")
    ap.add_argument("--chunk", type=int, default=1024)
    ap.add_argument("--autotune", action="store_true", help="pick the gravity chunk length by timing")
    ap.add_argument("--graph", action="store_true", help="replay steps from a HIP graph")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    worlds = batch_worlds(args.worlds, args.seed)
    B = sum(w.n_bodies for w in worlds)
    ar = Arena(bodies=B, islands=sum(w.n_islands for w in worlds), joints=sum(w.n_joints for w in worlds) + 1,
               inverses=sum(inverse_doubles(w) for w in worlds), tol2=1e-20, gravity_precision=args.precision,
               gravity_chunk=args.chunk, jacobian_precision=args.jacobian or args.precision, graph=args.graph)
    ids = [ar.admit(w) for w in worlds]
    assert all(i is not None for i in ids)
    if args.autotune:
        ar.step(1, 1e-2, 1.0)     # charges past the bootstrap frame
        ar.sync()
        ar.autotune_gravity()     # both apply to admissions from now on: admit the worlds again
        joint_mapping = ar.autotune_joints()
        for i in ids:
            ar.retire(i)
        ids = [ar.admit(w) for w in worlds]
        assert all(i is not None for i in ids)
    else:
        joint_mapping = "split"
    ar.step(args.warmup, 1e-2, 1.0)
    ar.sync()
    t0 = time.perf_counter()
    ar.step(args.steps, 1e-2, 1.0)
    ar.sync()
    el = time.perf_counter() - t0
    line = {"form": "batched", "worlds": len(worlds), "bodies": B, "steps": args.steps,
            "body_steps_per_s": B * args.steps / el, "ms_per_step": el * 1e3 / args.steps,
            "device_ms_per_step": ar.last_ms() / args.steps, "precision": args.precision,
            "jacobian": args.jacobian or args.precision,
            "gravity_chunk": ar.gravity_chunk, "joint_mapping": joint_mapping, "autotuned": bool(args.autotune),
            "graph": bool(args.graph),
            "islands": sum(w.n_islands for w in worlds), "pairs_per_step": int(sum(w.n_bodies ** 2 for w in worlds))}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
