"""The precision form: FP32 arithmetic where it is safe, checked against FP64.

Gravity: N point masses (positions uniform in the unit ball, masses lognormal(0, 0.5)
normalised to 1, G = 1, the size-core law) — forces with FP32 pairs and two-sum
compensated FP32 accumulators against the FP64 forces of the same charges: max and
RMS per-body relative error of the force vectors, and the interaction rate of both
kernels (device time of the gravity pass inside a stepped world).

Jointed: the 4,096-body chains workload scaled to `--chains` chains of 8 (65,536
bodies by default): the same rollout in FP64 and with FP32 Jacobian lanes and
compensated FP32 assembly (residual, Newton–Schulz, update in FP64), with and without
FP32 gravity — body-steps/s and the final-state relative error against the FP64
rollout (positions and momenta separately). The FP64 rollout itself is the device
path that the tests pin against the oracle at smaller sizes (the oracle cannot step
65,536 bodies for 200 steps in reasonable time).

    python -m exgame_amd.precision [--gravity-n 262144] [--chains 8192] [--steps 200] [--radius R]

The chain origins fill a ball of radius 50·(chains/512)^(1/3) unless --radius is given:
the 4,096-body workload's density at any size (126 at 8,192 chains).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import scenes
from .world import DeviceWorld, gravity_accumulate

H = 1e-2
TOL = 1e-10


def gravity_form(n: int, seed: int = 0, timing_steps: int = 3) -> dict:
    scene = scenes.plain_gravity(n=n, seed=seed)
    # charges as a step publishes them (the first step runs on the zero charges)
    w = DeviceWorld(scene, integrator="symplectic_euler")
    w.step(1, 1e-3, 1.0)
    w.sync()
    charges = w.charges()
    w.close()
    f64 = gravity_accumulate(charges, 1.0, "f64")[:, :3]
    f32 = gravity_accumulate(charges, 1.0, "f32c")[:, :3]
    mag = np.linalg.norm(f64, axis=1)
    err = np.linalg.norm(f32 - f64, axis=1) / np.maximum(mag, 1e-300)
    out = {"bodies": n, "max_rel_error": float(err.max()), "rms_rel_error": float(np.sqrt(np.mean(err ** 2))),
           "norm_rel_error": float(np.linalg.norm(f32 - f64) / np.linalg.norm(f64))}
    pairs = float(n) * (n - 1)
    for prec in ("f64", "f32c"):
        w = DeviceWorld(scene, integrator="symplectic_euler", gravity_precision=prec)
        w.step(2, 1e-3, 1.0)      # past the bootstrap frame and the code-object load
        w.sync()
        w.step(timing_steps, 1e-3, 1.0)
        w.sync()
        g_s = w.last_timing()["gravity_ms"] * 1e-3 / timing_steps
        w.close()
        out[prec] = {"ms_per_pass": g_s * 1e3, "interactions_per_s": pairs / g_s,
                     "tflops_at_20_per_interaction": 20.0 * pairs / g_s / 1e12}
    return out


def _rel(a, b) -> float:
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def jointed_form(chains: int, steps: int, radius: float = 50.0) -> dict:
    """FP64 against mixed precision on the same rollout, after one step and after
    `steps`. A dense many-body rollout is chaotic: any change of rounding grows, so the
    same rollout in FP64 with the gravity sums taken in another (equally valid) order
    — an arena with another chunk length — is reported beside it as the floor."""
    scene = scenes.jointed_chains(n_chains=chains, m=8, seed=0, radius=radius)
    body_steps = scene.n_bodies * steps
    runs = {}
    ref1 = ref = None
    for name, gp, jp in [("f64", "f64", "f64"), ("f32_jacobian", "f64", "f32c"),
                         ("f32_jacobian_f32c_gravity", "f32c", "f32c")]:
        w = DeviceWorld(scene, tol2=TOL * TOL, gravity_precision=gp, jacobian_precision=jp)
        w.step(1, H, 1.0)
        w.sync()
        one = w.state()
        w.close()
        w = DeviceWorld(scene, tol2=TOL * TOL, gravity_precision=gp, jacobian_precision=jp)
        w.sync()
        st0 = w.stats().astype(np.int64)
        t0 = time.perf_counter()
        w.step(steps, H, 1.0)
        w.sync()
        el = time.perf_counter() - t0
        st = w.state()
        tm = w.last_timing()
        stats = (w.stats().astype(np.int64) - st0).mean(axis=0) / steps
        w.close()
        r = {"body_steps_per_s": body_steps / el, "gravity": gp, "jacobian": jp,
             "island_ms_per_step": tm["island_ms"] / steps, "gravity_ms_per_step": tm["gravity_ms"] / steps,
             "newton_iterations_per_island_step": float(stats[1]), "spans_per_island_step": float(stats[0])}
        if ref is None:
            ref1, ref = one, st
        else:
            r["position_rel_error_1_step"] = _rel(one["position"], ref1["position"])
            r["momentum_rel_error_1_step"] = _rel(one["momentum"], ref1["momentum"])
            r["position_rel_error"] = _rel(st["position"], ref["position"])
            r["momentum_rel_error"] = _rel(st["momentum"], ref["momentum"])
            r["meets_1e-9"] = bool(r["position_rel_error"] <= 1e-9)
        runs[name] = r
    # the floor: FP64 again, gravity summed in another order
    from .arena import Arena, inverse_doubles
    ar = Arena(bodies=scene.n_bodies, islands=scene.n_islands, joints=scene.n_joints,
               inverses=inverse_doubles(scene), tol2=TOL * TOL, gravity_chunk=512)
    wid = ar.admit(scene)
    ar.step(1, H, 1.0)
    ar.sync()
    one = ar.state(wid)
    ar.step(steps - 1, H, 1.0)
    ar.sync()
    st = ar.state(wid)
    runs["f64_gravity_reordered"] = {
        "gravity": "f64 (chunks of 512 sources)", "jacobian": "f64",
        "position_rel_error_1_step": _rel(one["position"], ref1["position"]),
        "momentum_rel_error_1_step": _rel(one["momentum"], ref1["momentum"]),
        "position_rel_error": _rel(st["position"], ref["position"]),
        "momentum_rel_error": _rel(st["momentum"], ref["momentum"])}
    return {"bodies": scene.n_bodies, "islands": scene.n_islands, "joints": scene.n_joints, "steps": steps,
            "radius": radius, "runs": runs}


def main() -> int:
    ap = argparse.ArgumentParser(description="FP32 (compensated) against FP64: gravity forces and jointed rollouts.")
    ap.add_argument("--gravity-n", type=int, default=262144)
    ap.add_argument("--chains", type=int, default=8192)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--radius", type=float, default=None,
                    help="radius of the ball of chain origins (default: the 4,096-body workload's density)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--skip-gravity", action="store_true")
    args = ap.parse_args()
    line = {"form": "precision"}
    if not args.skip_gravity:
        line["gravity"] = gravity_form(args.gravity_n, args.seed)
    radius = args.radius if args.radius is not None else 50.0 * (args.chains / 512.0) ** (1.0 / 3.0)
    line["jointed"] = jointed_form(args.chains, args.steps, radius)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
