"""Throughput of the per-step dynamics hot path on one MI355X.

Workload (the plain implicit Newton configuration, f64): 4,096 rigid bodies in 512
chains of 8 (each chain one island), adjacent bodies joined by an axial
spring-damper (k = 1e3, c = 10) and a torsional damper (c = 5), N-body gravity
(G = 1, size-core law) on, h = 1e-2, Newton tolerance 1e-10 on the worst per-block
residual ratio, Newton–Schulz 2 warm / 24 cold, a 200-step rollout.

A bench step is one world step: gravity over all bodies, then every island's
implicit Newton solve, re-anchoring and charge publication. The timed region runs
K steps of a freshly uploaded world (state resident in HBM before the clock
starts), bracketed by device synchronisation. One JSON line on stdout.

Beside the headline line: the same rollout with the FP32 Jacobian and FP32 gravity
(`mixed_precision`), and the explicit configuration the reference runs on its CPU —
1,024 point masses, symplectic Euler, dt = 1e-3, 100 steps (`plain_gravity_explicit`,
eager and graph-replayed, with the oracle on its first 10 steps as its CPU baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

CHAINS, M, H, TOL = 512, 8, 1e-2, 1e-10
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def ns_flops_per_iteration(scene) -> np.ndarray:
    """Algorithmic flops of one Newton–Schulz iteration per island: the sparse
    R = A·X (2·36·Σ_i |nbr(i)|·m per block column) plus the dense X·(2I − R)
    (2·(6m)³)."""
    out = []
    for i in range(scene.n_islands):
        s, e = scene.island_start[i], scene.island_start[i + 1]
        m = e - s
        js, je = scene.island_joint_start[i], scene.island_joint_start[i + 1]
        nbr = [{k} for k in range(m)]
        for q in range(js, je):
            a, b = scene.joint_a[q] - s, scene.joint_b[q] - s
            nbr[a].add(b)
            nbr[b].add(a)
        sparse = 2 * 36 * 6 * m * sum(len(x) for x in nbr)
        out.append(sparse + 2 * (6 * m) ** 3)
    return np.array(out, dtype=np.float64)


def matvec_flops(scene) -> np.ndarray:
    m = np.diff(scene.island_start).astype(np.float64)
    return 2 * (6 * m) ** 2


def cpu_baseline(target_s: float = 10.0):
    """The oracle (C++ f64 restatement of the reference's algorithm) on a bounded
    sample of the same workload: 64 chains of 8, warm steps until ~target_s."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pyoracle
    from exgame_amd import scenes
    threads = pyoracle.max_threads()
    scene = scenes.jointed_chains(n_chains=64, m=M, seed=0)
    w = pyoracle.OracleWorld(scene, tol2=TOL * TOL)
    w.step(1, H, 1.0)  # the cold first step (seeding) is not part of the sample
    n, t0 = 0, time.perf_counter()
    while True:
        w.step(1, H, 1.0)
        n += 1
        el = time.perf_counter() - t0
        if el >= target_s or n >= 200:
            break
    return {"value": scene.n_bodies * n / el, "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle, 64 chains x 8 bodies (512 bodies), {n} warm implicit steps after 1 cold, "
                      f"{el:.1f} s on {threads} OpenMP threads"}


def cpu_plain_gravity_explicit(n: int, steps: int, dt: float):
    """The oracle on the first `steps` steps of the 1,024-body explicit rollout (a
    second, bounded sample of the CPU baseline leg)."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pyoracle
    from exgame_amd import scenes
    scene = scenes.plain_gravity(n=n, seed=0)
    w = pyoracle.OracleWorld(scene, integrator="symplectic_euler")
    t0 = time.perf_counter()
    w.step(steps, dt, 1.0)
    el = time.perf_counter() - t0
    return {"value": n * steps / el, "unit": "body-steps/s", "cores": pyoracle.max_threads(), "kind": "port",
            "sample": f"oracle, the first {steps} steps of the {n}-body explicit rollout, {el:.2f} s"}


def plain_gravity_explicit(n: int = 1024, steps: int = 100, dt: float = 1e-3, repeats: int = 5):
    """The reference's CPU-runnable configuration on the device: N point masses,
    explicit symplectic Euler, size-core gravity, dt = 1e-3, 100 steps; stepped
    eagerly and by replaying a captured graph of two steps, timed after two untimed
    steps (which capture the graph). A rollout lasts a few milliseconds, so each form
    runs it `repeats` times on fresh worlds and reports the median."""
    from exgame_amd import scenes
    from exgame_amd.world import DeviceWorld
    scene = scenes.plain_gravity(n=n, seed=0)
    out = {"workload": f"plain_gravity_explicit_f64: {n} point masses, symplectic Euler, size-core gravity G=1, "
                       f"dt={dt:g}, {steps} steps", "repeats": repeats}
    for graph in (False, True):
        w = DeviceWorld(scene, integrator="symplectic_euler", graph=graph)
        w.step(2, dt, 1.0)   # code objects (and the graph's capture)
        w.sync()
        w.close()
        times = []
        for _ in range(repeats):
            w = DeviceWorld(scene, integrator="symplectic_euler", graph=graph)
            w.step(2, dt, 1.0)   # untimed: the graph's capture (and the same two steps eagerly)
            w.sync()
            t0 = time.perf_counter()
            w.step(steps, dt, 1.0)
            w.sync()
            times.append(time.perf_counter() - t0)
            w.close()
        el = float(np.median(times))
        out["graph" if graph else "eager"] = {"value": n * steps / el, "unit": "body-steps/s",
                                              "ms_per_step": el * 1e3 / steps}
    return out


def measured_traffic():
    """HBM bytes per island-kernel launch from the committed counter passes (copy-
    calibrated FETCH_SIZE and WRITE_SIZE), or None when no such record exists."""
    path = os.path.join(HERE, "profiles", "island_step_traffic_jointed_chains_4096.json")
    try:
        with open(path) as f:
            return int(json.load(f)["bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def copy_calibration(nbytes: int) -> None:
    """One device-to-device copy of `nbytes` through the HIP runtime (its blit kernel
    reads and writes exactly nbytes), so HBM byte counters can be scaled."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
        raise RuntimeError("copy calibration: hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
        raise RuntimeError("copy calibration: hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the run, copy this many MiB device-to-device once (the known-size copy that "
                         "scales the HBM byte counters of a profile)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (worlds shard by island with no "
              f"exchange, gravity by target rows with one all-gather; this build runs on one GPU)", file=sys.stderr)
        return 2

    from exgame_amd import scenes, _lib
    from exgame_amd.world import DeviceWorld

    scene = scenes.jointed_chains(n_chains=CHAINS, m=M, seed=0)
    info = _lib.device_info()
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib << 20)

    # warm-up on a throwaway world (code objects, caches, clocks)
    if args.warmup > 0:
        wu = DeviceWorld(scene, tol2=TOL * TOL)
        wu.step(args.warmup, H, 1.0)
        wu.sync()
        wu.close()

    world = DeviceWorld(scene, tol2=TOL * TOL)
    world.sync()
    stats0 = world.stats().astype(np.int64)
    t0 = time.perf_counter()
    world.step(args.steps, H, 1.0)
    world.sync()
    el = time.perf_counter() - t0
    timing = world.last_timing()
    stats = world.stats().astype(np.int64) - stats0

    body_steps = scene.n_bodies * args.steps
    value = body_steps / el
    # roofline of the dominant kernel (the island solve): Newton–Schulz + matvec flops
    ns_it = stats[:, 2].astype(np.float64)
    newton_it = stats[:, 1].astype(np.float64)
    flops = float(np.sum(ns_it * ns_flops_per_iteration(scene)) + np.sum(newton_it * matvec_flops(scene)))
    island_s = timing["island_ms"] * 1e-3
    per_launch_s = island_s / max(args.steps, 1)
    achieved = flops / max(args.steps, 1) / per_launch_s / 1e12 if per_launch_s > 0 else 0.0

    line = {
        "metric": "body-steps/s",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 scene)",
        "config": {
            "workload": "jointed_chains_implicit_f64: 4096 bodies = 512 chains x 8, axial k=1e3 c=10 + torsional "
                        "c=5, N-body gravity G=1, h=1e-2, Newton tol 1e-10, NS 2/24, 200-step rollout",
            "bodies": scene.n_bodies, "islands": scene.n_islands, "joints": scene.n_joints, "h": H,
            "newton_tol": TOL, "device": info["name"], "cus": info["compute_units"],
        },
        "kernels_ms_per_step": {"gravity": timing["gravity_ms"] / args.steps,
                                "island_solve": timing["island_ms"] / args.steps},
        "solver_per_step": {"newton_iterations_per_island": float(newton_it.mean() / args.steps),
                            "ns_iterations_per_island": float(ns_it.mean() / args.steps),
                            "spans_per_island": float(stats[:, 0].mean() / args.steps)},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": measured_traffic(),
                     "kernel": "island_step_kernel (Newton-Schulz + matvec flops / avg launch)"},
    }
    # the same rollout in mixed precision, against the f64 final state
    ref = world.state()["position"]
    mixed = {}
    for name, gp, jp in [("f32_jacobian", "f64", "f32c"), ("f32_jacobian_f32c_gravity", "f32c", "f32c")]:
        w = DeviceWorld(scene, tol2=TOL * TOL, gravity_precision=gp, jacobian_precision=jp)
        w.step(min(args.warmup, 1), H, 1.0)   # code objects of this variant
        w.close()
        w = DeviceWorld(scene, tol2=TOL * TOL, gravity_precision=gp, jacobian_precision=jp)
        w.sync()
        t1 = time.perf_counter()
        w.step(args.steps, H, 1.0)
        w.sync()
        e1 = time.perf_counter() - t1
        pos = w.state()["position"]
        err = float(np.linalg.norm(pos - ref) / np.linalg.norm(ref))
        mixed[name] = {"value": body_steps / e1, "speedup_vs_f64": el / e1, "final_state_rel_error": err,
                       "meets_1e-9": bool(err <= 1e-9), "gravity": gp, "jacobian": jp}
        w.close()
    line["mixed_precision"] = mixed
    line["plain_gravity_explicit"] = plain_gravity_explicit()
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # the baseline is reported beside the result, never instead of it
            line["cpu_baseline"] = {"value": None, "error": str(e)}
        try:
            line["plain_gravity_explicit"]["cpu_baseline"] = cpu_plain_gravity_explicit(1024, 10, 1e-3)
        except Exception as e:
            line["plain_gravity_explicit"]["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
