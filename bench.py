"""Throughput of the per-step dynamics hot path on one MI355X.

Headline workload (the plain implicit Newton configuration, f64): 4,096 rigid bodies
in 512 chains of 8 (each chain one island), adjacent bodies joined by an axial
spring-damper (k = 1e3, c = 10) and a torsional damper (c = 5), N-body gravity
(G = 1, size-core law) on, h = 1e-2, Newton tolerance 1e-10 on the worst per-block
residual ratio, Newton–Schulz 2 warm / 24 cold.

A bench step is one world step: gravity over all bodies, then every island's
implicit Newton solve, re-anchoring and charge publication. The timed region runs
K steps of a freshly uploaded world (state resident in HBM before the clock
starts), bracketed by device synchronisation, replaying a captured graph of two
steps (every kernel of every step runs; the graph removes the host launches and the
gaps between dependent kernels). The same K steps run eagerly on another fresh
world give the per-kernel durations (events recorded by the dispatches themselves)
that the roofline uses, and the eager throughput beside it. One JSON line on stdout.

Beside the headline line:
  * mixed_precision — the same workload with FP32 Jacobians and FP32 gravity, timed
    as the headline over the same K steps, its final-state error taken over a fixed 200-step horizon
    against the FP64 device rollout (which the tests pin to the oracle), and a
    smaller world of the same shape (32 chains, 100 steps) against the oracle itself;
  * batched — 512 uneven worlds in one arena, mixed precision (body-steps/s and the
    ratio to the headline);
  * served — Poisson arrivals of rollout requests (256–32,768 bodies, 200–2,000
    steps), continuous batching, mixed precision: p50/p99 latency and body-steps/s,
    at λ = 2/s (below capacity) and at λ = 16/s (saturated);
  * plain_gravity_explicit — the reference's CPU-runnable configuration (1,024 point
    masses, symplectic Euler, dt = 1e-3, 100 steps);
  * cpu_baseline — the oracle (the C++ f64 restatement of the reference's algorithm)
    compiled for the host CPU, on all host cores and on one.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

CHAINS, M, H, TOL = 512, 8, 1e-2, 1e-10
ERROR_HORIZON = 200
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


def stage(msg: str) -> None:
    """One progress line on stderr as each part of the run finishes (stdout carries only
    the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return platform.processor() or "unknown"


_ORACLE = None


def oracle():
    """The oracle, compiled for this host's CPU on first use (the same arithmetic as
    the test build: -ffp-contract=off keeps every operation rounded as written)."""
    global _ORACLE
    if _ORACLE is None:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import pyoracle
        pyoracle.use_native()
        _ORACLE = pyoracle
    return _ORACLE


def cpu_baseline(steps_all: int = 2, target_single_s: float = 8.0):
    """The oracle built for the host CPU (-O3 -march=native, OpenMP over islands and
    gravity rows) on a bounded sample of the workload: the full 4,096 bodies on all
    host threads (one cold step untimed, then `steps_all` timed warm steps), and 64
    chains (512 bodies) on one thread (warm steps for about `target_single_s`)."""
    pyoracle = oracle()
    from exgame_amd import scenes
    threads = pyoracle.max_threads()
    scene = scenes.jointed_chains(n_chains=CHAINS, m=M, seed=0)
    w = pyoracle.OracleWorld(scene, tol2=TOL * TOL)
    w.step(1, H, 1.0)
    t0 = time.perf_counter()
    w.step(steps_all, H, 1.0)
    el = time.perf_counter() - t0
    out = {"value": scene.n_bodies * steps_all / el, "unit": "body-steps/s", "cores": threads, "kind": "port",
           "sample": f"oracle (-O3 -march=native, {threads} OpenMP threads), the 4096-body workload: "
                     f"{steps_all} warm implicit steps after 1 cold, {el:.1f} s",
           "cpu": cpu_model(), "host_threads": os.cpu_count()}
    pyoracle.set_threads(1)
    small = scenes.jointed_chains(n_chains=64, m=M, seed=0)
    w1 = pyoracle.OracleWorld(small, tol2=TOL * TOL)
    w1.step(1, H, 1.0)
    n, t0 = 0, time.perf_counter()
    while True:
        w1.step(1, H, 1.0)
        n += 1
        el1 = time.perf_counter() - t0
        if el1 >= target_single_s or n >= 100:
            break
    pyoracle.set_threads(threads)
    out["single_thread"] = {"value": small.n_bodies * n / el1, "unit": "body-steps/s", "cores": 1,
                            "sample": f"64 chains x 8 (512 bodies), {n} warm steps, {el1:.1f} s"}
    return out


def oracle_mixed_check(chains: int = 32, steps: int = 100):
    """Final-state error of the mixed-precision device path against the oracle itself,
    on a world of the headline's shape small enough for the oracle."""
    pyoracle = oracle()
    from exgame_amd import scenes
    from exgame_amd.world import DeviceWorld
    scene = scenes.jointed_chains(n_chains=chains, m=M, seed=0)
    orc = pyoracle.OracleWorld(scene, tol2=TOL * TOL)
    orc.step(steps, H, 1.0)
    ref = orc.state()
    out = {"world": f"{chains} chains x 8 ({scene.n_bodies} bodies), {steps} steps"}
    for name, gp, jp in [("f64", "f64", "f64"), ("f32_jacobian_f32c_gravity", "f32c", "f32c")]:
        w = DeviceWorld(scene, tol2=TOL * TOL, gravity_precision=gp, jacobian_precision=jp)
        w.step(steps, H, 1.0)
        st = w.state()
        w.close()
        out[name] = {"position_rel_error": float(np.linalg.norm(st["position"] - ref["position"]) /
                                                 np.linalg.norm(ref["position"])),
                     "momentum_rel_error": float(np.linalg.norm(st["momentum"] - ref["momentum"]) /
                                                 np.linalg.norm(ref["momentum"]))}
    return out


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


def batched_form(steps: int = 20, warmup: int = 2, worlds: int = 512):
    """512 uneven worlds (lognormal sizes, median 512, 60 % free bodies, 40 % chains of
    8) in one arena, mixed precision, stepped together."""
    from exgame_amd.arena import Arena, inverse_doubles
    from exgame_amd.batched import batch_worlds
    ws = batch_worlds(worlds, 0)
    B = sum(w.n_bodies for w in ws)
    ar = Arena(bodies=B, islands=sum(w.n_islands for w in ws), joints=sum(w.n_joints for w in ws) + 1,
               inverses=sum(inverse_doubles(w) for w in ws), tol2=TOL * TOL, gravity_precision="f32c",
               jacobian_precision="f32c", gravity_chunk=1024, graph=True)
    ids = [ar.admit(w) for w in ws]
    assert all(i is not None for i in ids)
    ar.step(1, H, 1.0)
    ar.sync()
    mapping = ar.autotune_joints()    # the mapping of later admissions: admit the worlds again
    for i in ids:
        ar.retire(i)
    ids = [ar.admit(w) for w in ws]
    assert all(i is not None for i in ids)
    ar.step(warmup, H, 1.0)
    ar.sync()
    t0 = time.perf_counter()
    ar.step(steps, H, 1.0)
    ar.sync()
    el = time.perf_counter() - t0
    ar.close()
    return {"workload": f"{len(ws)} worlds, {B} bodies (lognormal sizes 64-8192, median 512; 60% free bodies, "
                        f"40% chains of 8), mixed precision (f32 gravity and Jacobians), h=1e-2, tol 1e-10, "
                        f"two-step graph replay, island classes side by side", "bodies": B, "steps": steps,
            "joint_mapping": mapping, "value": B * steps / el, "unit": "body-steps/s",
            "ms_per_step": el * 1e3 / steps}


def served_form(requests: int = 16, rate: float = 2.0, seed: int = 0):
    """Poisson arrivals of rollout requests served with continuous batching."""
    from exgame_amd.serve import Server, arena_for, synthetic_requests
    reqs = synthetic_requests(requests, rate, seed)
    budget = 16 << 30
    ar = arena_for(reqs, budget, tol2=TOL * TOL, gravity_precision="f32c", jacobian_precision="f32c")
    rep = Server(ar, budget).run(reqs)
    ar.close()
    rep.update({"rate_per_s": rate, "bodies_range": [256, 32768], "steps_range": [200, 2000],
                "precision": "f32 gravity and Jacobians, f64 residual / Newton-Schulz / state"})
    return rep


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


def run_world(scene, steps, gp="f64", jp="f64", graph=False):
    """`steps` steps of a fresh world, timed. With `graph` the world first takes two
    untimed steps, which capture (and run) its two-step graph, and the timed steps
    replay it."""
    from exgame_amd.world import DeviceWorld
    w = DeviceWorld(scene, tol2=TOL * TOL, gravity_precision=gp, jacobian_precision=jp, graph=graph)
    if graph:
        w.step(2, H, 1.0)
    w.sync()
    stats0 = w.stats().astype(np.int64)
    t0 = time.perf_counter()
    w.step(steps, H, 1.0)
    w.sync()
    el = time.perf_counter() - t0
    res = {"elapsed": el, "timing": w.last_timing(), "stats": w.stats().astype(np.int64) - stats0, "state": w.state()}
    w.close()
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="skip the forms beside the headline (profiling)")
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

    # warm-up on a throwaway world (code objects, caches, clocks), both precisions
    if args.warmup > 0:
        for gp in ("f64", "f32c"):
            wu = DeviceWorld(scene, tol2=TOL * TOL, gravity_precision=gp, jacobian_precision=gp)
            wu.step(args.warmup, H, 1.0)
            wu.sync()
            wu.close()

    r = run_world(scene, args.steps)                  # eager: per-kernel durations, counters, state
    el, timing, stats = r["elapsed"], r["timing"], r["stats"]
    body_steps = scene.n_bodies * args.steps
    rg = run_world(scene, args.steps, graph=True)     # the timed region of the headline
    value = body_steps / rg["elapsed"]
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
        "ms_per_step": rg["elapsed"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 scene)",
        "config": {
            "workload": f"jointed_chains_implicit_f64: 4096 bodies = 512 chains x 8, axial k=1e3 c=10 + torsional "
                        f"c=5, N-body gravity G=1, h=1e-2, Newton tol 1e-10, NS 2/24, {args.steps} timed steps",
            "launch": "two-step graph replay", "bodies": scene.n_bodies, "islands": scene.n_islands,
            "joints": scene.n_joints, "h": H,
            "newton_tol": TOL, "device": info["name"], "cus": info["compute_units"],
        },
        "eager": {"value": body_steps / el, "ms_per_step": el * 1e3 / args.steps,
                  "note": "the same steps launched one kernel at a time (the run the kernel durations come from)"},
        "kernels_ms_per_step": {"gravity": timing["gravity_ms"] / args.steps,
                                "island_solve": timing["island_ms"] / args.steps},
        "solver_per_step": {"newton_iterations_per_island": float(newton_it.mean() / args.steps),
                            "ns_iterations_per_island": float(ns_it.mean() / args.steps),
                            "spans_per_island": float(stats[:, 0].mean() / args.steps)},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": measured_traffic(),
                     "kernel": "island_step_kernel (Newton-Schulz + matvec flops / avg launch)"},
    }
    # mixed precision: timed as the headline (graph replay, the same steps); error and
    # kernel durations from an eager rollout over a fixed horizon
    ref_h = r["state"] if args.steps == ERROR_HORIZON else run_world(scene, ERROR_HORIZON)["state"]
    mixed = {"error_horizon_steps": ERROR_HORIZON}
    for name, gp, jp in [("f32_jacobian", "f64", "f32c"), ("f32_jacobian_f32c_gravity", "f32c", "f32c")]:
        rm = run_world(scene, args.steps, gp, jp, graph=True)
        rh = run_world(scene, ERROR_HORIZON, gp, jp)
        st_h = rh["state"]
        err = float(np.linalg.norm(st_h["position"] - ref_h["position"]) / np.linalg.norm(ref_h["position"]))
        merr = float(np.linalg.norm(st_h["momentum"] - ref_h["momentum"]) / np.linalg.norm(ref_h["momentum"]))
        mixed[name] = {"value": body_steps / rm["elapsed"], "speedup_vs_f64": rg["elapsed"] / rm["elapsed"],
                       "final_state_rel_error_vs_f64_device": err, "momentum_rel_error_vs_f64_device": merr,
                       "meets_1e-9": bool(err <= 1e-9), "gravity": gp, "jacobian": jp,
                       "kernels_ms_per_step": {"gravity": rh["timing"]["gravity_ms"] / ERROR_HORIZON,
                                               "island_solve": rh["timing"]["island_ms"] / ERROR_HORIZON}}
    line["mixed_precision"] = mixed
    stage(f"headline {value:.4g} body-steps/s, mixed precision done")
    if not args.headline_only:
        try:
            line["mixed_precision"]["vs_oracle"] = oracle_mixed_check()
        except Exception as e:  # reported beside the result, never instead of it
            line["mixed_precision"]["vs_oracle"] = {"error": str(e)}
        stage("oracle mixed check done")
        b = batched_form()
        stage(f"batched {b['value']:.4g} body-steps/s")
        b["ratio_vs_headline_f64"] = b["value"] / value
        line["batched"] = b
        s = served_form()
        s["ratio_vs_headline_f64"] = s["body_steps_per_s"] / value
        s["note"] = "offered load below capacity: body-steps/s is the arrival rate's"
        line["served"] = s
        stage("served at 2/s done")
        s = served_form(requests=24, rate=16.0)
        s["ratio_vs_headline_f64"] = s["body_steps_per_s"] / value
        s["note"] = "arrivals outpace service: body-steps/s is the device's on this mix"
        line["served_saturated"] = s
        stage("served at 16/s done")
        line["plain_gravity_explicit"] = plain_gravity_explicit()
        stage("plain gravity explicit done")
        if not args.no_cpu_baseline:
            try:
                line["cpu_baseline"] = cpu_baseline()
            except Exception as e:  # the baseline is reported beside the result, never instead of it
                line["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
