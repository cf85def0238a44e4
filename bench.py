"""Benchmark of the batched XPBD world step on one MI355X.

Workload (the metric's configuration): 65,536 worlds of the quadruped, each on
block terrain of per-world difficulty U[0, 1) drawn from the engine RNG seeded
1, world i reset with seed i; dt = 1/30 s, 20 substeps, grid 64 x 0.4 m;
actions i.i.d. uniform [-1, 1) from the engine RNG seeded 0 (world-major,
joint-minor), target = default + 0.6 a.  Targets for every timed step are
generated before the timed region and are resident in HBM.

A "step" is one control step of every world (20 XPBD substeps each);
value = worlds x steps / seconds over exactly --steps timed steps.

Prints ONE JSON line (rank 0) with the roofline of the dominant kernel
(k_step) and the CPU baseline: the reference engine built from its own
sources with its shipped flags (oracle/_ref), timed on this host's cores on a
bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "arc-demo_amd"))

METRIC = "world-steps/s @65536 worlds (served vs plain GPU vs host CPU); p50/p99 step lat"
FLOP_PER_WORLD_STEP = 4.12e5     # algorithmic FLOP per world-step (op count of the reference)
BYTES_PER_WORLD_STEP = 2.8e3     # algorithmic HBM bytes per world-step (state in/out, targets, outputs)
PEAK_FP32 = 157.3e12             # MI355X vector FP32, FLOP/s
PEAK_HBM = 8.0e12                # MI355X HBM3E, B/s
NJ = 12


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--worlds", type=int, default=65536)
    ap.add_argument("--block", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--cpu-worlds", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "worlds shard by HBM budget across GPUs with no exchange inside a step", file=sys.stderr)
        return 2

    from arc import native

    n = args.worlds
    steps, warm = args.steps, args.warmup
    sim = native.Sim(n)
    sim.set_launch(args.block, 0)
    # synthetic inputs: difficulty per world, actions per step
    diff_u = _rng_stream(1, n)
    diff = (diff_u >> 8).astype(np.float64) * (1.0 / 16777216.0)
    ids = np.arange(n, dtype=np.int32)
    sim.reset(ids, ids.astype(np.uint32), diff)
    u = _rng_stream(0, n * NJ * (steps + warm))
    a = (u >> 8).astype(np.float64) * (1.0 / 8388608.0) - 1.0
    targets = (np.tile(sim.default_angles, n * (steps + warm))[: a.size] + 0.6 * a).astype(np.float32)
    dev = native.DeviceBuffer.from_array(targets)
    base = dev.ptr
    per = n * NJ
    # warmup
    if warm:
        sim.time_steps(base, warm, per)
    sim.sync()
    t0 = time.perf_counter()
    kernel_ms, wall_ms = sim.time_steps(base + 4 * per * warm, steps, per)
    sim.sync()
    t1 = time.perf_counter()
    secs = t1 - t0
    value = n * steps / secs
    avg_kernel_s = kernel_ms / 1e3 / steps
    achieved = n * FLOP_PER_WORLD_STEP / avg_kernel_s
    flags = sim.observe()[3]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "world-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": 1e3 * secs / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{n} worlds, per-world difficulty U[0,1), uniform random actions, resets off",
                   "worlds": n, "dt": 1.0 / 30.0, "substeps": 20, "grid": 64, "cell": 0.4,
                   "form": "per-step launch, one world per lane", "block": args.block},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "traffic": None,
                     "hbm_frac": n * BYTES_PER_WORLD_STEP / avg_kernel_s / PEAK_HBM,
                     "kernel": "k_step", "kernel_ms_avg": 1e3 * avg_kernel_s},
        "unstable_worlds": int(flags.sum()),
        "bytes_per_world": sim.bytes_per_world(),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_worlds, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


def _rng_stream(seed: int, n: int) -> np.ndarray:
    """The engine's xorshift128 stream (libarcsim's own generator)."""
    from arc import native
    return native.rng_stream(seed, n)


def cpu_baseline(n: int, seconds: float) -> dict:
    """Reference engine (oracle/_ref, shipped flags) on this host's cores."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as P
    flags = open("/proc/cpuinfo").read()
    kind = "r64_v4" if all(f in flags for f in ("avx512f", "avx512bw", "avx512dq", "avx512vl", "avx512cd")) else "r64_v3"
    if not P.available(kind):
        kind = "r64_v3" if P.available("r64_v3") else None
    if kind is None:
        return {"value": None, "unit": "world-steps/s", "note": "oracle/_ref not built"}
    threads = min(16, P.cpu_threads())
    L = P.Lib(kind)
    s = L.sim(n)
    d = (_rng_stream(1, n) >> 8).astype(np.float64) * (1.0 / 16777216.0)
    ids = np.arange(n, dtype=np.int32)
    s.reset(ids, ids.astype(np.uint32), d)
    defaults = s.robot_info()[0]
    rs = _rng_stream(0, n * NJ * 64)
    done, t0 = 0, time.perf_counter()
    k = 0
    while True:
        blk = rs[(k % 64) * n * NJ:((k % 64) + 1) * n * NJ]
        t = defaults[None, :] + 0.6 * ((blk >> 8).astype(np.float64) * (1.0 / 8388608.0) - 1.0).reshape(n, NJ)
        s.step(t, threads)
        done += n
        k += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    cpu = open("/proc/cpuinfo").read().split("model name")[1].split("\n")[0].strip(": \t") if "model name" in flags else "?"
    return {"value": done / el, "unit": "world-steps/s", "cores": threads, "kind": "reference",
            "sample": f"{n} worlds x {k} steps ({el:.1f} s), same seeds/difficulty/action law",
            "build": kind, "cpu": cpu}


if __name__ == "__main__":
    sys.exit(main())
