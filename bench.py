"""Benchmark of the batched XPBD world step on one MI355X.

Headline (value): the served form at 65,536 worlds -- 16 clients of 4,096
worlds each in one device-resident step server, each client keeping two step
requests in flight (the closed-loop figure, one in flight, is reported beside
it), every world on block terrain
of per-world difficulty U[0, 1) (engine RNG), world i reset with seed i,
dt = 1/30 s, 20 substeps, grid 64 x 0.4 m, actions i.i.d. uniform [-1, 1)
from the engine RNG (target = default + 0.6 a), fall rules and auto-reset on.
A "step" is one control step of all 65,536 worlds (one request of every
client); value = 65,536 x steps / seconds over exactly --steps timed steps
after --warmup untimed ones.

Beside it, in the same run:
  batched   the same worlds as one sim stepped by the batched launch
            (autotuned configuration) replayed from a hipGraph, resets on;
  plain     the plain form, one world per wavefront, one launch per step;
  latency   4,096 worlds at difficulty 0.5, seeds 0..4095, batched graph with
            resets: p50 / p99 of the per-step latency (host clock around
            launch + completion);
  serving   16 clients with world counts floor(exp(U ln 8192)), Poisson step
            arrivals at half the measured saturation rate: p50 / p99 request
            latency and throughput;
  roofline  the dominant kernel (k_step) timed with HIP events on the
            engine's stream over --steps launches at 65,536 worlds;
  cpu       the reference engine built from its sources with its shipped
            flags (oracle/_ref), on this host's cores, bounded sample.

Every pass draws fresh actions for every step it takes (target_pool): the
step kernel's cost depends on the action law.  Prints ONE JSON line.
`--only kernel` runs just the kernel timing pass (for counter collection
under rocprofv3); `--only forms` times every step form from one settled
state; `--only actions` times the kernel under laws that repeat a few action
sets against fresh actions.
"""

from __future__ import annotations

import argparse
import json
import math
import re
import sys
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "arc-demo_amd"))

METRIC = "world-steps/s @65536 worlds (served vs plain GPU vs host CPU); p50/p99 step lat"
FLOP_PER_WORLD_STEP = 4.12e5     # algorithmic FLOP per world-step (op count of the reference engine)
BYTES_PER_WORLD_STEP = 2.8e3     # algorithmic HBM bytes per world-step (state in/out, targets, outputs)
PEAK_FP32 = 157.3e12             # MI355X FP32 vector peak, FLOP/s
PEAK_HBM = 8.0e12                # MI355X HBM3E, B/s
NJ = 12
N = 65536
CLIENTS = 16
PROFILES = ROOT / "profiles"


def rng(seed: int, n: int) -> np.ndarray:
    from arc import native
    return native.rng_stream(seed, n)


def unit24(u: np.ndarray) -> np.ndarray:
    return (u >> 8).astype(np.float64) * (1.0 / 16777216.0)


def actions_to_targets(defaults: np.ndarray, u: np.ndarray) -> np.ndarray:
    a = (u >> 8).astype(np.float64) * (1.0 / 8388608.0) - 1.0
    return (np.tile(defaults, u.size // NJ) + 0.6 * a).astype(np.float32)


MAX_SETS = 512
TUNE_STEPS = 8   # timed steps per candidate in the autotuner (fewer let run-to-run noise pick a slower shape)
SETTLE = 100     # steps (fall rules and auto-reset on) every pass takes before --warmup and the timed ones:
                 # the contact load of freshly reset worlds settles after ~100 steps


def target_pool(defaults: np.ndarray, n: int, sets: int, seed: int = 0) -> np.ndarray:
    """`sets` independent draws of the action law for n worlds, float32
    (sets, n, 12); set k comes from the engine RNG seeded seed + k.  A pass
    steps with set k at its k-th step, so every timed step gets actions of its
    own: the step's cost depends on the action law (a law that cycles through
    a few sets makes the legs oscillate and the contact load grow; see
    --only actions)."""
    out = np.empty((sets, n, NJ), dtype=np.float32)
    d = np.asarray(defaults, dtype=np.float64)[None, :]
    for k in range(sets):
        # actions_to_targets, broadcast over the worlds instead of tiled
        a = (rng(seed + k, n * NJ) >> 8).reshape(n, NJ).astype(np.float64) * (1.0 / 8388608.0) - 1.0
        out[k] = d + 0.6 * a
    return out


def pool_sets(warm: int, steps: int) -> int:
    """sets of a pass's pool: one per step of its warm-up and timed steps (at most MAX_SETS, then cycled)"""
    return max(8, min(MAX_SETS, warm + steps))


def new_sim(n: int, difficulty=None):
    from arc import native
    s = native.Sim(n)
    ids = np.arange(n, dtype=np.int32)
    d = unit24(rng(1, n)) if difficulty is None else np.full(n, difficulty)
    s.reset(ids, ids.astype(np.uint32), d)
    return s


def settle(s, n, t, warm, sets):
    """SETTLE + --warmup steps with the fall rules and auto-reset on, target
    set k at step k; returns the steps taken"""
    per = n * NJ
    for k in range(SETTLE + warm):
        s.step_batch([0], [n], t.ptr + 4 * per * (k % sets), task=True, graph=True)
    s.sync()
    return SETTLE + warm


def warm_and_tune(s, n, t, warm, sets):
    """settle, then the autotuner on the settled worlds; t holds `sets`
    target sets.  Returns (launch, steps taken)."""
    k = settle(s, n, t, warm, sets)
    cfg, _ = s.autotune(TUNE_STEPS)
    return cfg, k


def apply_launch(s, launch: dict) -> None:
    """puts a sim in a launch configuration reported by launch()/autotune()"""
    s.set_form(launch["form"])
    if launch["form"] != "wave":
        s.set_launch(int(launch["block"]), 0 if launch["persistent"] else int(launch["grid_blocks"]))
        s.set_layout(int(launch["tile"]))
        s.set_persistent(bool(launch["persistent"]))


# the launch shape of a new sim (four lanes per world, 64-thread workgroups)
DEFAULT_LAUNCH = {"block": 64, "tile": 16, "grid_blocks": 0, "persistent": 0, "form": "leg"}


def kernel_pass(steps: int, warm: int, launch: dict | None = None) -> dict:
    """the step kernel at 65,536 worlds in the given launch configuration
    (the served headline's; a new sim's by default), from settled worlds: HIP
    events around each launch on the engine's stream"""
    from arc import native
    s = new_sim(N)
    S = pool_sets(SETTLE + warm, steps)
    t = native.DeviceBuffer.from_array(target_pool(s.default_angles, N, S))
    per = N * NJ
    k0 = settle(s, N, t, warm, S)
    apply_launch(s, launch or DEFAULT_LAUNCH)
    kms = 0.0
    for k in range(steps):
        m, _ = s.time_steps(t.ptr + 4 * per * ((k0 + k) % S), 1, per)
        kms += m
    avg_s = kms / 1e3 / steps
    return {"launch": s.launch(), "kernel_ms_avg": 1e3 * avg_s, "launches": steps,
            "settled_steps": k0}


def forms_pass(steps: int, warm: int) -> dict:
    """the step kernel of each form at 65,536 worlds (top-level keys) and at
    4,096 (under "worlds_4096"), each from a settled state reached in the leg
    form (all forms compute the same bits): HIP events around each launch"""
    from arc import native
    shapes = {"lane": 64, "leg": 64, "pair": 64, "shared": 256, "wave": 64}
    out = {}
    S = pool_sets(SETTLE + warm, steps)
    defaults = native.Sim(1).default_angles
    for n in (N, 4096):
        t = native.DeviceBuffer.from_array(target_pool(defaults, n, S))
        per = n * NJ
        res = {}
        for form, block in shapes.items():
            s = new_sim(n)
            s.set_form("leg")
            k0 = settle(s, n, t, warm, S)
            s.set_form(form)
            s.set_launch(block, 0)
            s.set_layout(16)
            # one untimed launch first: the form's first launch also loads its code
            s.time_steps(t.ptr + 4 * per * (k0 % S), 1, per)
            k0 += 1
            kms = 0.0
            for k in range(steps):
                m, _ = s.time_steps(t.ptr + 4 * per * ((k0 + k) % S), 1, per)
                kms += m
            res[form] = {"launch": s.launch(), "kernel_ms_avg": kms / steps}
            s.close()
        t.close()
        if n == N:
            out.update(res)
        else:
            out[f"worlds_{n}"] = res
    return out


def action_law_pass(steps: int, warm: int) -> dict:
    """the step kernel at 65,536 worlds under action laws that repeat with
    period p (p sets of the law in turn) and with fresh actions at every step:
    a new sim per law, SETTLE + --warmup steps of that law with resets, then
    --steps launches timed with HIP events"""
    from arc import native
    S = pool_sets(SETTLE + warm, steps)
    t = native.DeviceBuffer.from_array(target_pool(native.Sim(1).default_angles, N, S))
    per = N * NJ
    out = {}
    for p in (2, 4, 8, 32, S):
        s = new_sim(N)
        for k in range(SETTLE + warm):
            s.step_batch([0], [N], t.ptr + 4 * per * (k % p), task=True, graph=True)
        s.sync()
        kms = sum(s.time_steps(t.ptr + 4 * per * ((SETTLE + warm + k) % p), 1, per)[0] for k in range(steps))
        out["fresh" if p == S else f"period_{p}"] = {"period": p, "kernel_ms_avg": kms / steps,
                                                     "launch": s.launch()}
        s.close()
    t.close()
    return out


def batched_pass(steps: int, warm: int) -> dict:
    from arc import native
    s = new_sim(N)
    S = pool_sets(SETTLE + warm, steps)
    t = native.DeviceBuffer.from_array(target_pool(s.default_angles, N, S))
    per = N * NJ
    cfg, k0 = warm_and_tune(s, N, t, warm, S)
    t0 = time.perf_counter()
    for k in range(steps):
        s.step_batch([0], [N], t.ptr + 4 * per * ((k0 + k) % S), task=True, graph=True)
    s.sync()
    el = time.perf_counter() - t0
    stats = s.task_stats()
    return {"value": N * steps / el, "ms_per_step": 1e3 * el / steps, "launch": cfg,
            "resets": int(stats[1:].sum()), "bytes_per_world": s.bytes_per_world()}


def abi_pass(steps: int, warm: int) -> dict:
    """arc_sim_step through the reference ABI: host float64 targets in, the
    step, return on completion (PCIe-inclusive; never the headline)"""
    from arc import native
    s = new_sim(N)
    S = pool_sets(SETTLE + warm, 0)
    hp = target_pool(s.default_angles, N, S + steps + 1)
    t = native.DeviceBuffer.from_array(hp[:S])
    cfg, _ = warm_and_tune(s, N, t, warm, S)
    host = [hp[S + k].astype(np.float64) for k in range(steps + 1)]
    s.step(host[steps])
    s.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        s.step(host[k])
    s.sync()  # arc_sim_step returns with its step queued
    el = time.perf_counter() - t0
    return {"value": N * steps / el, "ms_per_step": 1e3 * el / steps, "launch": cfg,
            "path": "arc_sim_step(host float64 targets): convert, copy to HBM, step (the next call's conversion overlaps it)"}


def plain_pass(warm: int, steps: int = 3) -> dict:
    """the plain form from the settled state of the other passes (settled in
    the four-lanes-per-world form, which computes the same bits), one launch
    per control step, fall rules and auto-reset on"""
    from arc import native
    s = new_sim(N)
    S = pool_sets(SETTLE + warm, steps)
    t = native.DeviceBuffer.from_array(target_pool(s.default_angles, N, S))
    per = N * NJ
    k0 = settle(s, N, t, warm, S)
    s.set_form("wave")
    s.step_task(t.ptr + 4 * per * (k0 % S))  # the plain kernel's first launch outside the timed region
    s.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        s.step_task(t.ptr + 4 * per * ((k0 + 1 + k) % S))
    s.sync()
    el = time.perf_counter() - t0
    return {"value": N * steps / el, "ms_per_step": 1e3 * el / steps, "steps": steps, "settled_steps": k0 + 1,
            "form": "one world per wavefront, world staged in LDS, one launch per step"}


def config1_pass(steps: int = 1000) -> dict:
    """64 worlds on flat ground (difficulty 0), seeds 0..63, 1000 control
    steps with the fall rules and auto-reset on the device, one launch
    sequence per step from device-resident targets, timed end to end"""
    from arc import native
    n = 64
    s = new_sim(n, 0.0)
    t = native.DeviceBuffer.from_array(target_pool(s.default_angles, n, steps + 1, seed=300))
    per = n * NJ
    s.step_task(t.ptr + 4 * per * steps)  # first launches outside the timed region
    s.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        s.step_task(t.ptr + 4 * per * k)
    s.sync()
    el = time.perf_counter() - t0
    stats = s.task_stats()
    return {"worlds": n, "steps": steps, "value": n * steps / el, "ms_per_step": 1e3 * el / steps,
            "resets": int(stats[1:].sum()), "launch": s.launch()}


def latency_pass(steps: int, warm: int) -> dict:
    from arc import native
    n = 4096
    s = new_sim(n, 0.5)
    S = pool_sets(SETTLE + warm, steps)
    t = native.DeviceBuffer.from_array(target_pool(s.default_angles, n, S))
    per = n * NJ
    cfg, k0 = warm_and_tune(s, n, t, warm, S)
    lat = []
    for k in range(steps):
        t0 = time.perf_counter()
        s.step_batch([0], [n], t.ptr + 4 * per * ((k0 + k) % S), task=True, graph=True)
        s.sync()
        lat.append(1e3 * (time.perf_counter() - t0))
    lat = np.array(lat)
    return {"worlds": n, "difficulty": 0.5, "steps": steps, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
            "value": n / (lat.mean() / 1e3), "launch": cfg}


def served_pass(steps: int, warm: int) -> dict:
    """16 clients x 4,096 worlds; a step = one request of every client.
    Closed loop (each client waits for its step before sending the next) and
    pipelined (each client keeps two requests in flight), with the clients'
    targets resident in device memory (arcs_submit_device); the pipelined
    loop again with host targets (arcs_submit: staged and uploaded over PCIe
    inside the timed region, reported beside, never the headline)."""
    from arc import native, server
    per_world = native.Sim(1).bytes_per_world() + 2 * NJ * 4
    sv = server.StepServer(N * per_world * 1.0001)
    n = N // CLIENTS
    cids = [sv.open(n, seed0=c * n, difficulty=-1.0) for c in range(CLIENTS)]
    S = pool_sets(SETTLE + warm, steps)
    hp = target_pool(native.Sim(1).default_angles, N, S, seed=1000)
    dev = native.DeviceBuffer.from_array(hp)

    def run(k0, count, depth, device):
        """client c's step k: rows [c n, (c + 1) n) of set k mod S"""
        inflight = []
        for k in range(k0, k0 + count):
            j = k % S
            if device:
                inflight.append([sv.submit_device(c, dev.ptr + 4 * NJ * (j * N + c * n)) for c in cids])
            else:
                inflight.append([sv.submit(c, hp[j, c * n:(c + 1) * n]) for c in cids])
            if len(inflight) >= depth:
                for x in inflight.pop(0):
                    sv.wait(x)
        for ts in inflight:
            for x in ts:
                sv.wait(x)

    out = {"clients": CLIENTS, "worlds_per_client": n, "targets": "device-resident (arcs_submit_device)"}
    run(0, SETTLE + warm, 1, True)
    out["launch"], _ = sv.autotune(TUNE_STEPS)
    k = SETTLE + warm
    for depth, device, name in ((1, True, "closed_loop"), (2, True, "pipelined"), (2, False, "pipelined_host_pcie")):
        run(k, warm, depth, device)
        k += warm
        sv.reset_stats()
        t0 = time.perf_counter()
        run(k, steps, depth, device)
        el = time.perf_counter() - t0
        k += steps
        st = sv.stats()
        out[name] = {"value": N * steps / el, "ms_per_step": 1e3 * el / steps, "requests_in_flight_per_client": depth,
                     "targets": "device" if device else "host (PCIe-inclusive)",
                     "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "requests_per_batch": st["requests_per_batch"]}
    sv.close_server()
    dev.close()
    out["value"] = out["pipelined"]["value"]
    out["ms_per_step"] = out["pipelined"]["ms_per_step"]
    return out


def serving_pass(seconds: float) -> dict:
    """Clients log-uniform in [1, 8192] worlds: closed loop with one request in
    flight per client, saturation with two in flight per client, then Poisson
    arrivals at half the saturation rate."""
    from arc import native, server
    u = rng(4, CLIENTS).astype(np.float64) * (1.0 / 4294967296.0)
    sizes = [max(1, int(math.floor(math.exp(x * math.log(8192.0))))) for x in u]
    per_world = native.Sim(1).bytes_per_world() + 2 * NJ * 4
    sv = server.StepServer(sum(sizes) * per_world * 1.0001)
    cids = [sv.open(n, seed0=10000 * c, difficulty=-1.0) for c, n in enumerate(sizes)]
    defaults = native.Sim(1).default_angles
    pools = [target_pool(defaults, n, 64, seed=20000 + 100 * c) for c, n in enumerate(sizes)]
    used = [0] * CLIENTS

    def acts(c):
        """client c's next request: its next set of 64, in turn"""
        used[c] += 1
        return pools[c][used[c] % 64]

    for _ in range(30):  # settle the worlds, then tune the launch on them
        for x in [sv.submit(c, acts(c)) for c in cids]:
            sv.wait(x)
    launch, _ = sv.autotune(TUNE_STEPS)

    def loop(depth):
        """every client keeps `depth` requests in flight for a while"""
        stop = time.perf_counter() + min(1.0, seconds / 3)
        for c in cids:
            sv.wait(sv.submit(c, acts(c)))
        sv.reset_stats()

        def closed(c):
            q = [sv.submit(c, acts(c)) for _ in range(depth - 1)]
            while time.perf_counter() < stop:
                q.append(sv.submit(c, acts(c)))
                sv.wait(q.pop(0))
            for x in q:
                sv.wait(x)

        th = [threading.Thread(target=closed, args=(c,)) for c in cids]
        for x in th:
            x.start()
        for x in th:
            x.join()
        return sv.stats()

    one = loop(1)  # closed loop, one request in flight per client
    sat = loop(2)  # saturation: two in flight per client, a batch always queued
    rate = 0.5 * sat["requests"] / sat["seconds"]
    # open loop: merged Poisson arrivals, each to a uniformly chosen client
    g = np.random.default_rng(5)
    gaps = g.exponential(1.0 / rate, size=int(rate * seconds * 1.5) + 16)
    when = np.cumsum(gaps)
    when = when[when < seconds]
    who = g.integers(0, CLIENTS, size=when.size)
    sv.reset_stats()
    tickets = []
    t0 = time.perf_counter()
    for w, c in zip(when, who):
        d = t0 + w - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets.append(sv.submit(cids[c], acts(c)))
    for x in tickets:
        sv.wait(x)
    st = sv.stats()
    sv.close_server()
    return {"clients": CLIENTS, "worlds": int(sum(sizes)), "sizes": sizes, "launch": launch,
            "saturation_requests_per_s": sat["requests"] / sat["seconds"],
            "saturation_world_steps_per_s": sat["world_steps_per_s"],
            "saturation_requests_per_batch": sat["requests_per_batch"],
            "closed_loop_one_in_flight": {"requests_per_s": one["requests"] / one["seconds"],
                                          "world_steps_per_s": one["world_steps_per_s"], "p50_ms": one["p50_ms"],
                                          "p99_ms": one["p99_ms"], "requests_per_batch": one["requests_per_batch"]},
            "offered_requests_per_s": rate,
            "requests": int(st["requests"]), "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "max_ms": st["max_ms"],
            "world_steps_per_s": st["world_steps_per_s"], "requests_per_batch": st["requests_per_batch"]}


def kernel_name(launch: dict) -> str:
    return "k_step_leg" if launch.get("form") == "leg" else "k_step"


def rollout_pass(steps: int = 64, rollouts: int = 3, n: int = 8192) -> dict:
    """8,192 worlds at difficulty U[0,1), 64-step rollouts: a random-init
    2 x 256 tanh MLP in PyTorch reads the device task's observations and
    writes its actions, both exchanged zero-copy (DLPack out, device pointer
    in); rewards and dones gathered into a rollout buffer on the GPU.  Runs in
    its own process (torch's HIP runtime must load before libarcsim).  The
    policy and the task are ordered on the device (HIP events between torch's
    stream and the task's, no host wait inside a rollout); timed again with a
    host synchronisation at each hand-over, and with the draws from one numpy
    generator in the reference's order."""
    out = rollout_task(steps, rollouts, n, "world", "stream")
    host = rollout_task(steps, rollouts, n, "world", "host")
    npm = rollout_task(steps, rollouts, n, "numpy", "stream")
    keys = ("value", "ms_per_step", "mean_reward", "done_fraction")
    out["host_synchronised"] = {k: host[k] for k in keys}
    out["numpy_stream"] = {k: npm[k] for k in keys}
    return out


def rollout_task(steps: int, rollouts: int, n: int, mode: str, ordering: str) -> dict:
    import torch  # noqa: F401  (first: libarcsim then binds torch's HIP runtime)
    from arc import devtask
    cfg = devtask.TaskConfig(curriculum=False)
    t = devtask.DeviceTask(n, cfg, seed=0, rng=mode)
    t.set_levels(unit24(rng(1, n)))
    t.reset()
    cfgl, _ = t.sim.autotune(TUNE_STEPS)
    dev = torch.device("cuda")
    torch.manual_seed(0)
    policy = torch.nn.Sequential(torch.nn.Linear(t.obs_size, 256), torch.nn.Tanh(), torch.nn.Linear(256, 256),
                                 torch.nn.Tanh(), torch.nn.Linear(256, 12)).to(dev)
    obs = torch.from_dlpack(t.obs)            # the task's observation buffer, no copy
    rew = torch.from_dlpack(t.reward)
    done_a, done_b = torch.from_dlpack(t.terminated), torch.from_dlpack(t.truncated)
    buf_obs = torch.empty((steps, n, t.obs_size), device=dev)
    buf_rew = torch.empty((steps, n), device=dev)
    buf_done = torch.empty((steps, n), dtype=torch.bool, device=dev)
    act = torch.zeros((n, 12), device=dev)  # one actions buffer: the task's captured graph reads it in place
    cur = torch.cuda.current_stream()
    t.signal_stream(cur)                       # torch's work waits for the reset

    def rollout():
        with torch.no_grad():
            for k in range(steps):
                buf_obs[k].copy_(obs)
                act.copy_(policy(obs))
                if ordering == "stream":
                    t.wait_stream(cur)         # the task waits for the actions, on the device
                    t.step(act, graph=True)
                    t.signal_stream(cur)       # torch waits for the observations, on the device
                else:
                    cur.synchronize()          # actions complete before the task reads them
                    t.step(act, graph=True)
                    t.sync()                   # observations complete before torch reads them
                buf_rew[k].copy_(rew)
                buf_done[k].copy_((done_a | done_b).bool())
        cur.synchronize()

    rollout()  # warm (graph capture, kernels)
    t0 = time.perf_counter()
    for _ in range(rollouts):
        rollout()
    el = time.perf_counter() - t0
    t.close()
    return {"value": n * steps * rollouts / el, "unit": "world-steps/s", "worlds": n, "rollout_steps": steps,
            "rollouts": rollouts, "ms_per_step": 1e3 * el / (steps * rollouts), "launch": cfgl,
            "obs_size": t.obs_size, "mean_reward": float(buf_rew.mean()), "done_fraction": float(buf_done.float().mean()),
            "policy": "random-init MLP 195-256-256-12 tanh, fp32", "exchange": "DLPack (obs, rewards, dones), device pointer (actions)",
            "draws": mode, "ordering": "HIP events between torch's stream and the task's" if ordering == "stream"
            else "host synchronisation at each hand-over"}


def rollout_child() -> dict:
    import subprocess
    r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--only", "rollout"], capture_output=True,
                       text=True, timeout=600)
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode or not lines:
        return {"value": None, "error": (r.stderr or r.stdout)[-800:]}
    return json.loads(lines[-1])["rollout"]


def traffic_from_profile(kernel: str):
    """HBM bytes per launch of `kernel` at 65,536 worlds from the committed
    counter summary (FETCH_SIZE / WRITE_SIZE, calibrated), or None."""
    f = PROFILES / f"{kernel}_65536w_pmc.md"
    if not f.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def _ref_lib():
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as P
    flags = open("/proc/cpuinfo").read()
    kind = "r64_v4" if all(f in flags for f in ("avx512f", "avx512bw", "avx512dq", "avx512vl", "avx512cd")) else "r64_v3"
    if not P.available(kind):
        kind = "r64_v3" if P.available("r64_v3") else None
    return P, kind


def host_info() -> dict:
    """the CPUs this job sees: model, visible logical CPUs, physical cores,
    the affinity mask, OMP_NUM_THREADS, and the thread count the baselines
    use (the affinity mask, capped by OMP_NUM_THREADS when it is set: the
    share of the host's CPUs this job may use)"""
    import os
    text = open("/proc/cpuinfo").read()
    model = next((l.split(":", 1)[1].strip() for l in text.splitlines() if l.startswith("model name")), "?")
    cores = set()
    phys = core = None
    for l in text.splitlines():
        if l.startswith("physical id"):
            phys = l.split(":", 1)[1].strip()
        elif l.startswith("core id"):
            core = l.split(":", 1)[1].strip()
        elif not l.strip() and core is not None:
            cores.add((phys, core))
            phys = core = None
    if core is not None:
        cores.add((phys, core))
    affinity = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    use = min(affinity, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else affinity
    return {"cpu_model": model, "visible_cpus": os.cpu_count(), "physical_cores": len(cores) or None,
            "affinity_cpus": affinity, "omp_num_threads": omp, "threads_used": use}


def cpu_end_to_end(n: int, seconds: float, threads: int, max_steps: int | None = None) -> dict:
    """The reference's own step loop with its fall rules and auto-reset
    (arc/task.py:196-203 and arc/env.py:108-116 restated in numpy over the
    reference engine, oracle/_ref, shipped flags): step, observe, falls,
    reset of the fallen worlds; the host-side counterpart of the served path."""
    P, kind = _ref_lib()
    if kind is None:
        return {"value": None, "unit": "world-steps/s", "note": "oracle/_ref not built"}
    s = P.Lib(kind).sim(n)
    d = unit24(rng(1, n)) if max_steps is None else np.zeros(n)
    ids = np.arange(n, dtype=np.int32)
    s.reset(ids, ids.astype(np.uint32), d)
    defaults = s.robot_info()[0]
    half = 0.5 * 64 * 0.4
    torso = np.zeros(n, dtype=np.int64)
    episode = np.zeros(n, dtype=np.int64)
    done, k, resets, el = 0, 0, 0, 0.0
    while (el < seconds) if max_steps is None else (k < max_steps):
        # fresh actions every step (drawn outside the timed work)
        tg = actions_to_targets(defaults, rng(k, n * NJ)).astype(np.float64).reshape(n, NJ)
        t0 = time.perf_counter()
        s.step(tg, threads)
        root, rot, _, contacts, flags = s.observe()
        torso = np.where(contacts[:, 4] > 0, torso + 1, 0)
        bad = (flags != 0) | ~np.all(np.isfinite(root), axis=1)
        void = np.any(np.abs(root[:, :2]) > half, axis=1) | (root[:, 2] < -2.0)
        fell = bad | void | (rot[:, 2, 2] < 0.3) | (torso > 30)
        ended = np.flatnonzero(fell)
        if len(ended):
            episode[ended] += 1
            s.reset(ended.astype(np.int32), (ended + n * episode[ended]).astype(np.uint32), d[ended])
            torso[ended] = 0
            resets += len(ended)
        el += time.perf_counter() - t0
        done += n
        k += 1
    terrain = "flat (difficulty 0)" if max_steps is not None else "difficulty U[0,1)"
    return {"value": done / el, "unit": "world-steps/s", "cores": threads, "kind": "reference",
            "worlds": n, "steps": k,
            "sample": f"{n} worlds x {k} steps ({el:.1f} s), {terrain}: arc_sim_step, arc_sim_observe, numpy fall "
                      f"rules, arc_sim_reset of the fallen ({resets} resets); fresh uniform actions every step",
            "build": kind}


def cpu_baseline(n: int, seconds: float, threads: int) -> dict:
    """Reference engine (oracle/_ref, shipped flags) on `threads` of this
    host's cores: arc_sim_step only, timed as csrc/tools/bench.c:40-48 does."""
    P, kind = _ref_lib()
    if kind is None:
        return {"value": None, "unit": "world-steps/s", "note": "oracle/_ref not built"}
    s = P.Lib(kind).sim(n)
    d = unit24(rng(1, n))
    ids = np.arange(n, dtype=np.int32)
    s.reset(ids, ids.astype(np.uint32), d)
    defaults = s.robot_info()[0]
    done, k, el = 0, 0, 0.0
    while el < seconds:
        # fresh actions every step (drawn outside the timed work)
        tg = actions_to_targets(defaults, rng(k, n * NJ)).astype(np.float64).reshape(n, NJ)
        t0 = time.perf_counter()
        s.step(tg, threads)
        el += time.perf_counter() - t0
        done += n
        k += 1
    return {"value": done / el, "unit": "world-steps/s", "cores": threads, "kind": "reference", "worlds": n,
            "sample": f"{n} worlds x {k} steps ({el:.1f} s of stepping), per-world difficulty U[0,1), fresh "
                      "uniform actions every step (the GPU passes' law), arc_sim_step only", "build": kind}


def cpu_passes(seconds: float) -> dict:
    """the reference on all the CPU share this job may use and on half of it,
    at the metric's 65,536 worlds and at 64; config 1's end-to-end loop (64
    flat worlds, 1000 steps, fall rules and auto-reset) and the 65,536-world
    one on all of it"""
    host = host_info()
    full = host["threads_used"]
    half = max(1, full // 2)
    table = []
    for n, secs in ((N, seconds), (64, seconds / 2)):
        for th in (full, half):
            table.append(cpu_baseline(n, secs, th))
    return {"host": host, "table": table,
            "end_to_end_65536": cpu_end_to_end(N, seconds / 2, full),
            "end_to_end_config1_64": cpu_end_to_end(64, 0.0, full, max_steps=1000)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--latency-steps", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["all", "kernel", "rollout", "forms", "actions", "cpu"], default="all")
    ap.add_argument("--block", type=int, default=64, help="--only kernel: workgroup size of the leg form")
    ap.add_argument("--no-rollout", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested): worlds would shard by "
              "HBM budget across GPUs with no exchange inside a step, and no multi-GPU run is part of it",
              file=sys.stderr)
        return 2
    steps, warm = args.steps, args.warmup

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if args.only == "rollout":
        print(json.dumps({"rollout": rollout_pass()}), flush=True)
        return 0
    if args.only == "forms":
        print(json.dumps({"forms_pass": forms_pass(steps, warm)}), flush=True)
        return 0
    if args.only == "actions":
        print(json.dumps({"action_law_pass": action_law_pass(steps, warm)}), flush=True)
        return 0
    if args.only == "kernel":
        # a plain 1 GiB device copy first: the counters' reading for known traffic
        from arc import native
        a, b = native.DeviceBuffer(1 << 30), native.DeviceBuffer(1 << 30)
        native.sim_lib().arcx_memcpy(b.ptr, a.ptr, 1 << 30, 2)
        a.close()
        b.close()
        kp = kernel_pass(steps, warm, dict(DEFAULT_LAUNCH, block=args.block))
        print(json.dumps({"kernel_pass": kp}), flush=True)
        return 0
    if args.only == "cpu":
        print(json.dumps({"cpu": cpu_passes(args.cpu_seconds)}), flush=True)
        return 0

    log("served form, 65536 worlds")
    served = served_pass(steps, warm)
    log(f"served {served['value']:.4g} world-steps/s (closed loop {served['closed_loop']['value']:.4g}); "
        "batched graph form")
    batched = batched_pass(steps, warm)
    log(f"batched {batched['value']:.4g}; reference ABI path")
    abi = abi_pass(10, warm)
    log(f"abi {abi['value']:.4g}; kernel pass")
    kp = kernel_pass(steps, warm, served["launch"])
    log(f"{kernel_name(kp['launch'])} {kp['kernel_ms_avg']:.3f} ms; plain form")
    plain = plain_pass(warm)
    log(f"plain {plain['value']:.4g}; 64 flat worlds x 1000 steps")
    config1 = config1_pass()
    log(f"config 1 {config1['value']:.4g}; latency at 4096 worlds")
    lat = latency_pass(args.latency_steps, warm)
    log(f"latency p50 {lat['p50_ms']:.3f} p99 {lat['p99_ms']:.3f} ms; serving")
    serving = serving_pass(args.serve_seconds)
    log(f"serving p50 {serving['p50_ms']:.3f} p99 {serving['p99_ms']:.3f} ms")
    rollout = None
    if not args.no_rollout:
        rollout = rollout_child()
        log(f"rollout loop {rollout.get('value')}")

    avg_s = kp["kernel_ms_avg"] / 1e3
    achieved = N * FLOP_PER_WORLD_STEP / avg_s
    kname = kernel_name(kp["launch"])
    traffic = traffic_from_profile(kname)
    line = {
        "metric": METRIC,
        "value": served["value"],
        "unit": "world-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": served["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "65536 quadruped worlds served as 16 clients x 4096 (two step requests in flight "
                               "per client, targets resident in device memory), per-world difficulty U[0,1), "
                               "uniform random actions, fall rules + auto-reset on",
                   "worlds": N, "dt": 1.0 / 30.0, "substeps": 20, "grid": 64, "cell": 0.4,
                   "form": f"step server over the batched {served['launch'].get('form')}-form kernel "
                           f"(block {served['launch'].get('block')}, hipGraph)"},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "traffic": traffic,
                     "kernel": kname, "kernel_ms_avg": kp["kernel_ms_avg"], "worlds_per_launch": N,
                     "flop_per_world_step": FLOP_PER_WORLD_STEP,
                     "hbm_frac": N * BYTES_PER_WORLD_STEP / avg_s / PEAK_HBM},
        "forms": {"served": served, "batched_graph": batched, "plain": plain, "abi_pcie_inclusive": abi},
        "config1_64_flat_1000_steps": config1,
        "latency_4096": lat,
        "serving": serving,
        "rollout_8192": rollout,
        "launch": kp["launch"],
    }
    line["ratios"] = {"served_vs_plain": served["value"] / plain["value"],
                      "batched_vs_plain": batched["value"] / plain["value"]}
    if not args.no_cpu:
        log("cpu baselines")
        cpu = cpu_passes(args.cpu_seconds)
        head = cpu["table"][0]  # 65,536 worlds on the whole CPU share
        line["cpu_baseline"] = dict(head, host=cpu["host"])
        line["cpu"] = cpu
        if head.get("value"):
            line["ratios"]["served_vs_cpu"] = served["value"] / head["value"]
            e2e = cpu["end_to_end_config1_64"]
            if e2e.get("value"):
                line["ratios"]["config1_gpu_vs_cpu_end_to_end"] = config1["value"] / e2e["value"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
