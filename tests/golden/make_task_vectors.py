"""Writes tests/golden/task_vectors.npz from the reference's own checkpoint task.

The reference's Python task (arc/task.py, imported from the reference tree)
runs over its own ctypes binding (arc/native.py, imported from the reference
tree as well), whose library loader is pointed at the reference engine built
here from the reference's C sources by oracle/Makefile
(oracle/_ref/libref64s.so: FP64, no FMA contraction).  It is driven the way
its vectorised env drives it (arc/env.py:93-117): reset of every world, then
per step task.step(actions), observe, the ended worlds' reset and their
first observations.  The generator is np.random.default_rng(SEED), as
ArcTask(seed=SEED) makes it.

Stored per step: the raw actions, rewards, terminated, truncated, causes,
checkpoint counts, episode steps and difficulties, the seeds handed to the
engine's reset (in call order), the checkpoints (x, y, terrain z) and the
curriculum levels after the step and its resets, the generator's state
(PCG64 state, increment, 32-bit buffer), a SHA-1 of the float32 observations
of every step, and the observations themselves at a few steps.

Usage: python tests/golden/make_task_vectors.py   (needs the reference tree
and oracle/_ref; the committed .npz is what tests/test_task_golden.py reads)
"""

from __future__ import annotations

import ctypes
import hashlib
import importlib.util
import sys
import types
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
REF = Path("/root/reference/arc")
LIB = ROOT / "oracle" / "_ref" / "libref64s.so"

SEED, N, STEPS = 20260, 24, 40
OBS_AT = (0, 10, 20, 30, 40)
# a tightened task: checkpoints (wide radius and view), placements needing
# several rounds (|x|, |y| < 3.8 m), falls, truncations and level changes all
# happen within a few steps
CFG = dict(max_steps=17, checkpoint_radius=4.5, fov=170.0, border=9.0, beached_steps=2, flip_up_z=0.95,
           level_up_checkpoints=1, level_down_steps=10)


def load_reference():
    # the reference tree is read-only input: importing from it must leave no
    # __pycache__ behind
    sys.dont_write_bytecode = True
    pkg = types.ModuleType("refarc")
    pkg.__path__ = [str(REF)]
    sys.modules["refarc"] = pkg
    mods = {}
    for name in ("native", "task"):
        spec = importlib.util.spec_from_file_location(f"refarc.{name}", REF / f"{name}.py")
        mod = importlib.util.module_from_spec(spec)
        sys.modules[f"refarc.{name}"] = mod
        spec.loader.exec_module(mod)
        if name == "native":
            # the binding's loader, pointed at the reference engine built from its sources
            mod._load = lambda lib_name: ctypes.CDLL(str(LIB))
        mods[name] = mod
    return mods["native"], mods["task"]


def state_words(rng) -> np.ndarray:
    st = rng.bit_generator.state
    s, inc = st["state"]["state"], st["state"]["inc"]
    m = (1 << 64) - 1
    return np.array([s >> 64, s & m, inc >> 64, inc & m, st["has_uint32"], st["uinteger"]], dtype=np.uint64)


def obs_digest(obs: np.ndarray) -> np.ndarray:
    return np.frombuffer(hashlib.sha1(np.ascontiguousarray(obs, dtype=np.float32).tobytes()).digest(), np.uint8)


def main() -> None:
    if not LIB.exists():
        raise SystemExit(f"{LIB} missing: run `make -C oracle ref`")
    _, task = load_reference()
    cfg = task.TaskConfig(**CFG)
    t = task.ArcTask(N, cfg, threads=1, seed=SEED)
    resets = []
    engine_reset = t.sim.reset

    def record_reset(ids, seeds, difficulty):
        resets.append((np.array(ids, np.int64), np.array(seeds, np.uint32)))
        engine_reset(ids, seeds, difficulty)

    t.sim.reset = record_reset
    acts = np.random.default_rng(7).uniform(-1.3, 1.3, size=(STEPS, N, t.nj)).astype(np.float32)
    rec = {k: [] for k in ("reward", "terminated", "truncated", "cause", "checkpoints", "steps", "difficulty",
                           "target", "level", "rng", "obs_sha1", "reset_steps", "reset_ids", "reset_seeds")}
    obs_at = {}

    def snapshot(step, obs):
        rec["target"].append(np.concatenate([t.target, t.target_z[:, None]], 1))
        rec["level"].append(t.level.copy())
        rec["rng"].append(state_words(t.rng))
        rec["obs_sha1"].append(obs_digest(obs))
        if step in OBS_AT:
            obs_at[step] = obs.astype(np.float32)
        for ids, seeds in resets:
            rec["reset_steps"].append(np.full(len(ids), step, np.int64))
            rec["reset_ids"].append(ids)
            rec["reset_seeds"].append(seeds)
        resets.clear()

    t.reset(np.arange(N))
    snapshot(0, t.observe())
    for k in range(STEPS):
        reward, terminated, truncated, stats = t.step(acts[k])
        obs = t.observe()
        ended = np.flatnonzero(terminated | truncated)
        if len(ended):
            t.reset(ended)
            obs[ended] = t.observe()[ended]
        rec["reward"].append(reward)
        rec["terminated"].append(terminated)
        rec["truncated"].append(truncated)
        for key in ("cause", "checkpoints", "steps", "difficulty"):
            rec[key].append(np.asarray(stats[key]))
        snapshot(k + 1, obs)
    out = {"seed": np.array(SEED), "n": np.array(N), "actions": acts,
           "cfg_keys": np.array(list(CFG.keys())), "cfg_values": np.array(list(CFG.values()), np.float64)}
    for key, v in rec.items():
        out[key] = np.concatenate(v) if key.startswith("reset_") else np.stack(v)
    for step, o in obs_at.items():
        out[f"obs_{step}"] = o
    np.savez_compressed(HERE / "task_vectors.npz", **out)
    ev = {"resets": int(len(out["reset_ids"])), "falls": int(out["terminated"].sum()),
          "truncations": int(out["truncated"].sum()), "checkpoints": int(out["checkpoints"][-1].sum())}
    print("wrote", HERE / "task_vectors.npz", ev)


if __name__ == "__main__":
    main()
