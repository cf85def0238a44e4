"""Writes the golden vectors of tests/golden/ from the reference engine.

Runs the reference engine built from its own sources by oracle/Makefile
(oracle/_ref/libref32.so: FP32 build; libref64s.so: FP64, no FMA contraction)
on small, fixed scenarios and stores inputs and outputs as .npz files:

  rng.npz        xorshift128 streams (u32) and uniforms for a few seeds
  sim_<p>.npz    p in {32, 64}: 16 worlds at difficulties 0 / 0.5 / 1 and a
                 per-world mixed case: terrain after reset, full world state
                 after reset and after control steps 1, 2 and 10, the
                 observe outputs after step 10, heights at a probe grid
  kat_<p>.npz    single-world known-answer scenarios (free fall, pendulum,
                 resting sphere, block terrain) at fixed checkpoints

Usage: python tests/golden/make_fixtures.py   (needs oracle/_ref, i.e. the
reference tree; the committed .npz files are what the tests read)
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))

import pyoracle as P  # noqa: E402

N = 16
STEPS = (1, 2, 10)
CASES = {"d0": 0.0, "d05": 0.5, "d1": 1.0, "mixed": None}
OFFSETS = np.array([(f, l) for f in np.arange(-1.0, 1.41, 0.4) for l in np.arange(-1.0, 1.01, 0.4)],
                   dtype=np.float32).astype(np.float64)


def targets(n: int, steps: int) -> np.ndarray:
    u = P.Lib("o64").rng_stream(0, n * 12 * steps)
    d = P.Lib("o64").sim(1).robot_info()[0]
    return P.targets_from_actions(d, P.actions(u)).reshape(steps, n, 12)


def sim_fixture(kind: str) -> dict:
    L = P.Lib(kind)
    out = {}
    tg = targets(N, max(STEPS))
    out["targets"] = tg
    for name, d in CASES.items():
        diff = P.difficulties(1, N) if d is None else np.full(N, d)
        s = L.sim(N)
        ids = np.arange(N, dtype=np.int32)
        s.reset(ids, ids.astype(np.uint32), diff)
        out[f"{name}/difficulty"] = diff
        out[f"{name}/terrain"] = np.stack([s.terrain(i) for i in range(N)]).astype(np.float32 if kind == "r32" else np.float64)
        out[f"{name}/state0"] = s.states()
        for k in range(1, max(STEPS) + 1):
            s.step(tg[k - 1])
            if k in STEPS:
                out[f"{name}/state{k}"] = s.states()
        root, rot, joints, contacts, flags = s.observe()
        out[f"{name}/root"], out[f"{name}/rot"], out[f"{name}/joints"] = root, rot, joints
        out[f"{name}/contacts"], out[f"{name}/flags"] = contacts, flags
        out[f"{name}/heights"] = s.heights(OFFSETS)
    out["offsets"] = OFFSETS
    out["robot_info"] = np.concatenate(L.sim(1).robot_info())
    out["template"] = L.sim(1).state(-1)
    return out


def kat_fixture(kind: str) -> dict:
    L = P.Lib(kind)
    out = {}
    # free fall (csrc/tests/test_phys.c:23-35)
    w = L.world()
    w.terrain_flat(1, 1, 1.0, -1e6)
    w.add_body(2.0, L.inertia(0, 2.0, 0.1), (0, 0, 0))
    st = w.body(0)
    st[7] = 1.0
    w.set_body(0, st)
    w.step(steps=30)
    out["free_fall"] = w.body(0)
    # pendulum released from 0.1 rad, 600 fine steps (test_phys.c:61-101)
    w = L.world()
    w.terrain_flat(1, 1, 1.0, -1e6)
    w.add_body(1.0, L.inertia(0, 1.0, 0.05), (0, 0, -1.0))
    w.add_hinge(-1, 0, (0, 0, 0), (0, 1, 0))
    w.set_pose((0, 0, 0), (1, 0, 0, 0), [0.1])
    ang = []
    for _ in range(600):
        w.step(1.0 / 600.0, 1, 1)
        ang.append(w.joint_angle(0))
    out["pendulum_angle"] = np.array(ang)
    out["pendulum_body"] = w.body(0)
    # resting sphere (test_phys.c:188-206)
    w = L.world()
    w.terrain_flat(8, 8, 0.5, 0.0)
    w.add_body(1.0, L.inertia(0, 1.0, 0.1), (0.01, 0.02, 0.5))
    w.add_sphere(0, (0, 0, 0), 0.1)
    w.step(steps=60)
    out["resting_body"] = w.body(0)
    out["resting_sphere"] = w.sphere(0)
    # block terrain plateau (test_phys.c:251-283)
    w = L.world()
    w.terrain_flat(8, 8, 0.5, 0.0)
    for i in range(64):
        if i % 8 >= 4:
            w.terrain_set(i, 0.3)
    for pos in ((0.5, 0.0, 0.6), (-0.12, 0.25, 0.6), (0.02, -0.75, 0.6)):
        b = w.add_body(1.0, L.inertia(0, 1.0, 0.1), pos)
        w.add_sphere(b, (0, 0, 0), 0.1)
    w.step(steps=90)
    out["blocks_bodies"] = np.stack([w.body(i) for i in range(3)])
    # generated terrain used by the void test (csrc/tests/test_robot.c:168-188)
    w = L.world()
    w.terrain_generate(32, 0.4, 0.5, 9)
    out["void_terrain_info"] = w.terrain_info()
    return out


def main() -> None:
    for kind in ("r32", "r64s"):
        if not P.available(kind):
            raise SystemExit(f"{P.lib_path(kind)} missing: run `make ref` in oracle/ first")
    L = P.Lib("r64s")
    rng = {f"stream{s}": L.rng_stream(s, 4096) for s in (0, 1, 7, 12345, 0xFFFFFFFF)}
    rng.update({f"hash{x}": np.uint32(L.hash32(x)) for x in (0, 1, 2, 0x9E3779B9, 0xDEADBEEF)})
    rng["uniform32"] = P.Lib("r32").rng_uniform(3, 1024)
    rng["uniform64"] = L.rng_uniform(3, 1024)
    np.savez_compressed(HERE / "rng.npz", **rng)
    for kind, tag in (("r32", "32"), ("r64s", "64")):
        np.savez_compressed(HERE / f"sim_{tag}.npz", **sim_fixture(kind))
        np.savez_compressed(HERE / f"kat_{tag}.npz", **kat_fixture(kind))
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
