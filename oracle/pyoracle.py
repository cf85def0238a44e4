"""ctypes front end of the oracle (test infrastructure only).

Loads one of
  liboracle64.so / liboracle32.so      the CPU restatement (arc_oracle.c)
  _ref/libref64s.so / _ref/libref32.so the reference engine built from its own
                                       sources by oracle/Makefile (only where
                                       the reference tree exists)
  _ref/libref64_v3.so / _v4.so         the reference with its shipped flags
All export the same orc_* functions with a double API, so a scenario can be
replayed on any of them and compared.

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py use
this module; the product never imports it.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
STATE_SIZE = 13 * 13 + 12 * 4 + 21 * 8 + 2
NJ = 12

_d = ctypes.c_double
_i = ctypes.c_int
_u = ctypes.c_uint32
_p = ctypes.c_void_p


def lib_path(kind: str) -> Path:
    """kind: 'o64', 'o32', 'r64s', 'r32', 'r64_v3', 'r64_v4'."""
    if kind.startswith("o"):
        return HERE / f"liboracle{kind[1:]}.so"
    return HERE / "_ref" / f"libref{kind[1:]}.so"


def available(kind: str) -> bool:
    return lib_path(kind).exists()


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


class Lib:
    _cache: dict[str, "Lib"] = {}

    def __new__(cls, kind: str):
        if kind in cls._cache:
            return cls._cache[kind]
        self = super().__new__(cls)
        path = lib_path(kind)
        if not path.exists():
            raise OSError(f"{path} not built (run `make` / `make ref` in oracle/)")
        self.kind = kind
        self.lib = ctypes.CDLL(str(path))
        L = self.lib
        sig = {
            "orc_real_bytes": (_i,),
            "orc_rng_stream": (None, _u, _i, _p),
            "orc_hash32": (_u, _u),
            "orc_rng_uniform": (None, _u, _i, _p),
            "orc_world_new": (_p,),
            "orc_world_free": (None, _p),
            "orc_world_set_gravity": (None, _p, _d, _d, _d),
            "orc_world_set_mu": (None, _p, _d, _d),
            "orc_terrain_flat": (None, _p, _i, _i, _d, _d),
            "orc_terrain_generate": (None, _p, _i, _d, _d, _u),
            "orc_terrain_set": (None, _p, _i, _d),
            "orc_terrain_height": (_d, _p, _d, _d),
            "orc_terrain_info": (None, _p, _p),
            "orc_add_body": (_i, _p, _d, _p, _p, _p),
            "orc_add_sphere": (_i, _p, _i, _p, _d),
            "orc_add_hinge": (_i, _p, _i, _i, _p, _p),
            "orc_joint_set": (None, _p, _i, _p),
            "orc_joint_out": (None, _p, _i, _p),
            "orc_joint_angle": (_d, _p, _i),
            "orc_joint_rate": (_d, _p, _i),
            "orc_body_get": (None, _p, _i, _p),
            "orc_body_set": (None, _p, _i, _p),
            "orc_body_inertia": (None, _p, _i, _p),
            "orc_world_step": (None, _p, _d, _i, _i),
            "orc_world_set_pose": (None, _p, _p, _p, _p),
            "orc_world_energy": (_d, _p),
            "orc_world_unstable": (_i, _p),
            "orc_world_time": (_d, _p),
            "orc_world_counts": (_i, _p, _p),
            "orc_sphere_get": (None, _p, _i, _p),
            "orc_inertia": (None, _i, _d, _d, _d, _d, _p),
            "orc_sim_create": (_p, _i, _d, _i, _i, _d),
            "orc_sim_destroy": (None, _p),
            "orc_sim_num_worlds": (_i, _p),
            "orc_sim_robot_info": (None, _p, _p, _p, _p, _p),
            "orc_sim_world": (_p, _p, _i),
            "orc_sim_reset": (None, _p, _p, _i, _p, _p),
            "orc_sim_step": (None, _p, _p, _i),
            "orc_sim_observe": (None, _p, _p, _p, _p, _p, _p),
            "orc_sim_heights": (None, _p, _p, _i, _p),
            "orc_sim_terrain_at": (None, _p, _p, _p, _i, _p),
            "orc_state_size": (_i,),
            "orc_sim_state": (None, _p, _i, _p),
            "orc_sim_terrain": (None, _p, _i, _p),
            "orc_sim_num_contacts": (_i, _p, _i),
        }
        if kind.startswith("o"):
            sig["orc_sim_falls"] = (None, _p, _p, _p, _i, _d, _d)
        for name, (res, *args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = list(args)
        self.real_bytes = L.orc_real_bytes()
        cls._cache[kind] = self
        return self

    # -- RNG --------------------------------------------------------------
    def rng_stream(self, seed: int, n: int) -> np.ndarray:
        out = np.zeros(n, np.uint32)
        self.lib.orc_rng_stream(seed, n, out.ctypes.data)
        return out

    def rng_uniform(self, seed: int, n: int) -> np.ndarray:
        out = np.zeros(n)
        self.lib.orc_rng_uniform(seed, n, out.ctypes.data)
        return out

    def hash32(self, x: int) -> int:
        return self.lib.orc_hash32(x)

    def inertia(self, kind: int, mass: float, a: float, b: float = 0.0, c: float = 0.0) -> np.ndarray:
        out = np.zeros(3)
        self.lib.orc_inertia(kind, mass, a, b, c, out.ctypes.data)
        return out

    def world(self) -> "World":
        return World(self)

    def sim(self, n: int, dt: float = 1.0 / 30.0, substeps: int = 20, grid: int = 64,
            cell: float = 0.4) -> "Sim":
        return Sim(self, n, dt, substeps, grid, cell)


def _a(x, dtype=np.float64) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=dtype))


class World:
    """One hand-built world (the reference's arc_world API)."""

    def __init__(self, lib: Lib):
        self.L = lib.lib
        self.p = self.L.orc_world_new()

    def __del__(self):
        if getattr(self, "p", None):
            self.L.orc_world_free(self.p)
            self.p = None

    def gravity(self, x, y, z):
        self.L.orc_world_set_gravity(self.p, x, y, z)

    def mu(self, ms, md):
        self.L.orc_world_set_mu(self.p, ms, md)

    def terrain_flat(self, nx, ny, cell, height):
        self.L.orc_terrain_flat(self.p, nx, ny, cell, height)

    def terrain_generate(self, n, cell, difficulty, seed):
        self.L.orc_terrain_generate(self.p, n, cell, difficulty, seed)

    def terrain_set(self, i, h):
        self.L.orc_terrain_set(self.p, i, h)

    def height(self, x, y) -> float:
        return self.L.orc_terrain_height(self.p, x, y)

    def terrain_info(self) -> np.ndarray:
        out = np.zeros(6)
        self.L.orc_terrain_info(self.p, out.ctypes.data)
        return out

    def add_body(self, mass, inertia, pos, q=(1.0, 0.0, 0.0, 0.0)) -> int:
        I, p, qq = _a(inertia), _a(pos), _a(q)
        return self.L.orc_add_body(self.p, mass, I.ctypes.data, p.ctypes.data, qq.ctypes.data)

    def add_sphere(self, body, pos, rad) -> int:
        p = _a(pos)
        return self.L.orc_add_sphere(self.p, body, p.ctypes.data, rad)

    def add_hinge(self, parent, child, anchor, axis) -> int:
        a, x = _a(anchor), _a(axis)
        return self.L.orc_add_hinge(self.p, parent, child, a.ctypes.data, x.ctypes.data)

    def joint_set(self, j, lo, hi, kp=0.0, kd=0.0, max_torque=0.0, target=0.0):
        p = _a([lo, hi, kp, kd, max_torque, target])
        self.L.orc_joint_set(self.p, j, p.ctypes.data)

    def joint_out(self, j) -> np.ndarray:
        out = np.zeros(3)
        self.L.orc_joint_out(self.p, j, out.ctypes.data)
        return out

    def joint_angle(self, j) -> float:
        return self.L.orc_joint_angle(self.p, j)

    def joint_rate(self, j) -> float:
        return self.L.orc_joint_rate(self.p, j)

    def body(self, i) -> np.ndarray:
        out = np.zeros(13)
        self.L.orc_body_get(self.p, i, out.ctypes.data)
        return out

    def set_body(self, i, state13):
        s = _a(state13)
        self.L.orc_body_set(self.p, i, s.ctypes.data)

    def body_inertia(self, i) -> np.ndarray:
        out = np.zeros(3)
        self.L.orc_body_inertia(self.p, i, out.ctypes.data)
        return out

    def step(self, dt=1.0 / 30.0, substeps=20, steps=1):
        self.L.orc_world_step(self.p, dt, substeps, steps)

    def set_pose(self, pos, q, angles):
        p, qq, a = _a(pos), _a(q), _a(angles)
        self.L.orc_world_set_pose(self.p, p.ctypes.data, qq.ctypes.data, a.ctypes.data)

    def energy(self) -> float:
        return self.L.orc_world_energy(self.p)

    @property
    def unstable(self) -> int:
        return self.L.orc_world_unstable(self.p)

    @property
    def time(self) -> float:
        return self.L.orc_world_time(self.p)

    def counts(self) -> np.ndarray:
        out = np.zeros(4, np.int32)
        self.L.orc_world_counts(self.p, out.ctypes.data)
        return out

    def sphere(self, i) -> np.ndarray:
        out = np.zeros(8)
        self.L.orc_sphere_get(self.p, i, out.ctypes.data)
        return out


class Sim:
    """N worlds of the quadruped (the reference's arc_sim API)."""

    def __init__(self, lib: Lib, n, dt, substeps, grid, cell):
        self.lib = lib
        self.L = lib.lib
        self.n = n
        self.grid = min(grid, 64)
        self.p = self.L.orc_sim_create(n, dt, substeps, grid, cell)
        if not self.p:
            raise MemoryError("orc_sim_create failed")

    def __del__(self):
        if getattr(self, "p", None):
            self.L.orc_sim_destroy(self.p)
            self.p = None

    def robot_info(self):
        d, lo, hi, misc = np.zeros(NJ), np.zeros(NJ), np.zeros(NJ), np.zeros(3)
        self.L.orc_sim_robot_info(self.p, d.ctypes.data, lo.ctypes.data, hi.ctypes.data, misc.ctypes.data)
        return d, lo, hi, misc

    def reset(self, ids, seeds, difficulty):
        ids, seeds, diff = _a(ids, np.int32), _a(seeds, np.uint32), _a(difficulty)
        if len(ids):
            self.L.orc_sim_reset(self.p, ids.ctypes.data, len(ids), seeds.ctypes.data, diff.ctypes.data)

    def step(self, targets, threads: int = 1):
        t = _a(targets)
        assert t.size == self.n * NJ
        self.L.orc_sim_step(self.p, t.ctypes.data, threads)

    def observe(self):
        n = self.n
        root, rot, joints = np.zeros((n, 13)), np.zeros((n, 3, 3)), np.zeros((n, 36))
        contacts, flags = np.zeros((n, 6)), np.zeros(n, np.int32)
        self.L.orc_sim_observe(self.p, root.ctypes.data, rot.ctypes.data, joints.ctypes.data,
                               contacts.ctypes.data, flags.ctypes.data)
        return root, rot, joints, contacts, flags

    def heights(self, offsets) -> np.ndarray:
        o = _a(offsets).reshape(-1, 2)
        out = np.zeros((self.n, len(o)))
        self.L.orc_sim_heights(self.p, o.ctypes.data, len(o), out.ctypes.data)
        return out

    def terrain_at(self, ids, xy) -> np.ndarray:
        ids, xy = _a(ids, np.int32), _a(xy).reshape(-1, 2)
        out = np.zeros(len(ids))
        if len(ids):
            self.L.orc_sim_terrain_at(self.p, ids.ctypes.data, xy.ctypes.data, len(ids), out.ctypes.data)
        return out

    def state(self, i: int) -> np.ndarray:
        out = np.zeros(STATE_SIZE)
        self.L.orc_sim_state(self.p, i, out.ctypes.data)
        return out

    def states(self) -> np.ndarray:
        return np.stack([self.state(i) for i in range(self.n)])

    def terrain(self, i: int) -> np.ndarray:
        out = np.zeros(self.grid * self.grid + 1)
        self.L.orc_sim_terrain(self.p, i, out.ctypes.data)
        return out

    def num_contacts(self, i: int) -> int:
        return self.L.orc_sim_num_contacts(self.p, i)

    def falls(self, torso_steps: np.ndarray, beached_steps=30, flip_up_z=0.3, void_z=-2.0) -> np.ndarray:
        """Task fall rules (restatement only; see arc_oracle.c orc_sim_falls)."""
        assert torso_steps.dtype == np.int32 and torso_steps.flags.c_contiguous
        cause = np.zeros(self.n, np.int32)
        self.L.orc_sim_falls(self.p, torso_steps.ctypes.data, cause.ctypes.data, beached_steps,
                             flip_up_z, void_z)
        return cause


# --- the synthetic inputs shared by every harness ---------------------------

def actions(step_seed_stream: np.ndarray) -> np.ndarray:
    """u32 -> float-exact uniform [-1, 1): (u >> 8) * 2^-23 - 1."""
    return (step_seed_stream >> 8).astype(np.float64) * (1.0 / 8388608.0) - 1.0


def difficulties(seed: int, n: int, lib: Lib | None = None) -> np.ndarray:
    """Per-world difficulty U[0, 1): (u >> 8) * 2^-24 from the engine RNG."""
    u = (lib or Lib("o64")).rng_stream(seed, n)
    return (u >> 8).astype(np.float64) * (1.0 / 16777216.0)


def targets_from_actions(default: np.ndarray, a: np.ndarray, scale: float = 0.6) -> np.ndarray:
    """task.py:159-160 target map, rounded to float so both precisions see it."""
    return (default[None, :] + scale * a.reshape(-1, NJ)).astype(np.float32).astype(np.float64)


def cpu_threads() -> int:
    return len(os.sched_getaffinity(0))
