"""Shared test helpers: the oracle, the host-compiled engine harness and the
synthetic inputs every harness uses (seeds, actions, difficulties)."""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

import pyoracle as P

ROOT = Path(__file__).resolve().parent.parent
HOSTCHECK = ROOT / "tests" / "hostcheck" / "libhostcheck.so"
STATE = P.STATE_SIZE
NJ = 12


def default_angles_f64() -> np.ndarray:
    return P.Lib("o64").sim(1).robot_info()[0]


def action_stream(n_worlds: int, steps: int, seed: int = 0) -> np.ndarray:
    """Targets for `steps` control steps: uniform [-1, 1) actions from the
    engine RNG seeded `seed`, world-major, joint-minor; target = default +
    0.6 a rounded to float32 (task.py:159-160)."""
    u = P.Lib("o64").rng_stream(seed, n_worlds * NJ * steps)
    d = default_angles_f64()
    t = P.targets_from_actions(d, P.actions(u))
    return t.reshape(steps, n_worlds, NJ)


def same(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Elementwise equality treating NaN == NaN (and -0 == +0)."""
    return (a == b) | (np.isnan(a) & np.isnan(b))


# smallest normal binary32: differences below it are denormal noise around an
# exact zero, where a relative bound means nothing
FLOOR32 = float(np.finfo(np.float32).tiny)


def rel_close(a, b, rel: float = 1e-5, floor: float = FLOOR32) -> np.ndarray:
    """|a - b| <= rel |b| elementwise (pure relative), with an absolute floor
    at the smallest normal binary32 for values at or next to zero; NaN equals
    NaN, and an infinity only itself."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    with np.errstate(invalid="ignore"):
        d = np.abs(a - b)
        return same(a, b) | (d <= rel * np.abs(b)) | (d <= floor)


def max_rel(a, b) -> float:
    """Largest |a - b| / |b| over the finite pairs that differ (0 if none)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    k = ~same(a, b) & np.isfinite(a) & np.isfinite(b)
    if not k.any():
        return 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.max(np.abs(a[k] - b[k]) / np.maximum(np.abs(b[k]), FLOOR32)))


def _first_diff(ok: np.ndarray, a, b) -> str:
    bad = np.argwhere(~ok)
    i = tuple(bad[0])
    return f"{len(bad)} of {ok.size} values differ; first at {i}: {a[i]!r} vs {b[i]!r}; max rel {max_rel(a, b):.3g}"


def assert_same(a, b, what: str = "") -> None:
    """Bit-exact equality (NaN == NaN, -0 == +0), with the first difference
    and the largest relative error in the message."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shapes {a.shape} vs {b.shape}"
    ok = same(a, b)
    assert ok.all(), f"{what}: {_first_diff(ok, a, b)}"


def assert_rel(a, b, rel: float = 1e-5, what: str = "") -> None:
    """rel_close everywhere, with the first violation in the message."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert a.shape == b.shape, f"{what}: shapes {a.shape} vs {b.shape}"
    ok = rel_close(a, b, rel)
    assert ok.all(), f"{what}: {_first_diff(ok, a, b)}"


def world_type(real):
    """ctypes mirror of arc_world (include/arc_world.h, the reference's
    csrc/arc_phys.h:25-112 layout) with `real` for every double: c_double
    for the library's snapshot and the FP64 reference build, c_float for the
    reference compiled with double as float."""
    C = ctypes
    f = lambda *names: [(n, real) for n in names]

    class V3(C.Structure):
        _fields_ = f("x", "y", "z")

    class Q(C.Structure):
        _fields_ = f("w", "x", "y", "z")

    class M3(C.Structure):
        _fields_ = [("m", real * 9)]

    class Body(C.Structure):
        _fields_ = [("x", V3), ("x_prev", V3), ("q", Q), ("q_prev", Q), ("v", V3), ("w", V3)] + \
                   f("mass", "inv_mass") + [("inertia", V3), ("inv_inertia", V3), ("rot", M3), ("inv_iw", M3)]

    class Sphere(C.Structure):
        _fields_ = [("body", C.c_int), ("pos", V3), ("radius", real), ("touch", C.c_int), ("anchored", C.c_int),
                    ("anchor", V3), ("anchor_n", V3)]

    class Joint(C.Structure):
        _fields_ = [("parent", C.c_int), ("child", C.c_int), ("r1", V3), ("r2", V3), ("a1", V3), ("a2", V3),
                    ("b1", V3), ("b2", V3), ("q_rel0", Q)] + \
                   f("lo", "hi", "kp", "kd", "max_torque", "target", "angle", "rate", "torque", "torque_acc") + \
                   [("lim", real * 4)]

    class Contact(C.Structure):
        _fields_ = [("body", C.c_int), ("local", V3), ("n", V3)]

    class Visual(C.Structure):
        _fields_ = [("type", C.c_int), ("body", C.c_int), ("p0", V3), ("p1", V3)] + \
                   f("radius", "radius2", "aspect") + [("material", C.c_int), ("color", C.c_float * 3)]

    class Terrain(C.Structure):
        _fields_ = [("nx", C.c_int), ("ny", C.c_int)] + f("cell", "ox", "oy", "outside") + [("h", real * 4096)]

    class World(C.Structure):
        _fields_ = [("gravity", V3)] + f("mu_static", "mu_dynamic") + \
                   [(n, C.c_int) for n in ("n_bodies", "n_joints", "n_spheres", "n_contacts", "n_visuals",
                                           "unstable")] + \
                   [("time", real), ("bodies", Body * 16), ("joints", Joint * 16), ("spheres", Sphere * 48),
                    ("contacts", Contact * 128), ("visuals", Visual * 80), ("terrain", Terrain)]

    return World


def struct_values(obj, skip=()) -> dict:
    """Flattens a ctypes structure into {dotted name: float or int}."""
    out = {}

    def walk(o, prefix):
        if isinstance(o, ctypes.Structure):
            for name, _ in o._fields_:
                if name not in skip:
                    walk(getattr(o, name), f"{prefix}{name}.")
        elif isinstance(o, ctypes.Array):
            for k in range(len(o)):
                walk(o[k], f"{prefix}{k}.")
        else:
            out[prefix[:-1]] = o

    walk(obj, "")
    return out


class TaskMirror:
    """The device's fall rules and auto-reset replayed on an oracle sim
    (arc/task.py:196-203 for the causes, arc/env.py:114-116 for the reset of
    the ended worlds): after each step every fallen world is reset with seed
    global_index + n_seed * episode (its episode count after this fall) and
    its own difficulty, and its torso-contact counter restarts, as the
    device's k_falls / k_reset do.  `index` maps the oracle's worlds to their
    indices in the device sim (a sample, or a client's range), `n_seed` is
    the device sim's world count."""

    def __init__(self, oracle_sim, index, n_seed, difficulty, beached=30, flip=0.3, void=-2.0):
        self.o = oracle_sim
        self.index = np.asarray(index, np.int64)
        self.n_seed = int(n_seed)
        self.diff = np.asarray(difficulty, np.float64)
        self.rules = (beached, flip, void)
        self.torso = np.zeros(len(self.index), np.int32)
        self.episode = np.zeros(len(self.index), np.int64)
        self.falls = 0

    def after_step(self) -> np.ndarray:
        cause = self.o.falls(self.torso, *self.rules)
        fell = np.flatnonzero(cause)
        if len(fell):
            self.episode[fell] += 1
            seeds = ((self.index[fell] + self.n_seed * self.episode[fell]) & 0xFFFFFFFF).astype(np.uint32)
            self.o.reset(fell.astype(np.int32), seeds, self.diff[fell])
            self.torso[fell] = 0
            self.falls += len(fell)
        return cause


class HostCheck:
    """The device engine's code compiled for the host (tests/hostcheck)."""

    def __init__(self, n, dt=1.0 / 30.0, substeps=20, grid=64, cell=0.4):
        L = ctypes.CDLL(str(HOSTCHECK))
        vp = ctypes.c_void_p
        L.hc_create.restype = vp
        L.hc_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.hc_destroy.argtypes = [vp]
        L.hc_reset.argtypes = [vp, vp, ctypes.c_int, vp, vp]
        L.hc_step.argtypes = [vp, vp]
        L.hc_state.argtypes = [vp, ctypes.c_int, vp]
        L.hc_terrain.argtypes = [vp, ctypes.c_int, vp]
        L.hc_error.argtypes = [vp, ctypes.c_int]
        L.hc_error.restype = ctypes.c_int
        self.L, self.n, self.grid = L, n, min(grid, 64)
        self.p = L.hc_create(n, dt, substeps, grid, cell)
        assert self.p

    def __del__(self):
        if getattr(self, "p", None):
            self.L.hc_destroy(self.p)

    def reset(self, ids, seeds, diff):
        ids = np.ascontiguousarray(ids, np.int32)
        seeds = np.ascontiguousarray(seeds, np.uint32)
        diff = np.ascontiguousarray(diff, np.float64)
        self.L.hc_reset(self.p, ids.ctypes.data, len(ids), seeds.ctypes.data, diff.ctypes.data)

    def step(self, targets):
        t = np.ascontiguousarray(targets, np.float64)
        self.L.hc_step(self.p, t.ctypes.data)

    def state(self, i):
        out = np.zeros(STATE)
        self.L.hc_state(self.p, i, out.ctypes.data)
        return out

    def states(self):
        return np.stack([self.state(i) for i in range(self.n)])

    def terrain(self, i):
        out = np.zeros(self.grid * self.grid + 1)
        self.L.hc_terrain(self.p, i, out.ctypes.data)
        return out

    def error(self, i):
        return self.L.hc_error(self.p, i)


class TwinWorld:
    """A device world (arc.native.World) and an FP32-oracle world given the
    same calls; each read returns the device's value once it is asserted to be
    the oracle's, bit for bit."""

    def __init__(self):
        from arc import native
        self.g, self.o = native.World(), P.Lib("o32").world()

    def __getattr__(self, name):
        fg, fo = getattr(self.g, name), getattr(self.o, name)

        def call(*args, **kw):
            rg, ro = fg(*args, **kw), fo(*args, **kw)
            if rg is not None or ro is not None:
                assert_same(rg, ro, f"{name}{args}")
            return rg

        return call

    @property
    def unstable(self):
        assert bool(self.g.unstable) == bool(self.o.unstable)
        return self.g.unstable

    @property
    def time(self):
        assert_same(self.g.time, self.o.time, "time")
        return self.g.time


class TwinSim:
    """A device sim (arc.native.Sim, in `form`) and an FP32-oracle sim in
    lockstep, with the oracle front end's method signatures; observe()
    asserts every output equal to the oracle's."""

    def __init__(self, n, grid=64, cell=0.4, form=None):
        from arc import native
        self.g = native.Sim(n, grid=grid, cell=cell)
        if form:
            self.g.set_form(form)
        self.o = P.Lib("o32").sim(n, grid=grid, cell=cell)

    def robot_info(self):
        d, lo, hi, misc = self.o.robot_info()
        assert np.array_equal(self.g.default_angles, d) and np.array_equal(self.g.lo, lo)
        assert np.array_equal(self.g.hi, hi) and self.g.stand_height == misc[0]
        return d, lo, hi, misc

    def reset(self, ids, seeds, difficulty):
        ids = np.asarray(ids, np.int32)
        self.g.reset(ids, np.asarray(seeds, np.uint32), np.asarray(difficulty, np.float64))
        self.o.reset(ids, seeds, difficulty)

    def step(self, targets, threads: int = 1):
        self.g.step(np.ascontiguousarray(targets, np.float64))
        self.o.step(targets, threads)

    def observe(self):
        root, joints, contacts, flags = (np.array(a) for a in self.g.observe())
        rot = np.array(self.g.rot)
        ref = self.o.observe()
        for name, a, b in zip(("root", "rot", "joints", "contacts", "flags"), (root, rot, joints, contacts, flags),
                              ref):
            assert_same(a, b, f"observe {name}")
        return root, rot, joints, contacts, flags
