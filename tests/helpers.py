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
