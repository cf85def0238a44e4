"""Pins the oracle (oracle/arc_oracle.c) to the reference engine.

1. Golden vectors (tests/golden/*.npz, written by make_fixtures.py from the
   reference built from its own sources): the restatement reproduces them
   bit for bit in both precisions -- RNG streams, generated terrain, world
   state after reset and after 1 / 2 / 10 control steps, observe / heights
   outputs, single-world scenarios.
2. Where oracle/_ref exists (the build container), the same comparisons on
   fresh random scenarios against the live reference libraries.
A planted mistake in the oracle (a dropped term, a swapped operand, a wrong
threshold) changes these bits: the comparison is exact, not a tolerance.
"""

from pathlib import Path

import numpy as np
import pytest

import pyoracle as P
from helpers import same

GOLD = Path(__file__).resolve().parent / "golden"
CASES = ("d0", "d05", "d1", "mixed")


def _eq(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and bool(same(a, b).all())


def test_rng_golden():
    g = np.load(GOLD / "rng.npz")
    for kind in ("o64", "o32"):
        L = P.Lib(kind)
        for s in (0, 1, 7, 12345, 0xFFFFFFFF):
            assert np.array_equal(L.rng_stream(s, 4096), g[f"stream{s}"])
        for x in (0, 1, 2, 0x9E3779B9, 0xDEADBEEF):
            assert L.hash32(x) == int(g[f"hash{x}"])
    # binary32 rounding of u32 before scaling (float build) vs exact (double)
    assert np.array_equal(P.Lib("o32").rng_uniform(3, 1024), g["uniform32"])
    assert np.array_equal(P.Lib("o64").rng_uniform(3, 1024), g["uniform64"])


def test_rng_closed_form():
    """xorshift128 (Marsaglia 2003, shifts 11/8/19) with the splitmix-style
    seed hash, restated independently in numpy."""
    def h(x):
        x &= 0xFFFFFFFF
        x = ((x ^ (x >> 16)) * 0x7FEB352D) & 0xFFFFFFFF
        x = ((x ^ (x >> 15)) * 0x846CA68B) & 0xFFFFFFFF
        return x ^ (x >> 16)
    seed, s = 42, []
    for i in range(4):
        seed = h(seed + 0x9E3779B9 * (i + 1))
        s.append(seed or 1)
    out = []
    for _ in range(100):
        t = s[3]
        s[3], s[2], s[1] = s[2], s[1], s[0]
        t = (t ^ (t << 11)) & 0xFFFFFFFF
        t ^= t >> 8
        s[0] = t ^ s[0] ^ (s[0] >> 19)
        out.append(s[0])
    assert np.array_equal(P.Lib("o64").rng_stream(42, 100), np.array(out, np.uint32))


@pytest.mark.parametrize("prec", ["32", "64"])
@pytest.mark.parametrize("case", CASES)
def test_sim_golden(prec, case):
    g = np.load(GOLD / f"sim_{prec}.npz")
    L = P.Lib("o" + prec)
    n = g[f"{case}/state0"].shape[0]
    s = L.sim(n)
    ids = np.arange(n, dtype=np.int32)
    s.reset(ids, ids.astype(np.uint32), g[f"{case}/difficulty"])
    assert _eq(np.stack([s.terrain(i) for i in range(n)]), g[f"{case}/terrain"])
    assert _eq(s.states(), g[f"{case}/state0"])
    tg = g["targets"]
    for k in range(1, tg.shape[0] + 1):
        s.step(tg[k - 1])
        if f"{case}/state{k}" in g:
            assert _eq(s.states(), g[f"{case}/state{k}"]), f"state after step {k}"
    root, rot, joints, contacts, flags = s.observe()
    for name, v in (("root", root), ("rot", rot), ("joints", joints), ("contacts", contacts), ("flags", flags)):
        assert _eq(v, g[f"{case}/{name}"]), name
    assert _eq(s.heights(g["offsets"]), g[f"{case}/heights"])


@pytest.mark.parametrize("prec", ["32", "64"])
def test_robot_golden(prec):
    g = np.load(GOLD / f"sim_{prec}.npz")
    s = P.Lib("o" + prec).sim(1)
    assert _eq(np.concatenate(s.robot_info()), g["robot_info"])
    assert _eq(s.state(-1), g["template"])


@pytest.mark.parametrize("prec", ["32", "64"])
def test_kat_golden(prec):
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", GOLD / "make_fixtures.py")
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    got = mk.kat_fixture("o" + prec)
    g = np.load(GOLD / f"kat_{prec}.npz")
    for k in g.files:
        assert _eq(got[k], g[k]), k


# -- live comparison with the reference libraries ---------------------------

def _need(kind):
    if not P.available(kind):
        pytest.skip(f"{P.lib_path(kind)} not built (no reference tree on this machine)")


@pytest.mark.parametrize("prec", ["32", "64s"])
def test_live_reference_rollout(prec):
    """40 worlds, mixed difficulty, 30 steps, per-world seeds 1000.. : the
    restatement and the reference agree bit for bit at every step."""
    _need("r" + prec)
    okind = "o" + prec.rstrip("s")
    n, steps = 40, 30
    o, r = P.Lib(okind).sim(n), P.Lib("r" + prec).sim(n)
    diff = P.difficulties(5, n)
    ids = np.arange(n, dtype=np.int32)
    seeds = (ids + 1000).astype(np.uint32)
    o.reset(ids, seeds, diff)
    r.reset(ids, seeds, diff)
    u = P.Lib("o64").rng_stream(9, n * 12 * steps)
    d = P.Lib("o64").sim(1).robot_info()[0]
    tg = P.targets_from_actions(d, P.actions(u)).reshape(steps, n, 12)
    for k in range(steps):
        o.step(tg[k])
        r.step(tg[k])
        assert _eq(o.states(), r.states()), f"step {k + 1}"
    for a, b in zip(o.observe(), r.observe()):
        assert _eq(a, b)
    for i in range(n):
        assert o.num_contacts(i) == r.num_contacts(i)


@pytest.mark.parametrize("prec", ["32", "64s"])
def test_live_reference_partial_resets(prec):
    """Resets of a subset mid-rollout (the auto-reset path) and a grid of 48."""
    _need("r" + prec)
    okind = "o" + prec.rstrip("s")
    n = 12
    o, r = P.Lib(okind).sim(n, grid=48), P.Lib("r" + prec).sim(n, grid=48)
    ids = np.arange(n, dtype=np.int32)
    for s in (o, r):
        s.reset(ids, ids.astype(np.uint32) + 7, np.full(n, 0.7))
    u = P.Lib("o64").rng_stream(4, n * 12 * 8)
    d = P.Lib("o64").sim(1).robot_info()[0]
    tg = P.targets_from_actions(d, P.actions(u)).reshape(8, n, 12)
    for k in range(8):
        for s in (o, r):
            s.step(tg[k])
        if k == 3:
            sub = np.array([1, 5, 6], np.int32)
            for s in (o, r):
                s.reset(sub, np.array([99, 100, 101], np.uint32), np.array([0.0, 0.3, 1.0]))
        assert _eq(o.states(), r.states())


@pytest.mark.parametrize("prec", ["32", "64s"])
def test_live_reference_terrain_at(prec):
    """arc_sim_terrain_at over points inside, on edges of and outside the grid."""
    _need("r" + prec)
    okind = "o" + prec.rstrip("s")
    n = 6
    o, r = P.Lib(okind).sim(n), P.Lib("r" + prec).sim(n)
    ids = np.arange(n, dtype=np.int32)
    for s in (o, r):
        s.reset(ids, ids + 3, np.linspace(0, 1, n))
    m = 200
    rng = np.random.default_rng(0)
    pid = rng.integers(0, n, m).astype(np.int32)
    xy = rng.uniform(-14, 14, (m, 2)).astype(np.float32).astype(np.float64)
    xy[:10] = np.round(xy[:10] / 0.4) * 0.4  # column edges
    assert _eq(o.terrain_at(pid, xy), r.terrain_at(pid, xy))
