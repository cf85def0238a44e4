"""The device engine's own code (lane_engine.h / reset_engine.h / setup.cpp /
model.cpp), compiled for the host by tests/hostcheck, against the FP32 oracle
and the FP32 golden vectors.  Same binary32 operations in the same order, and
on the host the same glibc atan2/sin/cos: the comparison is bit for bit.
This pins the engine's algorithm on the CPU; the GPU tests then only have to
show the device build computes what this host build computes."""

from pathlib import Path

import numpy as np
import pytest

import pyoracle as P
from helpers import HOSTCHECK, HostCheck, action_stream, same

GOLD = Path(__file__).resolve().parent / "golden"

pytestmark = pytest.mark.skipif(not HOSTCHECK.exists(), reason="tests/hostcheck not built (run build())")


@pytest.mark.parametrize("case", ["d0", "d05", "d1", "mixed"])
def test_golden_fp32(case):
    g = np.load(GOLD / "sim_32.npz")
    n = g[f"{case}/state0"].shape[0]
    h = HostCheck(n)
    ids = np.arange(n)
    h.reset(ids, ids, g[f"{case}/difficulty"])
    assert same(np.stack([h.terrain(i) for i in range(n)]), g[f"{case}/terrain"].astype(np.float64)).all()
    assert same(h.states(), g[f"{case}/state0"]).all()
    tg = g["targets"]
    for k in range(1, tg.shape[0] + 1):
        h.step(tg[k - 1])
        if f"{case}/state{k}" in g:
            assert same(h.states(), g[f"{case}/state{k}"]).all(), f"step {k}"


def test_template():
    g = np.load(GOLD / "sim_32.npz")
    h = HostCheck(1)
    assert same(h.state(0), g["template"]).all()


@pytest.mark.parametrize("grid,cell", [(64, 0.4), (48, 0.4), (32, 0.25)])
def test_rollout_vs_oracle(grid, cell):
    """30 steps, 24 worlds, mixed difficulty, partial re-resets mid-way,
    other grid sizes and a finer cell (more columns per sphere)."""
    n, steps = 24, 30
    o = P.Lib("o32").sim(n, grid=grid, cell=cell)
    h = HostCheck(n, grid=grid, cell=cell)
    ids = np.arange(n)
    diff = P.difficulties(11, n)
    o.reset(ids, ids + 500, diff)
    h.reset(ids, ids + 500, diff)
    ts = action_stream(n, steps, seed=3)
    for k in range(steps):
        o.step(ts[k])
        h.step(ts[k])
        if k == 12:
            sub = np.array([0, 7, 8, 23])
            for s in (o, h):
                s.reset(sub, sub + 77, np.array([0.0, 0.2, 0.9, 1.0]))
        assert same(h.states(), o.states()).all(), f"step {k + 1}"
    assert all(h.error(i) == 0 for i in range(n))


def test_unstable_world_frozen():
    """A world pushed to NaN is flagged and then left alone (arc_sim.c:135-136)."""
    n = 4
    o = P.Lib("o32").sim(n)
    h = HostCheck(n)
    ids = np.arange(n)
    for s in (o, h):
        s.reset(ids, ids, np.full(n, 0.5))
    t = action_stream(n, 3)
    t[0, 2, :] = np.nan   # NaN targets clamp to NaN and poison world 2's drives
    for k in range(3):
        o.step(t[k])
        h.step(t[k])
        assert same(h.states(), o.states()).all()


@pytest.mark.parametrize("grid,cell", [(64, 0.4), (48, 0.4), (20, 0.25), (8, 0.4)])
def test_reset_split_matches_sequential(grid, cell):
    """The device's reset: draws made by 256 lanes jumping ahead in the
    xorshift128 stream (each lane's GF(2) matrix power applied from 4-bit
    window tables), and the footprint scan split over the lanes, equal the
    sequential stream and scan bit for bit."""
    import ctypes
    h = HostCheck(1, grid=grid, cell=cell)
    h.L.hc_reset_split_check.restype = ctypes.c_int
    h.L.hc_reset_split_check.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_int]
    for seed in (0, 1, 12345, 4294967295):
        for gen in (1, 0):
            assert h.L.hc_reset_split_check(h.p, seed, gen) == 0, (seed, gen)


def test_column_quotient_sequence():
    """The column windows' three-instruction quotient (q0 = a y, r = d q0 - a,
    q = q0 - r y with y = RN(1/d), two fused multiply-adds), evaluated on the
    host with the same correctly rounded FMA: equal to the IEEE quotient for
    every sampled numerator of its domain (zero, 2^-96 <= |a| < 2^100), for
    the column sizes the tests and the benchmark use; below 2^-96 the residual
    underflows and the sequence does differ for some numerators, which is why a
    sim enables it only where its numerators stay in the domain (the device
    check then covers all 2^32 patterns)."""
    import ctypes
    L = ctypes.CDLL(str(HOSTCHECK))
    f = L.hc_quotient_mismatches
    f.restype = ctypes.c_ulonglong
    f.argtypes = [ctypes.c_float, ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_int]
    for cell in (0.4, 0.25, 0.37):
        assert f(cell, 0, 1 << 32, 1009, 1) == 0, cell
        # every pattern of one binade in the middle of the domain, both signs
        assert f(cell, 0x3F800000, 0x40000000, 1, 1) == 0
        assert f(cell, 0xBF800000, 0xC0000000, 1, 1) == 0
    # below the domain the sequence is not exact (tiny numerators of both signs)
    assert f(0.4, 0, 0x0F800000, 3, 0) > 0
