"""GPU parity: the HIP engine through the arc_sim C ABI against the oracle.

The oracle here is the FP32 restatement (bit-exact with the reference engine
compiled in FP32, see test_oracle_ref.py).  The device engine evaluates the
same binary32 operations in the same order.  Its atan2 / sin / cos are the
device library's binary64 functions rounded once to binary32, as the FP32
reference rounds glibc's; two binary64 results an ulp or so apart round to
different binary32 values only when the exact value lies within that
distance of a rounding boundary (about 2^-29 of calls).

Bars: bit-exact (NaN == NaN, -0 == +0) for every state field, the observe
outputs, heights, terrain and the integer flags after reset and after one,
two, three and ten control steps; longer rollouts are compared in
test_gpu_rollout.py.
"""

import ctypes

import numpy as np
import pytest

import pyoracle as P
from helpers import STATE, action_stream, assert_same, max_rel, same

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from arc import native
    return native


def _pair(native, n, diff, seed0=0):
    g = native.Sim(n)
    o = P.Lib("o32").sim(n)
    ids = np.arange(n, dtype=np.int32)
    seeds = (np.arange(n) + seed0).astype(np.uint32)
    g.reset(ids, seeds, diff)
    o.reset(ids, seeds, diff)
    return g, o


def _states(sim, n):
    return np.stack([sim.state(i) for i in range(n)])


def test_library_loads_native(native):
    lib = native.sim_lib()
    assert lib.arcx_device_count() >= 1
    assert lib.arc_sim_num_joints() == 12 and lib.arc_sim_num_legs() == 4


def test_fast_reciprocal_exhaustive(native):
    """The step's 1 / x (one Newton step on the hardware reciprocal where that
    is exact, the IEEE division elsewhere) equals the IEEE quotient for every
    one of the 2^32 binary32 bit patterns."""
    bad = ctypes.c_ulonglong(12345)
    assert native.sim_lib().arcx_check_recip(ctypes.addressof(bad)) == 0
    assert bad.value == 0


@pytest.mark.parametrize("grid,cell", [(64, 0.4), (32, 0.25), (48, 0.37)])
def test_fast_column_quotient_checked(native, grid, cell):
    """A sim divides by its column size with the three-instruction quotient
    exactly when the exhaustive device check finds no numerator of the
    quotient's domain (zero, 2^-96 <= |a| < 2^100) where it differs from the
    IEEE quotient; for these column sizes the check passes, so the step kernels
    take the fast path (and the parity tests cover it)."""
    lib = native.sim_lib()
    s = native.Sim(4, grid=grid, cell=cell)
    try:
        mask = lib.arcx_fast_divisors(s.ptr)
    finally:
        s.close()
    bad = ctypes.c_ulonglong(12345)
    assert lib.arcx_check_quotient(ctypes.c_float(cell), ctypes.addressof(bad)) == 0
    assert bad.value == 0
    assert mask == 1


def test_template_matches_oracle(native):
    g = native.Sim(4)
    o = P.Lib("o32").sim(4)
    a = _states(g, 4)
    b = np.stack([o.state(-1)] * 4)
    assert same(a, b).all()
    d, lo, hi, misc = o.robot_info()
    assert np.array_equal(g.default_angles, d) and np.array_equal(g.lo, lo) and np.array_equal(g.hi, hi)
    assert g.stand_height == misc[0] and g.total_mass == misc[1]


@pytest.mark.parametrize("difficulty", [-0.5, 0.0, 1e-7, 0.5, 1.0, 1.7, "mixed"])
def test_reset_bit_exact(native, difficulty):
    """Terrain and pose bit-exact after a reset: at and below zero difficulty
    the terrain is flat and draws nothing; a tiny positive one draws the full
    stream; above one the amplitudes only grow."""
    n = 48
    diff = P.difficulties(1, n) if difficulty == "mixed" else np.full(n, difficulty)
    g, o = _pair(native, n, diff)
    for i in range(n):
        assert np.array_equal(g.terrain(i), o.terrain(i)), f"terrain of world {i}"
    a, b = _states(g, n), _states(o, n)
    # the pose uses binary64 sin/cos rounded to binary32
    assert_same(a, b, f"state after reset, difficulty {difficulty}")


def test_reset_seed_extremes(native):
    """Seeds at the ends of the 32-bit range (the hash seeding and the jump
    tables treat every seed alike): terrain and state after reset, then one
    step, against the oracle."""
    seeds = np.array([0, 1, 2 ** 31 - 1, 2 ** 31, 2 ** 32 - 2, 2 ** 32 - 1], dtype=np.uint32)
    n = len(seeds)
    diff = np.full(n, 0.8)
    ids = np.arange(n, dtype=np.int32)
    g = native.Sim(n)
    o = P.Lib("o32").sim(n)
    g.reset(ids, seeds, diff)
    o.reset(ids, seeds, diff)
    for i in range(n):
        assert np.array_equal(g.terrain(i), o.terrain(i)), f"terrain of seed {seeds[i]}"
    t = action_stream(n, 1, seed=3)[0]
    g.step(t)
    o.step(t)
    assert_same(_states(g, n), _states(o, n), "state after one step")


@pytest.mark.parametrize("difficulty", [0.0, 0.5, "mixed"])
def test_one_step_parity(native, difficulty):
    n = 128
    diff = P.difficulties(1, n) if difficulty == "mixed" else np.full(n, difficulty)
    g, o = _pair(native, n, diff)
    t = action_stream(n, 1)[0]
    g.step(t)
    o.step(t)
    a, b = _states(g, n), _states(o, n)
    print(f"one step, difficulty {difficulty}: bit-exact worlds {same(a, b).all(1).mean():.3f}, "
          f"max rel err {max_rel(a, b):.3g}")
    # every field, the integers (flags, touch counts, anchors' validity) among them
    assert_same(a, b, f"state after one step, difficulty {difficulty}")
    # the observe ABI agrees with the state read-back
    root, joints, contacts, flags = g.observe()
    ro, rt, jo, co, fo = o.observe()
    assert_same(root, ro, "observe root")
    assert_same(g.rot, rt, "observe rot")
    assert_same(joints, jo, "observe joints")
    assert_same(contacts, co, "observe contacts")
    assert_same(flags, fo, "observe flags")


def test_ten_steps_track(native):
    """Ten control steps: chaos would amplify any last-bit difference
    ~1e2-1e4x per step, so bit-exact states here show that no step rounded
    differently."""
    n = 64
    diff = P.difficulties(1, n)
    g, o = _pair(native, n, diff)
    ts = action_stream(n, 10)
    for t in ts:
        g.step(t)
        o.step(t)
    a, b = _states(g, n), _states(o, n)
    print(f"ten steps: bit-exact worlds {same(a, b).all(1).mean():.3f}, max rel err {max_rel(a, b):.3g}")
    assert_same(a, b, "state after ten steps")


def test_heights_and_terrain_at(native):
    n = 16
    diff = np.full(n, 0.8)
    g, o = _pair(native, n, diff)
    t = action_stream(n, 2)
    for k in range(2):
        g.step(t[k])
        o.step(t[k])
    offs = np.array([(f, l) for f in np.arange(-1.0, 1.41, 0.2) for l in np.arange(-1.0, 1.01, 0.2)])
    offs = offs.astype(np.float32).astype(np.float64)
    hg, ho = g.heights(offs), o.heights(offs)
    assert_same(hg, ho, "heights")
    ids = np.arange(n, dtype=np.int32)
    xy = np.stack([np.linspace(-13, 13, n), np.linspace(5, -5, n)], 1).astype(np.float32).astype(np.float64)
    assert np.array_equal(g.terrain_at(ids, xy), o.terrain_at(ids, xy))


@pytest.mark.parametrize("difficulty", [0.0, "mixed"])
def test_wave_form_matches_lane_form(native, difficulty):
    """The plain form (one world per wavefront, joints by dependency level,
    generic arithmetic) computes the lane form's results: equal values,
    signed zeros aside; and both equal the oracle's."""
    n = 96
    diff = P.difficulties(1, n) if difficulty == "mixed" else np.full(n, difficulty)
    lane, o = _pair(native, n, diff)
    lane.set_form("lane")
    wave = native.Sim(n)
    wave.set_form("wave")
    ids = np.arange(n, dtype=np.int32)
    wave.reset(ids, ids.astype(np.uint32), diff)
    ts = action_stream(n, 3)
    for k in range(3):
        lane.step(ts[k])
        wave.step(ts[k])
        o.step(ts[k])
        a, b = _states(wave, n), _states(lane, n)
        assert_same(a, b, f"wave vs lane after step {k + 1}")
        assert_same(a, _states(o, n), f"wave vs oracle after step {k + 1}")


@pytest.mark.parametrize("form", ["lane", "leg", "pair", "shared"])
@pytest.mark.parametrize("grid,cell", [(8, 0.4), (16, 0.4), (20, 0.25)])
def test_small_grids(native, grid, cell, form):
    """Grids where the robot's legs reach the border and beyond, in every
    register-resident form: on the 8-column grid no LDS tile fits and every
    column is read from HBM; on the others the tile is clamped to the padded
    grid and the legs read columns outside it."""
    n = 64
    diff = P.difficulties(1, n)
    g = native.Sim(n, grid=grid, cell=cell)
    g.set_form(form)
    o = P.Lib("o32").sim(n, grid=grid, cell=cell)
    ids = np.arange(n, dtype=np.int32)
    g.reset(ids, ids.astype(np.uint32), diff)
    o.reset(ids, ids.astype(np.uint32), diff)
    for i in range(n):
        assert np.array_equal(g.terrain(i), o.terrain(i)), f"terrain of world {i}"
    t = action_stream(n, 1)[0]
    g.step(t)
    o.step(t)
    assert_same(_states(g, n), _states(o, n), f"{form}, grid {grid} cell {cell}")


@pytest.mark.parametrize("form", ["lane", "leg", "pair", "shared"])
@pytest.mark.parametrize("dt,substeps,grid,cell", [(1.0 / 60.0, 10, 64, 0.4), (1.0 / 20.0, 32, 48, 0.5)])
def test_other_timesteps_and_grids(native, form, dt, substeps, grid, cell):
    """Step length, substep count, grid and cell other than the defaults: the
    compiled-in robot constants do not depend on them, the Model does."""
    n = 64
    diff = P.difficulties(2, n)
    g = native.Sim(n, dt=dt, substeps=substeps, grid=grid, cell=cell)
    g.set_form(form)
    o = P.Lib("o32").sim(n, dt, substeps, grid, cell)
    ids = np.arange(n, dtype=np.int32)
    g.reset(ids, (ids + 500).astype(np.uint32), diff)
    o.reset(ids, (ids + 500).astype(np.uint32), diff)
    ts = action_stream(n, 3, seed=4)
    for k in range(3):
        g.step(ts[k])
        o.step(ts[k])
        assert_same(_states(g, n), _states(o, n), f"{form}, dt {dt}, {substeps} substeps, step {k + 1}")


@pytest.mark.parametrize("form", ["lane", "wave"])
def test_fine_cells(native, form):
    """Cells of 10 cm: the robot's sphere windows add up to more than the
    128-record contact cap, so records follow the world's sphere order under
    one running total (per-body lists in the lane form, a single list in the
    plain form); the group forms, which need every window within the cap,
    refuse such a robot."""
    n = 48
    diff = P.difficulties(3, n)
    g = native.Sim(n, grid=64, cell=0.1)
    g.set_form(form)
    o = P.Lib("o32").sim(n, grid=64, cell=0.1)
    ids = np.arange(n, dtype=np.int32)
    g.reset(ids, ids.astype(np.uint32), diff)
    o.reset(ids, ids.astype(np.uint32), diff)
    for k, t in enumerate(action_stream(n, 5, seed=6)):
        g.step(t)
        o.step(t)
        assert_same(_states(g, n), _states(o, n), f"{form}, cell 0.1, step {k + 1}")
    with pytest.raises((native.SimError, ValueError)):
        g.set_form("leg")
