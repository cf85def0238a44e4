"""Edge cases of the step path on the GPU, through the C ABI.

- non-finite targets: NaN passes the reference's clamp and makes its world
  non-finite and unstable (frozen from then on), +-inf clamp to the joint
  limits; every form against the oracle (NaN == NaN), flags exact;
- host addresses where device targets are expected: refused before any
  kernel could read them, the sim left as it was;
- arc_sim_observe with any subset of its outputs (the others NULL, as the
  reference allows): each requested output equals the full read-out's, and
  the buffers not requested are left untouched;
- a sim of no worlds: create, step, observe, heights and a batch of no
  requests do nothing and report no error (the reference's loops are empty);
- world counts that fill no whole wavefront or quad group (1, 3, 17, 100):
  every form (lane, four lanes per world, two lanes per world, four lanes per
  world with shared root solves, one world per wavefront) bit-exact with the
  FP32 oracle after one step;
- reset inputs: an empty list (nothing changes), repeated ids (the last entry
  wins, as resetting in list order does), ids out of range (skipped, the
  others reset, the error reported), difficulties beyond the int8 terrain's
  range (skipped and reported; 6 itself exact against the oracle);
- the largest batch: 256 requests of uneven sizes, some of one world, step
  exactly what one step of all worlds does;
- the full size, 65,536 worlds: the forms agree bit for bit after a few steps
  and a repeated run gives the same bits (size-independent properties, at the
  shape the benchmark runs).
"""

import hashlib

import numpy as np
import pytest

import pyoracle as P
from helpers import action_stream, assert_same, same

pytestmark = pytest.mark.gpu

@pytest.fixture(scope="module")
def native():
    from arc import native
    return native


def _states(sim, n):
    return np.stack([sim.state(i) for i in range(n)])


def test_non_finite_targets(native):
    n, steps = 16, 3
    diff = P.difficulties(5, n)
    ids = np.arange(n, dtype=np.int32)
    ts = action_stream(n, steps, seed=8)
    ts[0, 2, 4] = np.nan
    ts[0, 5, :] = np.inf
    ts[0, 7, 1] = -np.inf
    ts[1, 9, 0] = np.nan
    o = P.Lib("o32").sim(n)
    o.reset(ids, ids.astype(np.uint32), diff)
    want = []
    for k in range(steps):
        o.step(ts[k])
        want.append((_states(o, n), o.observe()[4].copy()))
    assert want[-1][1][2] and want[-1][1][9] and not want[-1][1][5]  # NaN worlds unstable, inf world clamped
    for form in ("lane", "leg", "pair", "shared"):
        g = native.Sim(n)
        g.set_form(form)
        g.reset(ids, ids.astype(np.uint32), diff)
        for k in range(steps):
            g.step(ts[k])
            a, flags = _states(g, n), g.observe()[3]
            assert same(a, want[k][0]).all(), f"{form}, step {k + 1}"
            assert np.array_equal(flags, want[k][1]), f"{form}, step {k + 1}"
        g.close()


def test_host_addresses_refused(native):
    n = 8
    g = native.Sim(n)
    ids = np.arange(n, dtype=np.int32)
    g.reset(ids, ids.astype(np.uint32), np.full(n, 0.4))
    before = _states(g, n)
    host = np.ascontiguousarray(action_stream(n, 1)[0].astype(np.float32))
    # device memory too short for what the kernels would read: 12 floats
    # short of n x 12 (a whole allocation of libarcsim's, so its extent is
    # exactly n x 12 - 12 floats)
    short = native.DeviceBuffer(4 * 12 * (n - 1))
    full = native.DeviceBuffer.from_array(host)
    for call in (lambda: g.step_device(host.ctypes.data), lambda: g.step_task(host.ctypes.data),
                 lambda: g.time_steps(host.ctypes.data, 1, 0),
                 lambda: g.step_batch([0], [n], host.ctypes.data),
                 lambda: g.step_batch_rows([0, 4], [4, 4], [full.ptr, host.ctypes.data]),
                 lambda: g.step_device(short.ptr), lambda: g.step_task(short.ptr),
                 lambda: g.time_steps(full.ptr, 2, n * 12),
                 lambda: g.step_batch([0], [n], short.ptr),
                 lambda: g.step_batch_rows([0, 4], [4, 4], [full.ptr, short.ptr + 4 * 12 * (n - 4)])):
        with pytest.raises(native.SimError):
            call()
    g.sync()
    assert same(_states(g, n), before).all()
    # the same entry points with ranges that fit are accepted
    g.step_batch_rows([0, 4], [4, 4], [full.ptr, short.ptr + 4 * 12 * (n - 5)])
    g.step_batch([0], [n - 1], short.ptr)
    g.sync()
    g.close()


def test_observe_output_subsets(native):
    n = 37
    g = native.Sim(n)
    ids = np.arange(n, dtype=np.int32)
    g.reset(ids, ids.astype(np.uint32), P.difficulties(4, n))
    g.step(action_stream(n, 1)[0])
    root, joints, contacts, flags = (x.copy() for x in g.observe())
    rot = g.rot.copy()
    full = [root, rot, joints, contacts, flags]
    for k in range(5):
        bufs = [np.full_like(x, 7.0) if x.dtype != np.int32 else np.full_like(x, 7) for x in full]
        args = [b.ctypes.data if j == k else None for j, b in enumerate(bufs)]
        g._lib.arc_sim_observe(g.ptr, *args)
        for j, b in enumerate(bufs):
            if j == k:
                assert np.array_equal(b, full[j]), f"output {j}"
            else:
                assert np.all(b == 7), f"output {j} written though not requested"
    g.close()


def test_no_worlds(native):
    for form in ("lane", "leg", "shared"):
        g = native.Sim(0)
        g.set_form(form)
        g.reset(np.zeros(0, np.int32), np.zeros(0, np.uint32), np.zeros(0))
        g.step(np.zeros((0, 12)))
        root, joints, contacts, flags = g.observe()
        assert root.shape == (0, 13) and joints.shape == (0, 36) and flags.shape == (0,)
        assert g.heights(np.zeros((3, 2))).shape == (0, 3)
        buf = native.DeviceBuffer(64)
        g.step_batch([], [], buf.ptr, task=True)
        g.sync()
        g.close()


@pytest.mark.parametrize("n", [1, 3, 17, 100])
def test_ragged_world_counts_all_forms(native, n):
    diff = P.difficulties(2, n)
    ids = np.arange(n, dtype=np.int32)
    seeds = (np.arange(n) + 7).astype(np.uint32)
    t = action_stream(n, 1)[0]
    o = P.Lib("o32").sim(n)
    o.reset(ids, seeds, diff)
    o.step(t)
    b = _states(o, n)
    ref = None
    for form in ("lane", "leg", "pair", "shared", "wave"):
        g = native.Sim(n)
        g.set_form(form)
        g.reset(ids, seeds, diff)
        g.step(t)
        a = _states(g, n)
        assert_same(a, b, f"{form}, {n} worlds vs oracle")
        if ref is None:
            ref = a
        else:
            assert same(a, ref).all(), f"{form} differs from the lane form at {n} worlds"
        g.close()


def test_reset_edge_inputs(native):
    n = 8
    g = native.Sim(n)
    ids = np.arange(n, dtype=np.int32)
    g.reset(ids, ids.astype(np.uint32), np.full(n, 0.6))
    before = _states(g, n)
    terr = [g.terrain(i) for i in range(n)]
    # empty list: nothing changes
    g.reset(np.zeros(0, np.int32), np.zeros(0, np.uint32), np.zeros(0))
    assert same(_states(g, n), before).all()
    # repeated ids: the last entry per world wins
    g.reset(np.array([2, 5, 2], np.int32), np.array([100, 101, 102], np.uint32), np.array([0.3, 0.9, 1.0]))
    h = native.Sim(n)
    h.reset(ids, ids.astype(np.uint32), np.full(n, 0.6))
    h.reset(np.array([2, 5], np.int32), np.array([102, 101], np.uint32), np.array([1.0, 0.9]))
    assert same(_states(g, n), _states(h, n)).all()
    assert np.array_equal(g.terrain(2), h.terrain(2)) and not np.array_equal(g.terrain(2), terr[2])
    # out-of-range ids are skipped and reported; the valid ones are reset
    with pytest.raises(native.SimError):
        g.reset(np.array([-1, 3, n], np.int32), np.array([1, 55, 2], np.uint32), np.array([0.5, 0.5, 0.5]))
    h.reset(np.array([3], np.int32), np.array([55], np.uint32), np.array([0.5]))
    assert same(_states(g, n), _states(h, n)).all()
    # difficulty above 6 does not fit the device terrain: that entry is skipped and reported
    with pytest.raises(native.SimError):
        g.reset(np.array([1, 4], np.int32), np.array([9, 10], np.uint32), np.array([6.5, 0.7]))
    h.reset(np.array([4], np.int32), np.array([10], np.uint32), np.array([0.7]))
    assert same(_states(g, n), _states(h, n)).all()
    g.close()
    h.close()
    # the steepest accepted terrain, against the oracle
    o = P.Lib("o32").sim(4)
    g = native.Sim(4)
    for s in (o, g):
        s.reset(np.arange(4, dtype=np.int32), np.array([3, 77, 1234, 99999], np.uint32), np.full(4, 6.0))
    for i in range(4):
        assert np.array_equal(g.terrain(i), o.terrain(i)), f"terrain of world {i} at difficulty 6"
    g.close()


def test_largest_batch_equals_full_step(native):
    rng = np.random.default_rng(5)
    counts = rng.integers(1, 40, size=256)
    counts[::7] = 1  # single-world requests among them
    n = int(counts.sum())
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    order = rng.permutation(256)  # requests in any order
    ids = np.arange(n, dtype=np.int32)
    diff = P.difficulties(3, n)
    t = action_stream(n, 1)[0].astype(np.float32)
    for form in ("lane", "leg", "shared"):
        g = native.Sim(n)
        h = native.Sim(n)
        for s in (g, h):
            s.set_form(form)
            s.reset(ids, ids.astype(np.uint32), diff)
        buf = native.DeviceBuffer.from_array(np.ascontiguousarray(t))
        g.step_batch(starts[order], counts[order], buf.ptr)
        g.sync()
        h.step(t.astype(np.float64))
        assert same(_states(g, n), _states(h, n)).all(), form
        g.close()
        h.close()


def _run_full(native, form, steps, n=65536):
    s = native.Sim(n)
    s.set_form(form)
    u = native.rng_stream(1, n)
    s.reset(np.arange(n), np.arange(n), (u >> 8) * (1.0 / 16777216.0))
    a = (native.rng_stream(0, n * 12 * steps) >> 8) * (1.0 / 8388608.0) - 1.0
    host = (np.tile(s.default_angles, n * steps) + 0.6 * a).astype(np.float32)
    buf = native.DeviceBuffer.from_array(host)
    # step k reads floats [k n 12, (k + 1) n 12): all inside the buffer
    s.time_steps(buf.ptr, steps, n * 12)
    root, joints, contacts, flags = s.observe()
    h = hashlib.sha256()
    for x in (root, joints, contacts, flags):
        h.update(np.ascontiguousarray(x).tobytes())
    s.close()
    return h.hexdigest(), int(flags.sum()), float(np.abs(root[:, 7:10]).mean())


def test_full_size_forms_agree_and_repeat(native):
    ref = _run_full(native, "leg", 4)
    assert ref[1] == 0 and ref[2] > 0.0  # stable, moving
    assert _run_full(native, "leg", 4) == ref
    assert _run_full(native, "lane", 4) == ref
    assert _run_full(native, "pair", 4) == ref
    assert _run_full(native, "shared", 4) == ref
