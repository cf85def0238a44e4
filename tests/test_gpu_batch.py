"""GPU batched form: uneven request ranges in one launch, the persistent
work-queue distribution, hipGraph replays and the autotuner.

Every launch configuration runs the same per-world arithmetic, so the bar
here is bit-identity with the library's plain full-sim step (itself held to
the oracle by test_gpu_parity.py), plus the oracle directly for one batch."""

import numpy as np
import pytest

import pyoracle as P
from helpers import action_stream, assert_same

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from arc import native
    return native


def _sim(native, n, seed0=0):
    g = native.Sim(n)
    ids = np.arange(n, dtype=np.int32)
    g.reset(ids, (ids + seed0).astype(np.uint32), P.difficulties(1, n))
    return g


def _bits(sim, n):
    return np.stack([sim.state(i) for i in range(n)]).view(np.uint64)


def test_batch_steps_only_its_ranges(native):
    n = 200
    ranges = [(0, 5), (10, 27), (100, 64), (199, 1)]
    t = action_stream(n, 1)[0].astype(np.float32)
    dev = native.DeviceBuffer.from_array(t)
    g = _sim(native, n)
    before = _bits(g, n)
    g.step_batch([r[0] for r in ranges], [r[1] for r in ranges], dev.ptr)
    g.sync()
    after = _bits(g, n)
    full = _sim(native, n)
    full.step_device(dev.ptr)
    full.sync()
    ref = _bits(full, n)
    inside = np.zeros(n, bool)
    for s, c in ranges:
        inside[s:s + c] = True
    assert np.array_equal(after[inside], ref[inside])
    assert np.array_equal(after[~inside], before[~inside])
    # and the stepped worlds against the oracle
    o = P.Lib("o32").sim(n)
    ids = np.arange(n, dtype=np.int32)
    o.reset(ids, ids.astype(np.uint32), P.difficulties(1, n))
    o.step(t.astype(np.float64))
    a = after[inside].view(np.float64)
    b = np.stack([o.state(i) for i in np.flatnonzero(inside)])
    assert_same(a, b, "batched ranges vs oracle")


@pytest.mark.parametrize("form", ["lane", "leg"])
@pytest.mark.parametrize("block", [64, 128, 256])
def test_launch_configurations_bit_identical(native, block, form):
    n, steps = 300, 2
    ts = action_stream(n, steps).astype(np.float32)
    dev = native.DeviceBuffer.from_array(ts)
    per = n * 12 * 4
    base = _sim(native, n)
    base.set_form("lane")
    for k in range(steps):
        base.step_device(dev.ptr + k * per)
    base.sync()
    ref = _bits(base, n)
    for tile in (0, 16):
        for persistent in (False, True):
            g = _sim(native, n)
            g.set_form(form)
            g.set_launch(block, 0)
            g.set_layout(tile)
            g.set_persistent(persistent)
            # the group forms always stage the 16 x 16 tile; the layout switch is the lane form's
            assert g.launch()["block"] == block and g.launch()["tile"] == (tile if form == "lane" else 16)
            for k in range(steps):
                g.step_batch([0], [n], dev.ptr + k * per)
            g.sync()
            assert np.array_equal(_bits(g, n), ref), f"block {block} tile {tile} persistent {persistent}"


@pytest.mark.parametrize("persistent", [False, True])
def test_graph_replays_match_eager(native, persistent):
    """One captured graph, replayed with a different request table each step
    (and new target values uploaded between replays), against eager launches
    of the same batches; eager read-backs between the replays."""
    n = 256
    batches = [([0, 64], [32, 100]), ([5], [251]), ([0, 128, 200], [10, 60, 56]), ([0], [256]), ([77], [3])]
    t = native.DeviceBuffer(n * 12 * 4)
    gr, eg = _sim(native, n), _sim(native, n)
    for s in (gr, eg):
        s.set_persistent(persistent)
    for k, (starts, counts) in enumerate(batches):
        t.upload(action_stream(n, k + 1, seed=5)[k].astype(np.float32))
        gr.step_batch(starts, counts, t.ptr, graph=True)
        eg.step_batch(starts, counts, t.ptr, graph=False)
        gr.sync()
        eg.sync()
        assert np.array_equal(_bits(gr, n), _bits(eg, n)), f"replay {k}"
    assert gr.graph_captures() == 1


def test_batch_task_matches_step_task(native):
    """Full-range batches with the task (fall rules + auto-reset) captured in a
    graph against arcx_step_task, under tightened rules that terminate many
    worlds; and a partial batch applies the rules to its own worlds only."""
    n, steps = 128, 5
    ts = action_stream(n, steps, seed=3).astype(np.float32)
    dev = native.DeviceBuffer.from_array(ts)
    per = n * 12 * 4
    a, b = _sim(native, n), _sim(native, n)
    for s in (a, b):
        s.task_config(2, 0.97, -2.0)
    for k in range(steps):
        a.step_batch([0], [n], dev.ptr + k * per, task=True, graph=True)
        b.step_task(dev.ptr + k * per)
        a.sync()
        b.sync()
        assert np.array_equal(a.task_causes(), b.task_causes()), f"step {k}"
        assert np.array_equal(_bits(a, n), _bits(b, n)), f"step {k}"
    assert a.task_stats()[1:].sum() > 10
    assert np.array_equal(a.task_episodes(), b.task_episodes())
    # partial batch: worlds outside it keep their cause 0 and episode count
    c = _sim(native, n)
    c.task_config(2, 0.97, -2.0)
    for k in range(steps):
        c.step_batch([0], [n // 2], dev.ptr + k * per, task=True)
    c.sync()
    assert np.all(c.task_causes()[n // 2:] == 0) and np.all(c.task_episodes()[n // 2:] == 0)
    assert c.task_episodes()[: n // 2].sum() > 0


def test_autotune_restores_state(native):
    n = 512
    g = _sim(native, n)
    before = _bits(g, n)
    cfg, ms = g.autotune(2)
    assert np.array_equal(_bits(g, n), before)
    assert cfg["block"] in (64, 128, 256) and cfg["tile"] in (0, 16) and ms > 0.0
    assert g.launch() == cfg
    print(f"autotune at {n} worlds: {cfg}, {ms:.3f} ms/step")


def test_batch_rejects_bad_requests(native):
    n = 64
    g = _sim(native, n)
    dev = native.DeviceBuffer(n * 12 * 4)
    with pytest.raises(native.SimError):
        g.step_batch([0, 10], [20, 5], dev.ptr)  # overlap
    with pytest.raises(native.SimError):
        g.step_batch([60], [5], dev.ptr)  # past the end
    with pytest.raises(native.SimError):
        g.step_batch(list(range(257)), [0] * 257, dev.ptr)  # too many requests


@pytest.mark.parametrize("form,tile", [("leg", 16), ("leg", 0), ("pair", 16), ("shared", 16)])
def test_leg_form_matches_lane_form(native, form, tile):
    """Four lanes per world (one leg each) and two lanes per world (two legs
    each), torso replicated, hip joints chained through DPP broadcasts or
    solved by one wavefront per block through LDS (shared; 200 worlds leave a
    partial last block), against one world per lane: bit-identical states,
    flags, contact counts and joint outputs over several steps."""
    n, steps = 200, 4
    ts = action_stream(n, steps, seed=9).astype(np.float32)
    dev = native.DeviceBuffer.from_array(ts)
    per = n * 12 * 4
    lane, leg = _sim(native, n), _sim(native, n)
    lane.set_form("lane")
    leg.set_form(form)
    leg.set_layout(tile)
    lane.set_layout(tile)
    for k in range(steps):
        lane.step_device(dev.ptr + k * per)
        leg.step_device(dev.ptr + k * per)
        lane.sync()
        leg.sync()
        assert np.array_equal(_bits(leg, n), _bits(lane, n)), f"step {k}"
    assert np.array_equal(leg.observe()[3], lane.observe()[3])
    assert all(leg.world_error(i) == 0 for i in range(n))


def test_shared_roots_form_block_sizes(native):
    """The shared-root form with blocks of 128 threads (32 worlds) and 256
    (64 worlds) against one world per lane, 200 worlds (partial last blocks),
    bit for bit; launch shapes it cannot run are refused."""
    n, steps = 200, 3
    ts = action_stream(n, steps, seed=13).astype(np.float32)
    dev = native.DeviceBuffer.from_array(ts)
    per = n * 12 * 4
    lane = _sim(native, n)
    sims = []
    for block in (128, 256):
        g = _sim(native, n)
        g.set_form("shared")
        g.set_launch(block, 0)
        sims.append(g)
    with pytest.raises(ValueError):
        sims[0].set_launch(64, 0)
    with pytest.raises(ValueError):
        sims[0].set_persistent(True)
    for k in range(steps):
        lane.step_device(dev.ptr + k * per)
        lane.sync()
        for g in sims:
            g.step_device(dev.ptr + k * per)
            g.sync()
            assert np.array_equal(_bits(g, n), _bits(lane, n)), f"step {k}"


@pytest.mark.parametrize("form,persistent", [("leg", False), ("leg", True), ("pair", False), ("pair", True),
                                             ("shared", False)])
def test_leg_form_batches_and_graphs(native, form, persistent):
    n = 256
    batches = [([0, 64], [32, 100]), ([5], [251]), ([0, 128, 200], [10, 60, 56])]
    t = native.DeviceBuffer(n * 12 * 4)
    a, b = _sim(native, n), _sim(native, n)
    a.set_form(form)
    a.set_persistent(persistent)
    for k, (starts, counts) in enumerate(batches):
        t.upload(action_stream(n, k + 1, seed=6)[k].astype(np.float32))
        a.step_batch(starts, counts, t.ptr, task=True, graph=True)
        b.step_batch(starts, counts, t.ptr, task=True, graph=False)
        a.sync()
        b.sync()
        assert np.array_equal(_bits(a, n), _bits(b, n)), f"batch {k}"
        assert np.array_equal(a.task_causes(), b.task_causes())


def test_batch_rows_match_one_buffer(native):
    """Each request's targets in memory of its own (arcx_step_batch_rows, read
    where they lie) step the worlds exactly as the same targets in one
    world-major buffer: leg form replayed from one graph whose table changes
    its row addresses every step, and the lane and shared-root forms launched
    eagerly; with the fall rules and auto-reset on."""
    n, steps = 300, 3
    ts = action_stream(n, steps, seed=11).astype(np.float32)
    starts, counts = [0, 40, 170, 299], [25, 100, 90, 1]
    dev = native.DeviceBuffer.from_array(ts)
    per = n * 12 * 4
    bufs = [[native.DeviceBuffer.from_array(np.ascontiguousarray(ts[k, s0:s0 + c]))
             for s0, c in zip(starts, counts)] for k in range(steps)]
    ref, rows_graph, rows_lane, rows_shared = (_sim(native, n) for _ in range(4))
    rows_lane.set_form("lane")
    rows_shared.set_form("shared")
    for k in range(steps):
        ref.step_batch(starts, counts, dev.ptr + k * per, task=True, graph=True)
        rp = [b.ptr for b in bufs[k]]
        rows_graph.step_batch_rows(starts, counts, rp, task=True, graph=True)
        rows_lane.step_batch_rows(starts, counts, rp, task=True)
        rows_shared.step_batch_rows(starts, counts, rp, task=True)
        for s in (ref, rows_graph, rows_lane, rows_shared):
            s.sync()
        want = _bits(ref, n)
        for s, name in ((rows_graph, "graph"), (rows_lane, "lane"), (rows_shared, "shared")):
            assert np.array_equal(_bits(s, n), want), f"{name}, step {k}"
    assert rows_graph.graph_captures() == 1
    with pytest.raises(native.SimError):
        rows_graph.step_batch_rows([0], [5], [0])  # a request without targets
