"""The allocation-free step (arcx_step_device_noalloc) against the step it
varies (arcx_step_device), through the C ABI on the GPU.

The two run the same kernels on the same inputs; only the scratch differs
(the library's static per-GPU array instead of the sim's own contact-list
buffers and work counter), so the bar is bit-exact equality of every world's
state, in every form and both work distributions, over two calls in a row,
the second with a smaller sim (the array then holds the first call's lists,
which the second must not read).  Beyond the limit (8,192 worlds) the call
is refused before anything is launched, and a call allocates no device
memory."""

import numpy as np
import pytest

import pyoracle as P
from helpers import action_stream, assert_same

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from arc import native
    return native


def _states(sim, n):
    return np.stack([sim.state(i) for i in range(n)])


def _sims(native, n, form, persistent, seed0):
    sims = []
    ids = np.arange(n, dtype=np.int32)
    for _ in range(2):
        g = native.Sim(n)
        g.set_form(form)
        if persistent:
            g.set_persistent(True)
        g.reset(ids, (ids + seed0).astype(np.uint32), P.difficulties(3, n))
        sims.append(g)
    return sims


@pytest.mark.parametrize("form,persistent", [("leg", False), ("leg", True), ("lane", False), ("lane", True),
                                             ("pair", False), ("pair", True), ("shared", False), ("wave", False)])
def test_noalloc_matches_step_device(native, form, persistent):
    steps = 2
    for n, seed0 in ((600, 0), (37, 5000)):  # the second call steps a smaller sim
        ref, dut = _sims(native, n, form, persistent, seed0)
        t = np.ascontiguousarray(action_stream(n, steps, seed=seed0 + 1).astype(np.float32))
        dev = native.DeviceBuffer.from_array(t)
        for k in range(steps):
            ref.step_device(dev.ptr + k * n * 48)
            dut.step_device_noalloc(dev.ptr + k * n * 48)
        ref.sync()
        dut.sync()
        print(f"{form}{' persistent' if persistent else ''}, {n} worlds, {steps} steps")
        assert_same(_states(dut, n), _states(ref, n), f"{form}, persistent {persistent}, {n} worlds")
        ref.close()
        dut.close()


def test_noalloc_refuses_more_worlds_than_it_holds(native):
    n = native.NOALLOC_MAX_WORLDS + 1
    g = native.Sim(n)
    ids = np.arange(n, dtype=np.int32)
    g.reset(ids, ids.astype(np.uint32), np.full(n, 0.5))
    before = g.state(n - 1)
    dev = native.DeviceBuffer(n * 48)
    dev.upload(np.tile(g.default_angles.astype(np.float32), n))
    with pytest.raises(native.SimError, match="static scratch"):
        g.step_device_noalloc(dev.ptr)
    g.sync()
    assert_same(g.state(n - 1), before, "state after a refused call")
    g.close()


def test_noalloc_allocates_no_device_memory(native):
    """Free device memory is unchanged across allocation-free steps of the
    largest sim it takes, in the persistent distribution (whose regular
    entry point allocates a work counter and request tables on first use)."""
    import torch
    n = native.NOALLOC_MAX_WORLDS
    warm = native.Sim(64)
    warm.set_persistent(True)
    g = native.Sim(n)
    g.set_persistent(True)
    ids = np.arange(n, dtype=np.int32)
    g.reset(ids, ids.astype(np.uint32), np.full(n, 0.5))
    dev = native.DeviceBuffer(n * 48)
    dev.upload(np.tile(g.default_angles.astype(np.float32), n))
    warm.step_device(dev.ptr)  # loads the step kernels' code
    warm.sync()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        g.step_device_noalloc(dev.ptr)
    g.sync()
    free1 = torch.cuda.mem_get_info()[0]
    g.step_device(dev.ptr)  # the regular entry point: the persistent distribution's buffers
    g.sync()
    free2 = torch.cuda.mem_get_info()[0]
    print(f"free device memory: {free0} -> {free1} (no-alloc steps) -> {free2} (regular step)")
    assert free1 == free0
    assert free2 <= free1
