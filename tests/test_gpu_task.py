"""GPU task layer: fall rules and auto-reset on the device against the
oracle's restatement of arc/task.py:196-203 and arc/env.py:114-116.

Falls are rare under random actions (0.2-1.3 per 1000 world-steps), so the
test tightens the rules (flip when the torso's up axis tilts past z < 0.97,
beached after 2 steps of torso contact) to make many worlds terminate and
reset within a few steps.  Causes and reset seeds (seed = i + N * episode)
are integers and must agree exactly; states bit for bit."""

import numpy as np
import pytest

import pyoracle as P
from helpers import action_stream, assert_same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["lane", "leg", "pair", "shared"])
def test_falls_and_autoreset_match_oracle(form):
    from arc import native

    n, steps = 96, 6
    beached, flip, void = 2, 0.97, -2.0
    diff = P.difficulties(1, n)
    g = native.Sim(n)
    g.set_form(form)
    o = P.Lib("o32").sim(n)
    ids = np.arange(n, dtype=np.int32)
    g.reset(ids, ids.astype(np.uint32), diff)
    o.reset(ids, ids.astype(np.uint32), diff)
    g.task_config(beached, flip, void)
    ts = action_stream(n, steps, seed=2)
    dev = native.DeviceBuffer.from_array(ts.astype(np.float32).reshape(steps, -1))
    per = n * 12 * 4
    torso = np.zeros(n, np.int32)
    episode = np.zeros(n, np.int64)
    total = 0
    for k in range(steps):
        g.step_task(dev.ptr + k * per)
        g.sync()
        o.step(ts[k])
        cause_o = o.falls(torso, beached, flip, void)
        cause_g = g.task_causes()
        assert np.array_equal(cause_g, cause_o), f"step {k + 1}"
        fell = np.flatnonzero(cause_o)
        total += len(fell)
        if len(fell):
            episode[fell] += 1
            seeds = (fell + n * episode[fell]).astype(np.uint32)
            o.reset(fell.astype(np.int32), seeds, diff[fell])
        a = np.stack([g.state(i) for i in range(n)])
        b = np.stack([o.state(i) for i in range(n)])
        assert_same(a, b, f"{form}: state after step {k + 1}")
    assert total > 10, "the tightened rules should terminate many worlds"
    assert np.array_equal(g.task_episodes(), episode.astype(np.uint32))
    stats = g.task_stats()
    assert stats[1:].sum() == total
