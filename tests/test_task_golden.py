"""The task oracle (oracle/pytask.py) against vectors of the reference's own
Python task (tests/golden/task_vectors.npz, written by
tests/golden/make_task_vectors.py: arc/task.py of the reference over the
reference engine built from its sources, FP64, driven like arc/env.py's
step_wait, generator np.random.default_rng(seed)).

The oracle replays the same actions over the FP64 oracle engine (bit-exact
with that reference build) and must reproduce, bit for bit and step by step:
the seeds handed to the engine's resets, the generator's state (position in
the stream and 32-bit buffer), the checkpoints, rewards, terminations, causes,
checkpoint counts, episode steps, difficulties, curriculum levels and the
float32 observations.  This pins the restated draw order (checkpoint rounds,
reset seeds) and the restated task arithmetic on the reference's code."""

import hashlib
from pathlib import Path

import numpy as np
import pytest

import pyoracle as P
from pytask import OracleTask

GOLDEN = Path(__file__).resolve().parent / "golden" / "task_vectors.npz"
OFFSETS = np.array([(f, l) for f in np.arange(-1.0, 1.41, 0.2) for l in np.arange(-1.0, 1.01, 0.2)])


class Cfg:
    """arc/task.py:19-62 defaults"""
    dt, substeps, grid, cell = 1.0 / 30.0, 20, 64, 0.4
    max_steps, action_scale = 1000, 0.6
    checkpoint_radius, checkpoint_min_dist, checkpoint_max_dist, border = 0.5, 3.0, 6.0, 1.5
    w_progress, max_speed, fov, w_turn, w_checkpoint, w_time, w_fall = 10.0, 1.2, 45.0, 1.0, 10.0, 0.05, 20.0
    w_energy, w_action_rate, w_body_contact, w_torso_contact, w_body_rate, w_bounce = 0.002, 0.005, 0.02, 0.5, 0.01, 0.05
    flip_up_z, beached_steps, void_z = 0.3, 30, -2.0
    curriculum, difficulty, max_difficulty, level_step = True, 0.0, 1.0, 0.1
    level_up_checkpoints, level_down_steps = 3, 300


def _state(rng):
    st = rng.bit_generator.state
    s, inc = st["state"]["state"], st["state"]["inc"]
    m = (1 << 64) - 1
    return np.array([s >> 64, s & m, inc >> 64, inc & m, st["has_uint32"], st["uinteger"]], dtype=np.uint64)


def _sha1(obs):
    return np.frombuffer(hashlib.sha1(np.ascontiguousarray(obs, dtype=np.float32).tobytes()).digest(), np.uint8)


@pytest.mark.skipif(not P.available("o64"), reason="oracle not built (run build())")
def test_task_oracle_reproduces_reference_task():
    g = np.load(GOLDEN)
    cfg = type("C", (Cfg,), {})
    for k, v in zip(g["cfg_keys"], g["cfg_values"]):
        setattr(cfg, str(k), type(getattr(Cfg, str(k)))(v))
    n = int(g["n"])
    o = OracleTask(P.Lib("o64").sim(n, cfg.dt, cfg.substeps, cfg.grid, cfg.cell), cfg, OFFSETS, cfg.grid, cfg.cell,
                   rng=np.random.default_rng(int(g["seed"])))
    seeds = []
    engine_reset = o.sim.reset
    o.sim.reset = lambda ids, s, d: (seeds.append((np.array(ids), np.array(s, np.uint32))), engine_reset(ids, s, d))

    def check(step, obs):
        tgt = np.concatenate([o.target, o.target_z[:, None]], 1)
        assert np.array_equal(tgt, g["target"][step]), f"step {step}: checkpoints"
        assert np.array_equal(o.level, g["level"][step]), f"step {step}: levels"
        assert np.array_equal(_state(o.rng), g["rng"][step]), f"step {step}: generator state"
        assert np.array_equal(_sha1(obs), g["obs_sha1"][step]), f"step {step}: observations"
        if f"obs_{step}" in g.files:
            assert np.array_equal(obs.astype(np.float32), g[f"obs_{step}"])
        at = g["reset_steps"] == step
        got_ids = np.concatenate([i for i, _ in seeds]) if seeds else np.zeros(0, np.int64)
        got_seeds = np.concatenate([s for _, s in seeds]) if seeds else np.zeros(0, np.uint32)
        assert np.array_equal(got_ids, g["reset_ids"][at]) and np.array_equal(got_seeds, g["reset_seeds"][at]), \
            f"step {step}: reset seeds"
        seeds.clear()

    o.reset(np.arange(n))
    check(0, o.observe())
    for k in range(len(g["actions"])):
        reward, term, trunc, stats, _ = o.step(g["actions"][k])
        obs = o.observe()
        ended = np.flatnonzero(term | trunc)
        if len(ended):
            o.reset(ended)
            obs[ended] = o.observe()[ended]
        assert np.array_equal(reward, g["reward"][k]), f"step {k + 1}: rewards"
        assert np.array_equal(term, g["terminated"][k]) and np.array_equal(trunc, g["truncated"][k])
        for key in ("cause", "checkpoints", "steps", "difficulty"):
            assert np.array_equal(np.asarray(stats[key]), g[key][k]), f"step {k + 1}: {key}"
        check(k + 1, obs)
    # the vectors exercise every path the draws take
    assert g["checkpoints"].max() > 0 and g["truncated"].any() and g["terminated"].any()
    assert len(np.unique(g["level"])) > 1
