"""The SB3-style vectorised env over the device task (arc/vecenv.py) against
the reference's ArcVecEnv restated over the oracle (ArcVecEnv.step_wait,
arc/env.py:99-117, on oracle/pytask.py with numpy's generator): observations,
rewards, dones and the infos of ended worlds (terminal observation, time-limit
flag, checkpoints, difficulty, fell, cause), seeds and checkpoints from the
same np.random.default_rng stream, re-seeding through seed() + reset().
Integers, flags and causes exact; float32 outputs within 1e-5 relative
(floored at the smallest normal binary32)."""

import numpy as np
import pytest

import pyoracle as P
from helpers import action_stream, rel_close
from pytask import OracleTask

pytestmark = pytest.mark.gpu

CAUSES = ("", "unstable", "void", "flip", "beached")


def _close(a, b):
    """pure relative 1e-5, floored at the smallest normal binary32"""
    return bool(rel_close(a, b, 1e-5).all())


class OracleVecEnv:
    """arc/env.py:86-117 over the oracle task"""

    def __init__(self, n, cfg, offsets, seed):
        self.n = n
        self.task = OracleTask(P.Lib("o32").sim(n, cfg.dt, cfg.substeps, cfg.grid, cfg.cell), cfg, offsets,
                               cfg.grid, cfg.cell, rng=np.random.default_rng(seed))

    def reset(self, seed=None):
        if seed is not None:
            self.task.rng = np.random.default_rng(seed)
        self.task.reset(np.arange(self.n))
        return self.task.observe()

    def step(self, actions):
        reward, terminated, truncated, stats, _ = self.task.step(actions)
        obs = self.task.observe()
        done = terminated | truncated
        infos = [{} for _ in range(self.n)]
        ended = np.flatnonzero(done)
        for i in ended:
            infos[i] = {"terminal_observation": obs[i].copy(), "TimeLimit.truncated": bool(truncated[i]),
                        "checkpoints": int(stats["checkpoints"][i]), "difficulty": float(stats["difficulty"][i]),
                        "fell": bool(stats["fell"][i]), "cause": CAUSES[stats["cause"][i]]}
        if len(ended):
            self.task.reset(ended)
            obs[ended] = self.task.observe()[ended]
        return obs, reward, done, infos


def test_vecenv_matches_reference_env():
    import torch
    from arc.devtask import HEIGHT_OFFSETS, TaskConfig
    from arc.vecenv import DeviceVecEnv
    n, steps = 64, 36
    cfg = TaskConfig(max_steps=15, checkpoint_radius=3.0, fov=120.0, beached_steps=2, flip_up_z=0.95,
                     level_up_checkpoints=1, level_down_steps=10)
    env = DeviceVecEnv(n, cfg, seed=17, form="leg")
    ref = OracleVecEnv(n, cfg, HEIGHT_OFFSETS, 17)
    assert env.num_envs == n and env.observation_space.shape == (195,) and env.action_space.shape == (12,)
    assert _close(env.reset(), ref.reset())
    acts = action_stream(n, steps, seed=4)
    raw = ((np.asarray(acts) - ref.task.default_angles) / 0.6 * 1.3).astype(np.float32)
    ended_total = 0
    for k in range(steps):
        a = raw[k]
        # host actions, and every third step a device tensor read in place
        obs, rew, done, infos = env.step(torch.from_numpy(a).cuda() if k % 3 == 2 else a)
        robs, rrew, rdone, rinfos = ref.step(a)
        assert np.array_equal(done, rdone), f"step {k}"
        assert _close(rew, rrew) and _close(obs, robs), f"step {k}"
        for i in np.flatnonzero(rdone):
            d, r = infos[i], rinfos[i]
            assert set(d) == set(r)
            assert _close(d.pop("terminal_observation"), r.pop("terminal_observation")), f"step {k} world {i}"
            assert d == r, f"step {k} world {i}"
        assert all(not infos[i] for i in np.flatnonzero(~rdone))
        ended_total += int(rdone.sum())
    assert ended_total > 0
    # seed() then reset() restarts the generator (arc/env.py:93-95)
    env.seed(123)
    assert _close(env.reset(), ref.reset(seed=123))
    obs, rew, done, infos = env.step(raw[0])
    robs, rrew, rdone, _ = ref.step(raw[0])
    assert np.array_equal(done, rdone) and _close(obs, robs) and _close(rew, rrew)
    assert env.get_attr("num_envs") == [n] * n and env.env_is_wrapped(object, [0, 1]) == [False, False]
    env.close()
