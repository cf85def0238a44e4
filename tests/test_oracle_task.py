"""Pins of the task-layer oracle (oracle/pytask.py), on the CPU, by what the
definitions fix (tests/test_task_golden.py pins it further against vectors of
the reference's own task module): closed-form gaze angles, the
observation's layout and ranges on freshly reset worlds (upright torso,
zero velocities, joint noise of at most 0.1 rad, unit heading vector),
checkpoint geometry, the reward of a resting world, the termination
precedence, and the curriculum's rounding.  The GPU test then holds the
device task to it step by step."""

import numpy as np
import pytest

import pyoracle as P
from pytask import OracleTask, gaze_error


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


OFFSETS = np.array([(f, l) for f in np.arange(-1.0, 1.41, 0.2) for l in np.arange(-1.0, 1.01, 0.2)])


def test_height_offsets_are_the_13_by_11_grid():
    assert OFFSETS.shape == (143, 2)
    assert np.allclose(np.unique(OFFSETS[:, 0]), np.linspace(-1.0, 1.4, 13))
    assert np.allclose(np.unique(OFFSETS[:, 1]), np.linspace(-1.0, 1.0, 11))


def test_gaze_closed_forms():
    rot = np.repeat(np.eye(3)[None], 4, 0)
    pos = np.zeros((4, 3))
    tg = np.array([[2.0, 0.0], [0.0, 3.0], [-1.0, 0.0], [1.0, -1.0]])
    assert np.allclose(gaze_error(rot, pos, tg), [0.0, np.pi / 2, np.pi, np.pi / 4])
    c, s = np.cos(0.3), np.sin(0.3)
    yaw = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])[None]
    assert np.allclose(gaze_error(yaw, pos[:1], np.array([[c, s]])), 0.0)


def _task(n, cfg=Cfg):
    o = OracleTask(P.Lib("o64").sim(n, cfg.dt, cfg.substeps, cfg.grid, cfg.cell), cfg, OFFSETS, cfg.grid, cfg.cell)
    ids = np.arange(n)
    pos = np.zeros((n, 2))
    ang = np.linspace(-3.0, 3.0, n)
    tg = pos + 4.5 * np.stack([np.cos(ang), np.sin(ang)], 1)
    o.reset(ids, ids.astype(np.uint32), tg)
    return o, tg


def test_observation_layout_after_reset():
    o, tg = _task(16)
    obs = o.observe()
    assert obs.shape == (16, 195) and obs.dtype == np.float32
    assert np.allclose(obs[:, 0:3], [0.0, 0.0, -1.0], atol=1e-6)       # gravity in the body frame, upright
    assert np.all(obs[:, 3:9] == 0.0)                                   # velocities zero after a reset
    assert np.all(np.abs(obs[:, 9:21]) <= 0.1 + 1e-6)                   # joint noise of the reset
    assert np.all(obs[:, 21:33] == 0.0) and np.all(obs[:, 33:45] == 0.0)  # rates, previous action
    assert np.allclose(np.hypot(obs[:, 45], obs[:, 46]), 1.0, atol=1e-6)  # unit direction
    assert np.allclose(obs[:, 47], 4.5 / 6.0, atol=1e-6)                # distance / 6 (spawn at the origin)
    assert np.all(np.abs(obs[:, 48:191]) <= 1.0) and np.all(obs[:, 191:195] == 0.0)
    # the checkpoint's bookkeeping: distance, gaze, view
    assert np.allclose(o.prev_dist, 4.5)
    assert np.allclose(o.prev_gaze, gaze_error(o.rot, o.root, tg))
    assert np.array_equal(o.in_view, o.prev_gaze < np.radians(45.0))


def test_resting_reward_and_no_termination():
    """Flat ground, actions holding the default pose: the robot settles; the
    reward is the time cost plus small drive / body-rate terms, never a fall."""
    class Flat(Cfg):
        pass
    n = 8
    o = OracleTask(P.Lib("o64").sim(n), Flat, OFFSETS, 64, 0.4)
    ids = np.arange(n)
    o.level[:] = 0.0
    o.reset(ids, ids.astype(np.uint32), np.full((n, 2), 4.0))
    for _ in range(20):
        r, term, trunc, stats, hit = o.step(np.zeros((n, 12)), lambda h: np.zeros((len(h), 2)))
    assert not term.any() and not trunc.any() and len(hit) == 0
    assert np.all(r < 0.0) and np.all(r > -0.2)
    assert np.all(stats["cause"] == 0) and np.all(stats["steps"] == 20)


def test_termination_precedence_and_curriculum_rounding():
    """Causes: bad > void > flip > beached, and the curriculum's np.round(., 6)."""
    bad = np.array([1, 0, 0, 0, 1, 0], bool)
    void = np.array([1, 1, 0, 0, 0, 0], bool)
    flip = np.array([1, 1, 1, 0, 1, 0], bool)
    beached = np.array([1, 1, 1, 1, 0, 0], bool)
    cause = np.select([bad, void, flip, beached], [1, 2, 3, 4], 0)
    assert list(cause) == [1, 2, 3, 4, 1, 0]
    lvl = np.clip(np.round(np.array([0.1 + 0.1 * 1.0, 0.7 + 0.1 * 1.0, 0.0 - 0.1, 0.9 + 0.1 + 0.1]), 6), 0.0, 1.0)
    assert list(lvl) == [0.2, 0.8, 0.0, 1.0]
    # the device's rounding: rint(x * 1e6) / 1e6 is numpy's algorithm for decimals > 0
    x = np.array([0.30000000000000004, 0.7999999999999999, 0.15000049999999999, 0.25])
    assert np.array_equal(np.round(x, 6), np.rint(x * 1e6) / 1e6)
