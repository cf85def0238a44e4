"""Oracle of the task layer: a numpy restatement of the reference's
checkpoint task (arc/task.py:90-280) and of its vectorised env's auto-reset
order (arc/env.py:99-117), over the oracle sim (pyoracle.Sim).

Test infrastructure: only tests/ use it, as the checker of the device task
(include/arc_task.h).  The reference draws reset seeds and checkpoint
positions from one numpy generator (arc/task.py:93,117,136-137).  Given a
numpy Generator (`rng`), this oracle draws them as the reference does, for the
device's numpy mode; without one it takes them from the caller (the device's
per-world streams, read back by the test).  It restates everything else:
observations, rewards, terminations and their precedence, the checkpoint
rule's consequences (terrain height, distance, gaze, view), the curriculum.
Pinned bit for bit, draws included, against the reference's own task module
run over the reference engine (tests/test_task_golden.py,
tests/golden/task_vectors.npz).
"""

from __future__ import annotations

import numpy as np

NJ, LEGS = 12, 4


def gaze_error(rot: np.ndarray, pos: np.ndarray, target: np.ndarray) -> np.ndarray:
    """arc/task.py:80-87"""
    fwd = rot[:, :2, 0]
    d = target - pos[:, :2]
    cross = fwd[:, 0] * d[:, 1] - fwd[:, 1] * d[:, 0]
    return np.abs(np.arctan2(cross, np.sum(fwd * d, axis=1)))


class OracleTask:
    def __init__(self, sim, cfg, offsets: np.ndarray, grid: int, cell: float, rng=None):
        self.sim, self.cfg, self.offsets, self.rng = sim, cfg, offsets, rng
        n = sim.n
        self.n = n
        d, _, _, misc = sim.robot_info()
        self.default_angles, self.stand_height = d, float(misc[0])
        self.half_extent = 0.5 * min(grid, 64) * cell
        self.target = np.zeros((n, 2))
        self.target_z = np.zeros(n)
        self.prev_dist = np.zeros(n)
        self.prev_gaze = np.zeros(n)
        self.in_view = np.zeros(n, dtype=bool)
        self.prev_action = np.zeros((n, NJ))
        self.torso_steps = np.zeros(n, dtype=np.int64)
        self.steps = np.zeros(n, dtype=np.int64)
        self.checkpoints = np.zeros(n, dtype=np.int64)
        self.level = np.full(n, 0.0 if cfg.curriculum else cfg.difficulty)
        self.difficulty = self.level.copy()
        self._observe()

    def _observe(self):
        self.root, self.rot, self.joints, self.contacts, self.flags = self.sim.observe()

    def _set_targets(self, ids, target):
        """arc/task.py:145-150 with the placed checkpoints given"""
        pos = self.root[ids, :2]
        self.target[ids] = target
        self.target_z[ids] = self.sim.terrain_at(ids, target)
        self.prev_dist[ids] = np.linalg.norm(target - pos, axis=1)
        self.prev_gaze[ids] = gaze_error(self.rot[ids], pos, target)
        self.in_view[ids] = self.prev_gaze[ids] < np.radians(self.cfg.fov)

    def _draw_targets(self, ids):
        """arc/task.py:127-144 on self.rng: rounds over the pending worlds, each
        drawing all angles, then all distances; inside-the-border candidates
        are taken; after 20 rounds the rest go halfway to the centre"""
        cfg = self.cfg
        pos = self.root[ids, :2]
        limit = self.half_extent - cfg.border
        out = 0.5 * pos
        pending = np.arange(len(ids))
        for _ in range(20):
            if len(pending) == 0:
                break
            ang = self.rng.uniform(-np.pi, np.pi, len(pending))
            dist = self.rng.uniform(cfg.checkpoint_min_dist, cfg.checkpoint_max_dist, len(pending))
            cand = pos[pending] + dist[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1)
            inside = np.all(np.abs(cand) < limit, axis=1)
            out[pending[inside]] = cand[inside]
            pending = pending[~inside]
        return out

    def reset(self, ids, seeds=None, targets=None):
        """arc/task.py:108-125; seeds and first checkpoints given, or drawn
        from self.rng (seeds first, then the checkpoints after the reset)"""
        ids = np.asarray(ids, dtype=np.int64)
        if not len(ids):
            return
        if seeds is None:
            seeds = self.rng.integers(0, 2**32, size=len(ids), dtype=np.uint32)
        self.last_seeds = seeds
        self.difficulty[ids] = self.level[ids]
        self.sim.reset(ids, seeds, self.difficulty[ids])
        self.prev_action[ids] = 0.0
        self.torso_steps[ids] = 0
        self.steps[ids] = 0
        self.checkpoints[ids] = 0
        self._observe()
        self._set_targets(ids, self._draw_targets(ids) if targets is None else targets)

    def step(self, actions, new_targets=None):
        """arc/task.py:155-239; new_targets(ids) gives the checkpoints placed
        for the worlds that reached theirs (None: drawn from self.rng).
        Returns reward, terminated,
        truncated, stats, and the worlds that got a new checkpoint."""
        cfg, sim = self.cfg, self.sim
        a = np.clip(np.asarray(actions, dtype=np.float64), -1.0, 1.0)
        sim.step(self.default_angles + cfg.action_scale * a)
        self._observe()
        root, joints, contacts, flags, rot = self.root, self.joints, self.contacts, self.flags, self.rot
        nj, nl = NJ, LEGS
        self.steps += 1
        pos = root[:, :3]
        to_target = self.target - pos[:, :2]
        dist = np.hypot(to_target[:, 0], to_target[:, 1])
        gaze = gaze_error(rot, pos, self.target)
        cos_fov = np.cos(np.radians(cfg.fov))
        view = np.clip((np.cos(gaze) - cos_fov) / (1.0 - cos_fov), 0.0, 1.0)
        self.in_view = gaze < np.radians(cfg.fov)
        progress = np.minimum(self.prev_dist - dist, cfg.max_speed * cfg.dt)
        progress = np.where(progress > 0, progress * view, progress)
        turn = self.prev_gaze - gaze
        reached = (dist < cfg.checkpoint_radius) & self.in_view
        body_rate = np.matmul(root[:, None, 10:13], rot)[:, 0, :2]
        torso = contacts[:, nl] > 0
        self.torso_steps = np.where(torso, self.torso_steps + 1, 0)
        power = np.abs(joints[:, nj:2 * nj] * joints[:, 2 * nj:]).sum(1)
        rate = np.square(a - self.prev_action).sum(1)
        reward = (cfg.w_progress * progress + cfg.w_turn * turn + cfg.w_checkpoint * reached - cfg.w_time
                  - cfg.w_energy * power * cfg.dt - cfg.w_action_rate * rate
                  - cfg.w_body_contact * (contacts[:, nl + 1] > 0) - cfg.w_torso_contact * torso
                  - cfg.w_body_rate * np.square(body_rate).sum(1) - cfg.w_bounce * np.square(root[:, 9]))
        bad = (flags != 0) | ~np.all(np.isfinite(root), axis=1)
        void = np.any(np.abs(pos[:, :2]) > self.half_extent, axis=1) | (pos[:, 2] < cfg.void_z)
        flip = rot[:, 2, 2] < cfg.flip_up_z
        beached = self.torso_steps > cfg.beached_steps
        fell = flip | beached | void | bad
        cause = np.select([bad, void, flip, beached], [1, 2, 3, 4], 0)
        reward = np.where(fell, reward - cfg.w_fall, reward)
        reward = np.where(bad, -cfg.w_fall, reward)
        self.prev_dist = dist
        self.prev_gaze = gaze
        self.prev_action = a
        self.checkpoints += reached
        hit = np.flatnonzero(reached & ~fell)
        if len(hit):
            self._set_targets(hit, self._draw_targets(hit) if new_targets is None else new_targets(hit))
        terminated = fell
        truncated = (self.steps >= cfg.max_steps) & ~terminated
        stats = {"checkpoints": self.checkpoints.copy(), "difficulty": self.difficulty.copy(),
                 "fell": fell.copy(), "cause": cause, "steps": self.steps.copy()}
        done = np.flatnonzero(terminated | truncated)
        if len(done) and cfg.curriculum:
            up = self.checkpoints[done] >= cfg.level_up_checkpoints
            down = fell[done] & (self.checkpoints[done] == 0) & (self.steps[done] < cfg.level_down_steps)
            lvl = self.level[done] + cfg.level_step * (up.astype(float) - down.astype(float))
            self.level[done] = np.clip(np.round(lvl, 6), 0.0, cfg.max_difficulty)
        return reward.astype(np.float32), terminated, truncated, stats, hit

    def observe(self) -> np.ndarray:
        """arc/task.py:250-280"""
        sim, nj, nl = self.sim, NJ, LEGS
        root, joints, contacts, rot = self.root, self.joints, self.contacts, self.rot
        gravity = -rot[:, 2, :]
        lin = np.matmul(root[:, None, 7:10], rot)[:, 0]
        ang = np.matmul(root[:, None, 10:13], rot)[:, 0]
        heading = np.arctan2(rot[:, 1, 0], rot[:, 0, 0])
        d = self.target - root[:, :2]
        c, s = np.cos(heading), np.sin(heading)
        fwd, left = c * d[:, 0] + s * d[:, 1], -s * d[:, 0] + c * d[:, 1]
        dist = np.maximum(np.hypot(fwd, left), 1e-6)
        target = np.stack([fwd / dist, left / dist, np.minimum(dist, 6.0) / 6.0], 1)
        heights = sim.heights(self.offsets)
        rel_heights = np.clip(root[:, 2:3] - self.stand_height - heights, -1.0, 1.0)
        obs = np.concatenate([gravity, lin, 0.25 * ang, joints[:, :nj] - self.default_angles,
                              0.1 * joints[:, nj:2 * nj], self.prev_action, target, rel_heights,
                              contacts[:, :nl]], axis=1, dtype=np.float32)
        obs[~np.isfinite(obs)] = 0.0
        return obs
