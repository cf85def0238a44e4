"""A Stable-Baselines3-style vectorised env over the device task.

`DeviceVecEnv` has the surface of the reference's `ArcVecEnv`
(arc/env.py:70-140): `reset()` -> observations, `step_async(actions)` /
`step_wait()` -> (obs, reward, done, infos) with SB3's auto-reset convention
(the ended worlds' last observations in `infos[i]["terminal_observation"]`,
`"TimeLimit.truncated"`, and the task's episode statistics), `seed`,
`get_attr` / `set_attr` / `env_method` / `env_is_wrapped`, `close`.  The task
runs on the GPU (`arc.devtask.DeviceTask`); this class only copies its outputs
to the host in SB3's layout.  With rng="numpy" (the default) the env draws
checkpoints and reset seeds from np.random.default_rng(seed) in the
reference's order, so an env built with the same seed reproduces the Python
env's seeds and checkpoints.

When stable_baselines3 is importable the class derives from its VecEnv, and
the spaces are gymnasium Boxes; otherwise it stands alone with Box-like
spaces of the same shape, bounds and dtype.  Rendering is not part of the
device path: `render` and `get_images` return None.
"""

from __future__ import annotations

from typing import Any, Iterable

import numpy as np

from .devtask import CAUSES, DeviceTask, TaskConfig

try:  # pragma: no cover - absent in this image
    from stable_baselines3.common.vec_env.base_vec_env import VecEnv as _VecEnvBase
except ImportError:  # stands alone
    _VecEnvBase = None

try:  # pragma: no cover - absent in this image
    from gymnasium import spaces as _spaces
except ImportError:
    _spaces = None


class Box:
    """The fields of gymnasium.spaces.Box that callers of a VecEnv read."""

    def __init__(self, low: float, high: float, shape: tuple, dtype=np.float32):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.low = np.full(self.shape, low, dtype=self.dtype)
        self.high = np.full(self.shape, high, dtype=self.dtype)

    def contains(self, x) -> bool:
        x = np.asarray(x)
        return x.shape == self.shape and bool(np.all((x >= self.low) & (x <= self.high)))

    def __repr__(self) -> str:
        return f"Box({self.low.flat[0]}, {self.high.flat[0]}, {self.shape}, {self.dtype})"


def _box(low, high, shape):
    if _spaces is not None:
        return _spaces.Box(low, high, shape=shape, dtype=np.float32)
    return Box(low, high, shape)


class _StandaloneVecEnv:
    """SB3 VecEnv's bookkeeping when stable_baselines3 is absent: sizes,
    spaces, per-env seeds and options consumed by the next reset."""

    def __init__(self, num_envs: int, observation_space, action_space):
        self.num_envs = num_envs
        self.observation_space = observation_space
        self.action_space = action_space
        self.reset_infos: list[dict[str, Any]] = [{} for _ in range(num_envs)]
        self._seeds: list[int | None] = [None] * num_envs
        self._options: list[dict[str, Any]] = [{} for _ in range(num_envs)]

    def step(self, actions):
        self.step_async(actions)
        return self.step_wait()

    def seed(self, seed: int | None = None):
        if seed is None:
            seed = int(np.random.randint(0, 2**32 - 1, dtype=np.uint32))
        self._seeds = [seed + idx for idx in range(self.num_envs)]
        return self._seeds

    def set_options(self, options=None) -> None:
        if options is None:
            options = {}
        self._options = options if isinstance(options, list) else [options] * self.num_envs

    def _reset_seeds(self) -> None:
        self._seeds = [None] * self.num_envs

    def _reset_options(self) -> None:
        self._options = [{} for _ in range(self.num_envs)]

    def _get_indices(self, indices) -> Iterable[int]:
        if indices is None:
            return range(self.num_envs)
        if isinstance(indices, int):
            return [indices]
        return indices


_Base = _VecEnvBase if _VecEnvBase is not None else _StandaloneVecEnv


class DeviceVecEnv(_Base):
    """ArcVecEnv's interface (arc/env.py:70-140) over `DeviceTask`."""

    metadata = {"render_modes": []}

    def __init__(self, num_envs: int, config: TaskConfig | None = None, threads: int = 1, seed: int | None = None,
                 render_mode: str | None = None, form: str | None = None, rng: str = "numpy", graph: bool = True):
        # threads: accepted for the reference's signature; the GPU ignores it
        self.task = DeviceTask(num_envs, config, seed=seed, form=form, rng=rng)
        self._rng_mode, self._graph = rng, graph
        obs_space = _box(-np.inf, np.inf, (self.task.obs_size,))
        act_space = _box(-1.0, 1.0, (self.task.nj,))
        self.render_mode = render_mode
        super().__init__(num_envs, obs_space, act_space)
        self._actions: np.ndarray | None = None

    # -- the VecEnv protocol ---------------------------------------------------

    def reset(self) -> np.ndarray:
        """arc/env.py:93-98: a pending seed() re-seeds the task's generator"""
        if self._seeds[0] is not None:
            if self._rng_mode == "numpy":
                self.task.seed_numpy(self._seeds[0])
            else:
                raise ValueError("seed(): the per-world streams are seeded at construction; use rng='numpy'")
        self._reset_seeds()
        self.task.reset()
        self._reset_options()
        return self.task.obs.numpy()

    def step_async(self, actions) -> None:
        self._actions = actions

    def step_wait(self):
        """arc/env.py:99-117: step, observe, terminal observations and infos of
        the ended worlds, their reset and first observations (all on the GPU)"""
        a = self._actions
        if hasattr(a, "data_ptr") and getattr(a, "is_cuda", False):
            # a device tensor: read in place once its producer has finished
            import torch
            torch.cuda.current_stream(a.device).synchronize()
            self.task.step(a, graph=self._graph)
        else:
            if hasattr(a, "numpy"):  # a host tensor
                a = a.numpy()
            self.task.actions_upload(np.asarray(a, dtype=np.float32).reshape(self.num_envs, self.task.nj))
            self.task.step(None, graph=self._graph)
        t = self.task
        obs = t.obs.numpy()
        reward = t.reward.numpy()
        terminated = t.terminated.numpy().astype(bool)
        truncated = t.truncated.numpy().astype(bool)
        done = terminated | truncated
        infos: list[dict[str, Any]] = [{} for _ in range(self.num_envs)]
        ended = np.flatnonzero(done)
        if len(ended):
            term_obs = t.terminal_obs.numpy()
            info = t.info.numpy()
            diff = t.difficulty.numpy()
            for i in ended:
                infos[i] = {
                    "terminal_observation": term_obs[i].copy(),
                    "TimeLimit.truncated": bool(truncated[i]),
                    "checkpoints": int(info[1, i]),
                    "difficulty": float(diff[i]),
                    "fell": bool(terminated[i]),
                    "cause": CAUSES[info[0, i]],
                }
        return obs, reward, done, infos

    def close(self) -> None:
        self.task.close()

    def render(self, mode: str | None = None):
        return None

    def get_images(self):
        return [None]

    def get_attr(self, attr_name, indices=None):
        return [getattr(self, attr_name)] * len(list(self._get_indices(indices)))

    def set_attr(self, attr_name, value, indices=None) -> None:
        setattr(self, attr_name, value)

    def env_method(self, method_name, *method_args, indices=None, **method_kwargs):
        return [getattr(self, method_name)(*method_args, **method_kwargs) for _ in self._get_indices(indices)]

    def env_is_wrapped(self, wrapper_class, indices=None):
        return [False] * len(list(self._get_indices(indices)))
