"""The checkpoint-reaching task on the device (include/arc_task.h).

`DeviceTask` mirrors the reference's `ArcTask` + `ArcVecEnv.step_wait`
(arc/task.py:90-280, arc/env.py:93-117): `reset()` returns the first
observations, `step(actions)` returns (obs, reward, done, infos-arrays) with
the ended worlds already reset and re-observed, their last observations in
`terminal_obs`.  Everything stays on the GPU: the buffers are exported with
the DLPack protocol (`torch.from_dlpack(task.obs)`), and `step` takes any
device pointer or `__dlpack__`-capable float32 [N, 12] array of actions, so a
policy on the same GPU exchanges them without host copies.

Import torch before this module (torch's HIP runtime is then the one
libarcsim binds to; see arc.native.DeviceBuffer).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, fields

import numpy as np

from .native import Sim, sim_lib

_vp = ctypes.c_void_p
_int = ctypes.c_int

CAUSES = ("", "unstable", "void", "flip", "beached")

# the reference's height sample grid around the torso, (forward, left) in m
# (arc/task.py:73-76): 13 x 11 offsets, forward outer
_F = np.arange(-1.0, 1.41, 0.2)
_L = np.arange(-1.0, 1.01, 0.2)
HEIGHT_OFFSETS = np.array([(f, l) for f in _F for l in _L])


@dataclass
class TaskConfig:
    """arc/task.py:19-62 (field names and defaults kept)"""
    dt: float = 1.0 / 30.0
    substeps: int = 20
    grid: int = 64
    cell: float = 0.4
    max_steps: int = 1000
    action_scale: float = 0.6
    checkpoint_radius: float = 0.5
    checkpoint_min_dist: float = 3.0
    checkpoint_max_dist: float = 6.0
    border: float = 1.5
    w_progress: float = 10.0
    max_speed: float = 1.2
    fov: float = 45.0
    w_turn: float = 1.0
    w_checkpoint: float = 10.0
    w_time: float = 0.05
    w_fall: float = 20.0
    w_energy: float = 0.002
    w_action_rate: float = 0.005
    w_body_contact: float = 0.02
    w_torso_contact: float = 0.5
    w_body_rate: float = 0.01
    w_bounce: float = 0.05
    flip_up_z: float = 0.3
    beached_steps: int = 30
    void_z: float = -2.0
    curriculum: bool = True
    difficulty: float = 0.0
    max_difficulty: float = 1.0
    level_step: float = 0.1
    level_up_checkpoints: int = 3
    level_down_steps: int = 300


class _Config(ctypes.Structure):
    _fields_ = [("max_steps", _int), ("action_scale", ctypes.c_double),
                ("checkpoint_radius", ctypes.c_double), ("checkpoint_min_dist", ctypes.c_double),
                ("checkpoint_max_dist", ctypes.c_double), ("border", ctypes.c_double),
                ("w_progress", ctypes.c_double), ("max_speed", ctypes.c_double), ("fov_deg", ctypes.c_double),
                ("w_turn", ctypes.c_double), ("w_checkpoint", ctypes.c_double), ("w_time", ctypes.c_double),
                ("w_fall", ctypes.c_double), ("w_energy", ctypes.c_double), ("w_action_rate", ctypes.c_double),
                ("w_body_contact", ctypes.c_double), ("w_torso_contact", ctypes.c_double),
                ("w_body_rate", ctypes.c_double), ("w_bounce", ctypes.c_double), ("flip_up_z", ctypes.c_double),
                ("void_z", ctypes.c_double), ("beached_steps", _int), ("curriculum", _int),
                ("difficulty", ctypes.c_double), ("max_difficulty", ctypes.c_double),
                ("level_step", ctypes.c_double), ("level_up_checkpoints", _int), ("level_down_steps", _int)]


def _task_lib():
    lib = sim_lib()
    if not hasattr(lib, "_task_declared"):
        for name, res, args in [
            ("arct_default_config", None, [_vp]),
            ("arct_create", _vp, [_vp, _vp, _vp, _int, ctypes.c_uint]),
            ("arct_destroy", None, [_vp]),
            ("arct_obs_size", _int, [_vp]),
            ("arct_buffer", _vp, [_vp, _int]),
            ("arct_reset_all", _int, [_vp]),
            ("arct_step", _int, [_vp, _vp, _int]),
            ("arct_sync", _int, [_vp]),
            ("arct_targets", _int, [_vp, _vp]),
            ("arct_levels", _int, [_vp, _vp]),
            ("arct_set_levels", _int, [_vp, _vp]),
            ("arct_last_resets", _int, [_vp, _vp, _vp]),
            ("arct_seed_numpy", _int, [_vp, _vp, _int]),
            ("arct_wait_stream", _int, [_vp, _vp]),
            ("arct_signal_stream", _int, [_vp, _vp]),
            ("arct_numpy_state", _int, [_vp, _vp]),
        ]:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib._task_declared = True
    return lib


# -- DLPack export of the task's device buffers -----------------------------

class _DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class _DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", _vp), ("device", _DLDevice), ("ndim", ctypes.c_int32), ("dtype", _DLDataType),
                ("shape", ctypes.POINTER(ctypes.c_int64)), ("strides", ctypes.POINTER(ctypes.c_int64)),
                ("byte_offset", ctypes.c_uint64)]


class _DLManagedTensor(ctypes.Structure):
    pass


_DELETER = ctypes.CFUNCTYPE(None, ctypes.POINTER(_DLManagedTensor))
_DLManagedTensor._fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", _vp), ("deleter", _DELETER)]

_KDL_ROCM = 10
_DTYPES = {"f32": (2, 32), "f64": (2, 64), "u8": (1, 8), "i32": (0, 32)}
_LIVE: dict[int, tuple] = {}  # exported tensors kept alive until their deleter runs

_PyCapsule_New = ctypes.pythonapi.PyCapsule_New
_PyCapsule_New.restype = ctypes.py_object
_PyCapsule_New.argtypes = [_vp, ctypes.c_char_p, _vp]


@_DELETER
def _dl_deleter(mt):
    _LIVE.pop(ctypes.addressof(mt.contents), None)


class DeviceArray:
    """A view of device memory the task owns, exported through DLPack."""

    def __init__(self, owner, ptr: int, shape: tuple, dtype: str, device_id: int = 0):
        self._owner, self.ptr, self.shape, self.dtype, self.device_id = owner, ptr, tuple(shape), dtype, device_id

    def __dlpack_device__(self):
        return (_KDL_ROCM, self.device_id)

    def __dlpack__(self, stream=None, **kwargs):
        self._owner.sync()
        shape = (ctypes.c_int64 * len(self.shape))(*self.shape)
        mt = _DLManagedTensor()
        code, bits = _DTYPES[self.dtype]
        mt.dl_tensor = _DLTensor(self.ptr, _DLDevice(_KDL_ROCM, self.device_id), len(self.shape),
                                 _DLDataType(code, bits, 1), shape, None, 0)
        mt.manager_ctx = None
        mt.deleter = _dl_deleter
        addr = ctypes.addressof(mt)
        _LIVE[addr] = (mt, shape, self._owner)
        # the consumer takes the tensor and calls its deleter when done with it
        return _PyCapsule_New(addr, b"dltensor", None)

    def numpy(self) -> np.ndarray:
        """a host copy"""
        self._owner.sync()
        dt = {"f32": np.float32, "f64": np.float64, "u8": np.uint8, "i32": np.int32}[self.dtype]
        out = np.empty(self.shape, dtype=dt)
        if sim_lib().arcx_memcpy(out.ctypes.data, self.ptr, out.nbytes, 0):
            raise RuntimeError("device to host copy failed")
        return out


def _pointer(actions) -> int:
    if actions is None:
        return 0
    if isinstance(actions, int):
        return actions
    if hasattr(actions, "data_ptr"):  # a torch tensor
        if not actions.is_contiguous() or str(actions.dtype) != "torch.float32":
            raise ValueError("actions must be a contiguous float32 tensor")
        return actions.data_ptr()
    if isinstance(actions, DeviceArray):
        return actions.ptr
    raise TypeError("actions: a device pointer, a torch tensor or a DeviceArray")


def seed_words(seed: int | None) -> np.ndarray:
    """numpy's SeedSequence entropy of `seed` as little-endian 32-bit words
    (None: fresh OS entropy, as np.random.default_rng(None))"""
    if seed is None:
        seed = np.random.SeedSequence().entropy
    seed = int(seed)
    if seed < 0:
        raise ValueError("seed must be non-negative")
    w = []
    while True:
        w.append(seed & 0xFFFFFFFF)
        seed >>= 32
        if not seed:
            return np.array(w, dtype=np.uint32)


class DeviceTask:
    """rng="world": checkpoints and reset seeds from per-world streams seeded by
    `seed`; rng="numpy": from np.random.default_rng(seed) in the reference's
    order (the Python task's draws for the same seed)."""

    def __init__(self, num_worlds: int, config: TaskConfig | None = None, seed: int | None = 0,
                 form: str | None = None, rng: str = "world"):
        self.cfg = cfg = config or TaskConfig()
        self._lib = _task_lib()
        self.sim = Sim(num_worlds, cfg.dt, cfg.substeps, cfg.grid, cfg.cell)
        if form:
            self.sim.set_form(form)
        c = _Config()
        for f in fields(cfg):
            name = "fov_deg" if f.name == "fov" else f.name
            if name in dict(_Config._fields_):
                setattr(c, name, int(getattr(cfg, f.name)) if name in ("curriculum",) else getattr(cfg, f.name))
        off = np.ascontiguousarray(HEIGHT_OFFSETS, dtype=np.float64)
        if rng not in ("world", "numpy"):
            raise ValueError(f"rng: 'world' or 'numpy', not {rng!r}")
        self._ptr = self._lib.arct_create(self.sim.ptr, ctypes.byref(c), off.ctypes.data, len(off),
                                          (seed or 0) & 0xFFFFFFFF)
        if not self._ptr:
            raise RuntimeError("arct_create failed")
        if rng == "numpy":
            self.seed_numpy(seed)
        n, self.num_worlds, self.nj = num_worlds, num_worlds, 12
        self.obs_size = self._lib.arct_obs_size(self._ptr)
        buf = lambda k: self._lib.arct_buffer(self._ptr, k)
        self.obs = DeviceArray(self, buf(0), (n, self.obs_size), "f32")
        self.terminal_obs = DeviceArray(self, buf(1), (n, self.obs_size), "f32")
        self.reward = DeviceArray(self, buf(2), (n,), "f32")
        self.terminated = DeviceArray(self, buf(3), (n,), "u8")
        self.truncated = DeviceArray(self, buf(4), (n,), "u8")
        self.info = DeviceArray(self, buf(5), (3, n), "i32")  # cause, checkpoints, steps
        self.difficulty = DeviceArray(self, buf(6), (n,), "f64")
        self.actions = DeviceArray(self, buf(7), (n, 12), "f32")

    def seed_numpy(self, seed: int | None) -> None:
        """from now on draw from np.random.default_rng(seed), as the reference's
        ArcTask.rng (arc/task.py:93, arc/env.py:93-95)"""
        w = seed_words(seed)
        if self._lib.arct_seed_numpy(self._ptr, w.ctypes.data, len(w)):
            self.sim._check()

    def numpy_state(self) -> dict:
        """the generator's state in the layout of numpy's bit_generator.state"""
        out = np.zeros(6, dtype=np.uint64)
        if self._lib.arct_numpy_state(self._ptr, out.ctypes.data):
            self.sim._check()
        return {"bit_generator": "PCG64",
                "state": {"state": int(out[0]) << 64 | int(out[1]), "inc": int(out[2]) << 64 | int(out[3])},
                "has_uint32": int(out[4]), "uinteger": int(out[5])}

    @staticmethod
    def _stream(stream) -> int:
        return int(getattr(stream, "cuda_stream", stream))

    def wait_stream(self, stream) -> None:
        """the task's next work waits, on the device, for the work queued so far
        on `stream` (a torch.cuda.Stream or a HIP stream handle)"""
        if self._lib.arct_wait_stream(self._ptr, self._stream(stream)):
            self.sim._check()

    def signal_stream(self, stream) -> None:
        """work queued on `stream` from now on waits, on the device, for the
        task's queued work"""
        if self._lib.arct_signal_stream(self._ptr, self._stream(stream)):
            self.sim._check()

    def sync(self) -> None:
        self._lib.arct_sync(self._ptr)
        self.sim._check()

    def reset(self) -> DeviceArray:
        if self._lib.arct_reset_all(self._ptr):
            self.sim._check()
        return self.obs

    def actions_upload(self, a: np.ndarray) -> None:
        """host actions (N, 12) into the task's own actions buffer"""
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.shape == (self.num_worlds, 12)
        if sim_lib().arcx_memcpy(self.actions.ptr, a.ctypes.data, a.nbytes, 1):
            raise RuntimeError("host to device copy failed")

    def step(self, actions=None, graph: bool = True) -> DeviceArray:
        """one env step; returns the observations (device), the other outputs
        are in reward / terminated / truncated / info / terminal_obs"""
        if self._lib.arct_step(self._ptr, _pointer(actions), 1 if graph else 0):
            self.sim._check()
        return self.obs

    def targets(self) -> np.ndarray:
        out = np.zeros((self.num_worlds, 3))
        self._lib.arct_targets(self._ptr, out.ctypes.data)
        return out

    def set_levels(self, levels: np.ndarray) -> None:
        lv = np.ascontiguousarray(levels, dtype=np.float64)
        assert lv.shape == (self.num_worlds,)
        if self._lib.arct_set_levels(self._ptr, lv.ctypes.data):
            self.sim._check()

    def levels(self) -> np.ndarray:
        out = np.zeros(self.num_worlds)
        self._lib.arct_levels(self._ptr, out.ctypes.data)
        return out

    def last_resets(self) -> tuple[np.ndarray, np.ndarray]:
        ids = np.zeros(self.num_worlds, dtype=np.int32)
        seeds = np.zeros(self.num_worlds, dtype=np.uint32)
        k = self._lib.arct_last_resets(self._ptr, ids.ctypes.data, seeds.ctypes.data)
        order = np.argsort(ids[:k], kind="stable")
        return ids[:k][order], seeds[:k][order]

    def close(self) -> None:
        if getattr(self, "_ptr", None):
            self._lib.arct_destroy(self._ptr)
            self._ptr = None
        self.sim.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
