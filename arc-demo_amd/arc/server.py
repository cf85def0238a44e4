"""ctypes bindings of the step server (include/arc_server.h).

One device-resident sim holds the worlds of many clients; clients are
admitted by HBM budget, their step requests are batched as they arrive and
each request's latency (submit to completion) is recorded."""

from __future__ import annotations

import ctypes
import functools

import numpy as np

from .native import sim_lib

_vp = ctypes.c_void_p
_int = ctypes.c_int
_dbl = ctypes.c_double
_ll = ctypes.c_longlong


class AdmissionError(RuntimeError):
    pass


@functools.cache
def server_lib() -> ctypes.CDLL:
    lib = sim_lib()
    for name, res, args in [
        ("arcs_create", _vp, [_dbl, _dbl, _int, _int, _dbl]),
        ("arcs_destroy", None, [_vp]),
        ("arcs_capacity", _int, [_vp]),
        ("arcs_used", _int, [_vp]),
        ("arcs_open", _int, [_vp, _int, ctypes.c_uint, _dbl]),
        ("arcs_close", _int, [_vp, _int]),
        ("arcs_client_range", _int, [_vp, _int, ctypes.POINTER(_int), ctypes.POINTER(_int)]),
        ("arcs_submit", _ll, [_vp, _int, _vp]),
        ("arcs_submit_device", _ll, [_vp, _int, _vp]),
        ("arcs_wait", _dbl, [_vp, _ll]),
        ("arcs_stats", _int, [_vp, _vp]),
        ("arcs_reset_stats", None, [_vp]),
        ("arcs_set_window", _int, [_vp, _dbl]),
        ("arcs_autotune", _int, [_vp, _int, _vp, ctypes.POINTER(_dbl)]),
        ("arcs_observe", _int, [_vp, _int, _vp]),
    ]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


STAT_KEYS = ("requests", "batches", "world_steps", "worlds_per_batch", "p50_ms", "p99_ms", "max_ms",
             "seconds", "world_steps_per_s", "requests_per_batch")


class StepServer:
    def __init__(self, hbm_budget_bytes: float, dt: float = 1.0 / 30.0, substeps: int = 20, grid: int = 64,
                 cell: float = 0.4):
        self._lib = server_lib()
        self._ptr = self._lib.arcs_create(float(hbm_budget_bytes), dt, substeps, grid, cell)
        if not self._ptr:
            raise RuntimeError("arcs_create failed (no GPU, or the budget holds no world)")

    @property
    def capacity(self) -> int:
        return self._lib.arcs_capacity(self._ptr)

    @property
    def used(self) -> int:
        return self._lib.arcs_used(self._ptr)

    def open(self, n_worlds: int, seed0: int = 0, difficulty: float = -1.0) -> int:
        cid = self._lib.arcs_open(self._ptr, n_worlds, seed0, difficulty)
        if cid < 0:
            raise AdmissionError(f"{n_worlds} worlds do not fit ({self.used} of {self.capacity} held)")
        return cid

    def close(self, cid: int) -> None:
        if self._lib.arcs_close(self._ptr, cid):
            raise KeyError(cid)

    def _count(self, cid: int) -> int:
        """World count of an open client, read from the server (which owns the
        client table) so that no Python-side copy can go stale."""
        return self.range(cid)[1]

    def range(self, cid: int) -> tuple[int, int]:
        s, c = _int(), _int()
        if self._lib.arcs_client_range(self._ptr, cid, ctypes.byref(s), ctypes.byref(c)):
            raise KeyError(cid)
        return s.value, c.value

    def submit(self, cid: int, targets: np.ndarray) -> int:
        t = np.ascontiguousarray(targets, dtype=np.float32)
        n = self._count(cid)
        if t.size != n * 12:
            raise ValueError(f"client {cid} needs {n} x 12 targets")
        ticket = self._lib.arcs_submit(self._ptr, cid, t.ctypes.data)
        if ticket < 0:
            raise RuntimeError("arcs_submit failed")
        return ticket

    def submit_device(self, cid: int, targets) -> int:
        """Queues a step whose targets already live in device memory: a
        float32 tensor on the GPU (n x 12 elements, contiguous), a
        native.DeviceBuffer, or a raw device address.  The memory must be
        written before the call and left unchanged until wait(ticket)."""
        w = self._count(cid)
        n = w * 12
        if hasattr(targets, "data_ptr"):
            if not targets.is_cuda or str(targets.dtype) != "torch.float32" or not targets.is_contiguous():
                raise ValueError("device targets must be a contiguous float32 GPU tensor")
            if targets.numel() != n:
                raise ValueError(f"client {cid} needs {w} x 12 targets")
            ptr = targets.data_ptr()
        elif hasattr(targets, "ptr"):
            if getattr(targets, "nbytes", 4 * n) < 4 * n:
                raise ValueError(f"client {cid} needs {w} x 12 targets")
            ptr = targets.ptr
        else:
            ptr = int(targets)
        ticket = self._lib.arcs_submit_device(self._ptr, cid, ptr)
        if ticket < 0:
            raise RuntimeError("arcs_submit_device failed")
        return ticket

    def wait(self, ticket: int) -> float:
        ms = self._lib.arcs_wait(self._ptr, ticket)
        if ms < 0:
            raise RuntimeError(f"ticket {ticket} failed or unknown")
        return ms

    def step(self, cid: int, targets: np.ndarray) -> float:
        return self.wait(self.submit(cid, targets))

    def stats(self) -> dict:
        out = np.zeros(10)
        self._lib.arcs_stats(self._ptr, out.ctypes.data)
        return dict(zip(STAT_KEYS, out.tolist()))

    def reset_stats(self) -> None:
        self._lib.arcs_reset_stats(self._ptr)

    def autotune(self, steps: int = 3) -> tuple[dict, float]:
        from .native import _launch_dict
        out = np.zeros(5, dtype=np.int32)
        ms = _dbl(0.0)
        if self._lib.arcs_autotune(self._ptr, steps, out.ctypes.data, ctypes.byref(ms)):
            raise RuntimeError("arcs_autotune failed")
        return _launch_dict(out), ms.value

    def set_window(self, microseconds: float) -> None:
        if self._lib.arcs_set_window(self._ptr, microseconds):
            raise ValueError(microseconds)

    def observe(self, cid: int) -> np.ndarray:
        out = np.zeros((self._count(cid), 13))
        if self._lib.arcs_observe(self._ptr, cid, out.ctypes.data):
            raise RuntimeError("arcs_observe failed")
        return out

    def close_server(self) -> None:
        if self._ptr:
            self._lib.arcs_destroy(self._ptr)
            self._ptr = None

    def __del__(self):
        self.close_server()
