"""Python handle of a batch of scenes (mpm_batch_* of include/mlsmpm.h).

A batch steps all of its member scenes with one persistent launch per stage
(captured in a HIP graph); each member keeps its own Scene handle for reset
and read-back, and its results are bit-identical to stepping it alone.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import BatchTuning, check, load
from .scene import Scene


class Batch:
    def __init__(self, device: int = 0, tuning: BatchTuning | dict | None = None):
        self._h = C.c_void_p()
        check(load().mpm_batch_create(device, C.byref(self._h)))
        self.members: list[Scene] = []
        if tuning is not None:
            self.set_tuning(tuning)

    def close(self):
        if self._h:
            load().mpm_batch_destroy(self._h)
            self._h = C.c_void_p()
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, scene: Scene):
        check(load().mpm_batch_add(self._h, scene.handle))
        self.members.append(scene)

    def remove(self, scene: Scene):
        check(load().mpm_batch_remove(self._h, scene.handle))
        self.members.remove(scene)

    def __len__(self) -> int:
        n = C.c_uint32()
        check(load().mpm_batch_count(self._h, C.byref(n)))
        return n.value

    def tuning(self) -> BatchTuning:
        t = BatchTuning()
        check(load().mpm_batch_get_tuning(self._h, C.byref(t)))
        return t

    def set_tuning(self, t: BatchTuning | dict):
        if isinstance(t, dict):
            cur = self.tuning()
            for k, v in t.items():
                setattr(cur, k, v)
            t = cur
        check(load().mpm_batch_set_tuning(self._h, C.byref(t)))

    def step(self, n: int = 1):
        check(load().mpm_batch_step(self._h, n))

    def sync(self):
        check(load().mpm_batch_sync(self._h))

    def enable_timing(self, on: bool = True):
        check(load().mpm_batch_enable_timing(self._h, 1 if on else 0))

    def step_times(self) -> np.ndarray:
        cnt = C.c_uint32()
        check(load().mpm_batch_step_times(self._h, None, 0, C.byref(cnt)))
        ms = np.zeros(max(cnt.value, 1), np.float32)
        check(load().mpm_batch_step_times(self._h, ms.ctypes.data_as(C.c_void_p), cnt.value,
                                          C.byref(cnt)))
        return ms[: cnt.value]
