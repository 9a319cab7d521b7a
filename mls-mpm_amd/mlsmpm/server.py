"""Python handle of the serving form (mpm_server_* of include/mlsmpm.h):
a request loop with continuous batching on one device."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import (Config, MPM_REQ_CANCELLED, MPM_REQ_DONE, MPM_REQ_FAILED, RequestInfo,
                   ServerConfig, ServerStats, StepRecord, check, load)

_FINAL = (MPM_REQ_DONE, MPM_REQ_FAILED, MPM_REQ_CANCELLED)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Server:
    def __init__(self, hbm_budget: int, device: int = 0, max_active: int = 0,
                 chunk_steps: int = 8, item_blocks: int = 0, resolve_blocks: int = 0,
                 aux_blocks: int = 0, tile_ppt_by_class: dict | None = None,
                 batch_flags: int = 0):
        """tile_ppt_by_class: {particle-count class: work-item size} applied to
        requests that leave tile_ppt at 0 (autotune.size_class gives a
        request's class; autotune.load_table reads a tuned table).
        batch_flags: the batch's tuning flags (0: graphs replayed for member
        sets that last more than one chunk; MPM_BATCH_NO_GRAPH: eager)."""
        cfg = ServerConfig(device=device, max_active=max_active, hbm_budget=int(hbm_budget),
                           chunk_steps=chunk_steps, item_blocks=item_blocks,
                           resolve_blocks=resolve_blocks, aux_blocks=aux_blocks,
                           batch_flags=batch_flags)
        for c, ppt in (tile_ppt_by_class or {}).items():
            cfg.tile_ppt_by_class[int(c)] = int(ppt)
        self._h = C.c_void_p()
        check(load().mpm_server_create(C.byref(cfg), C.byref(self._h)))
        self._n = {}

    def close(self):
        if self._h:
            load().mpm_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, cfg: Config, pos, steps: int, vel=None, C9=None, mass=None) -> int:
        pos = np.ascontiguousarray(np.asarray(pos, np.float32).reshape(-1, 3))
        n = pos.shape[0]
        f = lambda a, k: None if a is None else np.ascontiguousarray(
            np.asarray(a, np.float32).reshape(n, k))
        vel, C9 = f(vel, 3), f(C9, 9)
        mass = None if mass is None else np.ascontiguousarray(np.asarray(mass, np.float32).reshape(n))
        rid = C.c_uint64()
        check(load().mpm_server_submit(self._h, C.byref(cfg), n, _p(pos), _p(vel), _p(C9),
                                       _p(mass), steps, C.byref(rid)))
        self._n[rid.value] = n
        return rid.value

    def poll(self, rid: int) -> dict:
        info = RequestInfo()
        check(load().mpm_server_poll(self._h, rid, C.byref(info)))
        return {k: getattr(info, k) for k, _ in RequestInfo._fields_}

    def wait(self, rid: int, timeout: float = -1.0) -> int:
        st = C.c_int()
        check(load().mpm_server_wait(self._h, rid, timeout, C.byref(st)))
        return st.value

    def result(self, rid: int) -> dict:
        n = self._n.pop(rid)
        out = {"pos4": np.empty((n, 4), np.float32), "vel4": np.empty((n, 4), np.float32),
               "C9": np.empty((n, 9), np.float32), "density": np.empty(n, np.float32)}
        check(load().mpm_server_result(self._h, rid, _p(out["pos4"]), _p(out["vel4"]),
                                       _p(out["C9"]), _p(out["density"])))
        return out

    def cancel(self, rid: int):
        check(load().mpm_server_cancel(self._h, rid))

    def stats(self) -> dict:
        s = ServerStats()
        check(load().mpm_server_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def step_times(self) -> np.ndarray:
        cnt = C.c_uint32()
        check(load().mpm_server_step_times(self._h, None, 0, C.byref(cnt)))
        ms = np.zeros(max(cnt.value, 1), np.float32)
        check(load().mpm_server_step_times(self._h, _p(ms), cnt.value, C.byref(cnt)))
        return ms[: cnt.value]

    def step_records(self) -> np.ndarray:
        """Per batched step: device ms, resident scenes, scenes being admitted
        meanwhile, resident particles."""
        cnt = C.c_uint32()
        check(load().mpm_server_step_records(self._h, None, 0, C.byref(cnt)))
        rec = (StepRecord * max(cnt.value, 1))()
        check(load().mpm_server_step_records(self._h, rec, cnt.value, C.byref(cnt)))
        dt = np.dtype([("ms", np.float32), ("resident", np.uint32), ("admitting", np.uint32),
                       ("pad", np.uint32), ("resident_particles", np.uint64)])
        return np.frombuffer(bytes(rec), dtype=dt)[: cnt.value].copy()
