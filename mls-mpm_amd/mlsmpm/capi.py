"""ctypes binding of include/mlsmpm.h (libmlsmpm.so).

The library is the only compute path: there is no CPU fallback.  Loading
fails loudly when libmlsmpm.so has not been built, and every call that
returns an MPM_ERR_* code raises MpmError with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os

from .build import LIB

MPM_OK = 0
MPM_ERR_INVALID = -1
MPM_ERR_RANGE = -2
MPM_ERR_CAPACITY = -3
MPM_ERR_HIP = -4
MPM_ERR_NOMEM = -5
MPM_ERR_UNSUPPORTED = -6
MPM_ERR_STATE = -7
MPM_ERR_NOT_FOUND = -8

MPM_RAYCAST_NOALLOC_MAX = 1024
MPM_EXPORT_POS = 1
MPM_EXPORT_VEL = 2
MPM_EXPORT_DENSITY = 4
MPM_EXPORT_DIFFUSE = 8
MPM_EXPORT_ALL = 15
MPM_EXPORT_RESET_ORDER = 16

MPM_FLAG_KEEP_GRID = 1
MPM_FLAG_NO_LDS = 2
MPM_FLAG_NO_GRAPH = 4

ERROR_NAMES = {
    MPM_ERR_INVALID: "MPM_ERR_INVALID", MPM_ERR_RANGE: "MPM_ERR_RANGE",
    MPM_ERR_CAPACITY: "MPM_ERR_CAPACITY", MPM_ERR_HIP: "MPM_ERR_HIP",
    MPM_ERR_NOMEM: "MPM_ERR_NOMEM", MPM_ERR_UNSUPPORTED: "MPM_ERR_UNSUPPORTED",
    MPM_ERR_STATE: "MPM_ERR_STATE", MPM_ERR_NOT_FOUND: "MPM_ERR_NOT_FOUND",
}


class MpmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


class Config(C.Structure):
    """mpm_config (include/mlsmpm.h)."""

    _fields_ = [
        ("grid_x", C.c_uint32), ("grid_y", C.c_uint32), ("grid_z", C.c_uint32),
        ("capacity", C.c_uint32), ("diffuse_capacity", C.c_uint32), ("sort_period", C.c_uint32),
        ("flags", C.c_uint32), ("device", C.c_int32),
        ("dt", C.c_float), ("gravity", C.c_float), ("rest_density", C.c_float),
        ("stiffness", C.c_float), ("eos_power", C.c_float), ("viscosity", C.c_float),
        ("hard_min", C.c_float), ("hard_max_x", C.c_float), ("hard_max_y", C.c_float),
        ("hard_max_z", C.c_float), ("wall_min", C.c_float), ("wall_stiffness", C.c_float),
        ("lookahead_k", C.c_float),
        ("ke_threshold", C.c_float), ("pressure_threshold", C.c_float),
        ("crest_dot_threshold", C.c_float), ("spawn_rate_k", C.c_float),
        ("lifetime_foam", C.c_float), ("buoyancy", C.c_float), ("drag", C.c_float),
        ("density_spray_max", C.c_float), ("density_bubble_min", C.c_float),
        ("seed", C.c_uint32), ("tile_ppt", C.c_uint32), ("tile_margin", C.c_uint32),
        ("tile_threads", C.c_uint32), ("g2p_threads", C.c_uint32),
        ("lifetime_mode", C.c_uint32), ("reserved", C.c_uint32 * 1),
    ]

    def copy(self) -> "Config":
        c = Config()
        C.pointer(c)[0] = self
        return c


MPM_BATCH_NO_GRAPH = 1
MPM_BATCH_GRAPH_REUSED = 2


class BatchTuning(C.Structure):
    """mpm_batch_tuning: persistent blocks per batched kernel class (0 = one
    block per unit), steps per captured graph, flags."""
    _fields_ = [("item_blocks", C.c_uint32), ("resolve_blocks", C.c_uint32),
                ("aux_blocks", C.c_uint32), ("graph_steps", C.c_uint32), ("flags", C.c_uint32),
                ("reserved", C.c_uint32 * 3)]


MPM_REQ_QUEUED, MPM_REQ_RUNNING, MPM_REQ_DONE, MPM_REQ_FAILED, MPM_REQ_CANCELLED = range(5)


class ServerConfig(C.Structure):
    """mpm_server_config."""
    _fields_ = [("device", C.c_int32), ("max_active", C.c_uint32), ("hbm_budget", C.c_uint64),
                ("chunk_steps", C.c_uint32), ("item_blocks", C.c_uint32),
                ("resolve_blocks", C.c_uint32), ("aux_blocks", C.c_uint32),
                ("tile_ppt_by_class", C.c_uint8 * 16), ("batch_flags", C.c_uint32),
                ("pad", C.c_uint32)]


class ServerStats(C.Structure):
    """mpm_server_stats."""
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("cancelled", C.c_uint64), ("queued", C.c_uint32), ("active", C.c_uint32),
                ("bytes_resident", C.c_uint64), ("particle_steps", C.c_uint64),
                ("batched_steps", C.c_uint64), ("resident_steps", C.c_uint64),
                ("busy_seconds", C.c_double),
                ("elapsed_seconds", C.c_double),
                ("enqueue_seconds", C.c_double), ("admit_seconds", C.c_double),
                ("wait_seconds", C.c_double), ("retire_seconds", C.c_double),
                ("enroll_seconds", C.c_double), ("gpu_step_seconds", C.c_double),
                ("chunks", C.c_uint64), ("admitted", C.c_uint64),
                ("graph_steps", C.c_uint64), ("graph_captures", C.c_uint64),
                ("admit_create_seconds", C.c_double), ("admit_load_seconds", C.c_double),
                ("retire_read_seconds", C.c_double), ("retire_release_seconds", C.c_double)]


class StepRecord(C.Structure):
    """mpm_step_record."""
    _fields_ = [("ms", C.c_float), ("resident", C.c_uint32), ("admitting", C.c_uint32),
                ("pad", C.c_uint32), ("resident_particles", C.c_uint64)]


class RequestInfo(C.Structure):
    """mpm_request_info."""
    _fields_ = [("state", C.c_int32), ("n", C.c_uint32), ("steps", C.c_uint32),
                ("steps_done", C.c_uint32), ("footprint", C.c_uint64), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("tile_ppt", C.c_uint32),
                ("pad", C.c_uint32)]


class Hand(C.Structure):
    """mpm_hand: the push of FluidGPU.updateHand (fluid-gpu.js:479-529)."""
    _fields_ = [("pos", C.c_float * 3), ("radius", C.c_float), ("vel", C.c_float * 3),
                ("active", C.c_uint32), ("strength", C.c_float), ("eye", C.c_float * 3)]


assert C.sizeof(Hand) == 48


class Export(C.Structure):
    """mpm_export: device pointers of the renderer buffers (0 = skip)."""
    _fields_ = [("pos4", C.c_void_p), ("vel4", C.c_void_p), ("density", C.c_void_p),
                ("diffuse48", C.c_void_p)]


class Frame(C.Structure):
    """mpm_frame: one exported frame in pinned host memory."""
    _fields_ = [("pos4", C.c_void_p), ("vel4", C.c_void_p), ("density", C.c_void_p),
                ("diffuse48", C.c_void_p), ("n", C.c_uint32), ("top", C.c_int32),
                ("steps", C.c_uint64), ("ticket", C.c_uint64)]


# Every entry point declared in include/mlsmpm.h, with its argument types.
_VP = C.c_void_p
_SIGNATURES = {
    "mpm_version": ([], C.c_char_p),
    "mpm_last_error": ([], C.c_char_p),
    "mpm_config_default": ([_VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32], C.c_int),
    "mpm_config_from_aspect": ([_VP, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32], C.c_int),
    "mpm_create": ([_VP, _VP], C.c_int),
    "mpm_destroy": ([_VP], C.c_int),
    "mpm_reset": ([_VP, C.c_uint32, _VP, _VP, _VP, _VP], C.c_int),
    "mpm_step": ([_VP, C.c_uint32], C.c_int),
    "mpm_sync": ([_VP], C.c_int),
    "mpm_prepare": ([_VP], C.c_int),
    "mpm_get_counts": ([_VP, _VP, _VP], C.c_int),
    "mpm_read_particles": ([_VP, _VP, _VP, _VP, _VP], C.c_int),
    "mpm_read_grid": ([_VP, _VP, _VP, _VP, _VP], C.c_int),
    "mpm_read_diffuse": ([_VP, _VP, _VP, _VP], C.c_int),
    "mpm_write_diffuse": ([_VP, _VP, _VP, C.c_int32], C.c_int),
    "mpm_diffuse_alive": ([_VP, _VP], C.c_int),
    "mpm_set_params": ([_VP, _VP], C.c_int),
    "mpm_get_config": ([_VP, _VP], C.c_int),
    "mpm_set_seed": ([_VP, C.c_uint32], C.c_int),
    "mpm_set_walls": ([_VP, _VP], C.c_int),
    "mpm_set_hand": ([_VP, _VP], C.c_int),
    "mpm_delete_near": ([_VP, _VP, _VP, C.c_float, _VP], C.c_int),
    "mpm_raycast": ([_VP, _VP, _VP, C.c_float, _VP, _VP], C.c_int),
    "mpm_raycast_noalloc": ([_VP, C.c_uint32, _VP, _VP, _VP, _VP, _VP], C.c_int),
    "mpm_append_particles": ([_VP, C.c_uint32, _VP, _VP], C.c_int),
    "mpm_export_device": ([_VP, _VP, C.c_uint32], C.c_int),
    "mpm_frame_begin": ([_VP, C.c_uint32, _VP], C.c_int),
    "mpm_frame_poll": ([_VP, C.c_uint64, C.c_int, _VP, _VP], C.c_int),
    "mpm_get_stream": ([_VP, _VP], C.c_int),
    "mpm_enable_timing": ([_VP, C.c_int], C.c_int),
    "mpm_stage_times": ([_VP, _VP], C.c_int),
    "mpm_step_times": ([_VP, _VP, C.c_uint32, _VP], C.c_int),
    "mpm_tile_stats": ([_VP, _VP], C.c_int),
    "mpm_batch_create": ([C.c_int, _VP], C.c_int),
    "mpm_batch_destroy": ([_VP], C.c_int),
    "mpm_batch_add": ([_VP, _VP], C.c_int),
    "mpm_batch_remove": ([_VP, _VP], C.c_int),
    "mpm_batch_count": ([_VP, _VP], C.c_int),
    "mpm_batch_set_tuning": ([_VP, _VP], C.c_int),
    "mpm_batch_get_tuning": ([_VP, _VP], C.c_int),
    "mpm_batch_step": ([_VP, C.c_uint32], C.c_int),
    "mpm_batch_sync": ([_VP], C.c_int),
    "mpm_batch_enable_timing": ([_VP, C.c_int], C.c_int),
    "mpm_batch_step_times": ([_VP, _VP, C.c_uint32, _VP], C.c_int),
    "mpm_server_create": ([_VP, _VP], C.c_int),
    "mpm_server_destroy": ([_VP], C.c_int),
    "mpm_server_submit": ([_VP, _VP, C.c_uint32, _VP, _VP, _VP, _VP, C.c_uint32, _VP], C.c_int),
    "mpm_server_poll": ([_VP, C.c_uint64, _VP], C.c_int),
    "mpm_server_wait": ([_VP, C.c_uint64, C.c_double, _VP], C.c_int),
    "mpm_server_result": ([_VP, C.c_uint64, _VP, _VP, _VP, _VP], C.c_int),
    "mpm_server_cancel": ([_VP, C.c_uint64], C.c_int),
    "mpm_server_get_stats": ([_VP, _VP], C.c_int),
    "mpm_server_step_times": ([_VP, _VP, C.c_uint32, _VP], C.c_int),
    "mpm_server_step_records": ([_VP, _VP, C.c_uint32, _VP], C.c_int),
}

_lib = None


def lib_path() -> str:
    """The library this process loads: $MLSMPM_LIB if set, else the in-tree build."""
    return os.environ.get("MLSMPM_LIB") or LIB


def load() -> C.CDLL:
    """Load libmlsmpm.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run __graft_entry__.build() "
                               "or python -m mlsmpm.build (there is no CPU fallback)")
        L = C.CDLL(path)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    """Names of the entry points this binding declares."""
    return list(_SIGNATURES)


def check(rc: int) -> int:
    if rc != MPM_OK:
        msg = load().mpm_last_error().decode(errors="replace")
        raise MpmError(rc, msg)
    return rc


def default_config(gx: int, gy: int, gz: int, capacity: int, **kw) -> Config:
    c = Config()
    check(load().mpm_config_default(C.byref(c), gx, gy, gz, capacity))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def config_from_aspect(ax: float, ay: float, az: float, radius: float, capacity: int) -> Config:
    c = Config()
    check(load().mpm_config_from_aspect(C.byref(c), ax, ay, az, radius, capacity))
    return c
