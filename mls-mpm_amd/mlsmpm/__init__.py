"""mlsmpm -- MI355X (gfx950) MLS-MPM fluid step behind a C ABI.

Python side of include/mlsmpm.h: ctypes binding (capi), scene handle
(scene.Scene), batches of scenes (batch.Batch), the request server
(server.Server), seeded scene layouts
(scenes) and the build recipe (build).
The compute path is libmlsmpm.so only.
"""
from .capi import (Config, MpmError, MPM_FLAG_NO_GRAPH, MPM_FLAG_KEEP_GRID, MPM_FLAG_NO_LDS,
                   config_from_aspect, default_config, load, lib_path)
from .scene import DREC_DTYPE, Scene
from .batch import Batch
from .server import Server
from .capi import BatchTuning, MPM_BATCH_NO_GRAPH, MPM_BATCH_GRAPH_REUSED

__all__ = [
    "Config", "MpmError", "Scene", "Batch", "Server", "BatchTuning", "MPM_BATCH_NO_GRAPH", "MPM_BATCH_GRAPH_REUSED", "DREC_DTYPE", "default_config", "config_from_aspect", "load",
    "lib_path", "MPM_FLAG_KEEP_GRID", "MPM_FLAG_NO_LDS", "MPM_FLAG_NO_GRAPH",
]
