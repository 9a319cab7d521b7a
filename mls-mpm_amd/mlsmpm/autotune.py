"""Autotuner of the tile shape (work-item size, drift margin) of single scenes
and of the launch shapes of a batch (persistent blocks per batched kernel
class), by timing candidate configurations on the device.

    tune_tile(grid, pos)        -> (ppt, margin) fastest for this scene
    tune_tiles(scenes)          -> ppt per scene (by particle-count class)
    tune_batch(batch)           -> BatchTuning fastest for this batch

Results can be stored as JSON (save / load) so a deployment ships its tuned
table instead of re-timing at start-up.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import capi
from .batch import Batch
from .scene import Scene

TILE_CANDIDATES = ((8, 2), (4, 2), (16, 2), (8, 1), (4, 1))
ITEM_BLOCKS = (0, 1024, 2048, 4096, 8192)
RESOLVE_BLOCKS = (0, 512, 1024, 2048, 4096)
AUX_BLOCKS = (0, 256, 1024, 2048)


def _time_scene(grid, pos, vel, ppt, margin, steps) -> float:
    cfg = capi.default_config(*grid, len(pos), tile_ppt=ppt, tile_margin=margin, sort_period=1)
    with Scene(cfg) as s:
        s.reset(pos, vel)
        s.step(1)
        s.enable_timing(True)
        s.step(steps)
        ms = float(np.median(s.step_times()))
    return ms


def tune_tile(grid, pos, vel=None, candidates=TILE_CANDIDATES, steps: int = 8,
              margins=None) -> tuple[int, int]:
    """Work-item size and drift margin with the lowest median step time."""
    best, best_ms = None, math.inf
    for ppt, m in candidates:
        if margins is not None and m not in margins:
            continue
        ms = _time_scene(grid, pos, vel, ppt, m, steps)
        if ms < best_ms:
            best, best_ms = (ppt, m), ms
    return best


def size_class(n: int) -> int:
    """Particle-count class: floor(log2(n)) in steps of two octaves."""
    return int(math.log2(max(n, 1))) // 2


def tune_tiles(mix, margin: int = 2, steps: int = 8) -> dict:
    """Work-item size per particle-count class of a scene mix (all with one
    margin, as a batch requires): the median scene of each class is timed
    with every candidate size.  Returns {class: ppt}."""
    by_class: dict[int, list] = {}
    for grid, pos, vel, *_ in mix:
        by_class.setdefault(size_class(len(pos)), []).append((grid, pos, vel))
    out = {}
    for c, members in sorted(by_class.items()):
        members.sort(key=lambda m: len(m[1]))
        grid, pos, vel = members[len(members) // 2]
        out[c] = tune_tile(grid, pos, vel, [(p, margin) for p in (4, 8, 16)], steps)[0]
    return out


def _time_batch(b: Batch, steps: int) -> float:
    b.enable_timing(True)
    b.step(steps)
    ms = float(np.median(b.step_times()))
    b.enable_timing(False)
    return ms


def tune_batch(b: Batch, steps: int = 6, rounds: int = 2) -> capi.BatchTuning:
    """Coordinate descent over the persistent block counts of the work-item,
    resolve and per-brick/per-particle kernel classes (timed eagerly, one HIP
    event pair per batched step).  Leaves the batch set to the result."""
    t = b.tuning()
    cur = {"item_blocks": t.item_blocks, "resolve_blocks": t.resolve_blocks,
           "aux_blocks": t.aux_blocks}
    space = {"item_blocks": ITEM_BLOCKS, "resolve_blocks": RESOLVE_BLOCKS,
             "aux_blocks": AUX_BLOCKS}
    b.step(1)  # first steps after a reset are not batched
    b.set_tuning(cur)
    best_ms = _time_batch(b, steps)
    for _ in range(rounds):
        changed = False
        for k, cands in space.items():
            for v in cands:
                if v == cur[k]:
                    continue
                trial = dict(cur, **{k: v})
                b.set_tuning(trial)
                ms = _time_batch(b, steps)
                if ms < best_ms * 0.99:
                    best_ms, cur, changed = ms, trial, True
        if not changed:
            break
    b.set_tuning(cur)
    return b.tuning()


def save(path: str, table: dict) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")


def load(path: str) -> dict | None:
    try:
        with open(path) as f:
            return json.load(f)
    except FileNotFoundError:
        return None
