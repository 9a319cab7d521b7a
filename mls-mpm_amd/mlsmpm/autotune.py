"""Autotuner of the tile shape (work-item size, drift margin) of single scenes
and of the launch shapes of a batch (persistent blocks per batched kernel
class), by timing candidate configurations on the device.

    tune_tile(grid, pos)          -> (ppt, margin) fastest for this scene alone
    tune_tiles(mix)               -> {class: ppt} for scenes stepped alone
    tune_tiles_batched(mix)       -> {class: ppt} for the mix stepped as one batch
    tune_batch(batch)             -> BatchTuning fastest for this batch

The best work-item size depends on the form: a scene stepped alone gains from
more, smaller items (more blocks in flight), while a batch has blocks enough
and pays for the extra ghost cores of small items, so the two are tuned
separately, by particle-count class (size_class).  Results are stored as JSON
(save / load; TABLE is the table tuned on MI355X that the package ships) so a
deployment applies a tuned table instead of re-timing at start-up:
apply_tiles sets a config's tile_ppt from one, and mlsmpm.Server takes the
batched table for its requests.

    python -m mlsmpm.autotune [--out PATH]    re-tunes the shipped table
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


def _mix_batch_ms(mix, ppt_of, steps: int, tuning=None) -> float:
    """Median batched step of the whole mix, each scene created with
    tile_ppt = ppt_of(n), after its first steps (graphs off, one event pair
    per step)."""
    sc = []
    try:
        for grid, pos, vel, *_ in mix:
            s = Scene(capi.default_config(*grid, len(pos), sort_period=1,
                                          tile_ppt=ppt_of(len(pos))))
            s.reset(pos, vel)
            sc.append(s)
        with Batch(tuning=tuning) as b:
            for s in sc:
                b.add(s)
            b.step(2)
            return _time_batch(b, steps)
    finally:
        for s in sc:
            s.close()


def tune_tiles_batched(mix, candidates=(8, 4, 16), steps: int = 12, rounds: int = 1,
                       tuning=None, verbose: bool = False) -> dict:
    """Work-item size per particle-count class for the mix stepped as one
    batch: coordinate descent from candidates[0] for every class, one class
    at a time, each candidate timed on the whole batch (a change is kept
    when it is at least 1% faster).  Returns {class: ppt}."""
    classes = sorted({size_class(len(p)) for _, p, *_ in mix})
    cur = {c: candidates[0] for c in classes}
    best = _mix_batch_ms(mix, lambda n: cur[size_class(n)], steps, tuning)
    for _ in range(rounds):
        changed = False
        for c in classes:
            for ppt in candidates:
                if ppt == cur[c]:
                    continue
                trial = {**cur, c: ppt}
                ms = _mix_batch_ms(mix, lambda n: trial[size_class(n)], steps, tuning)
                if verbose:
                    print(f"[autotune] batched {trial}: {ms:.3f} ms (best {best:.3f})", flush=True)
                if ms < best * 0.99:
                    best, cur, changed = ms, trial, True
        if not changed:
            break
    return cur


def apply_tiles(cfg: capi.Config, table: dict | None) -> capi.Config:
    """cfg with tile_ppt set from a {class: ppt} table (keys may be strings,
    as JSON stores them) for its capacity's class; left as it is when the
    table has no entry or cfg already names a size."""
    if table and not cfg.tile_ppt:
        ppt = {int(k): int(v) for k, v in table.items()}.get(size_class(cfg.capacity))
        if ppt:
            cfg.tile_ppt = ppt
    return cfg


def _time_batch(b: Batch, steps: int) -> float:
    b.enable_timing(True)
    b.step(steps)
    ms = float(np.median(b.step_times()))
    b.enable_timing(False)
    return ms


def tune_batch(b: Batch, steps: int = 6, rounds: int = 2) -> capi.BatchTuning:
    """Coordinate descent over the persistent block counts of the work-item,
    resolve and per-brick kernel classes (timed eagerly, one HIP
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


# The table this package ships, tuned on one MI355X with main() below.
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_mi355x.json")


def load_table(path: str = TABLE, key: str = "batch_ppt_by_class") -> dict | None:
    """One {class: ppt} table of a tuned file (None when absent)."""
    t = load(path)
    if not t or key not in t:
        return None
    return {int(k): int(v) for k, v in t[key].items()}


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


def main() -> int:
    """Tunes the shipped table on the device: the 256-scene batched mix of the
    benchmark (scene_mix(3, 256)) for the batched work-item sizes, the median
    scene of each class alone for the single-scene ones."""
    import argparse

    from . import scenes

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--scenes", type=int, default=256)
    a = ap.parse_args()
    mix = scenes.scene_mix(3, a.scenes)
    alone = tune_tiles(mix, steps=16)
    print(f"[autotune] alone {alone}", flush=True)
    batched = tune_tiles_batched(mix, verbose=True)
    table = {"device": "MI355X (gfx950)", "mix": f"scene_mix(3, {a.scenes})",
             "alone_ppt_by_class": alone, "batch_ppt_by_class": batched}
    save(a.out, table)
    print(json.dumps(table, sort_keys=True), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
