"""Seeded particle layouts for the benchmark and test scenes.

The reference fills blocks with a 0.5-cell lattice offset by a quarter cell
and a random jitter (FluidGPU.fillBlock, fluid-gpu.js:107-128, unseeded
Math.random, jitter +/-0.025).  The scenes here use the same lattice with a
seeded jitter so runs are reproducible.
"""
from __future__ import annotations

import math

import numpy as np


def lattice_block(lo, hi, jitter: float, rng: np.random.Generator) -> np.ndarray:
    """Points lo + 0.25 + 0.5 k (k >= 0, < hi) per axis, x fastest, plus
    uniform jitter in [-jitter, jitter]; float32 (n, 3)."""
    axes = [np.arange(lo[a] + 0.25, hi[a], 0.5, dtype=np.float64) for a in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    if jitter:
        pos = pos + rng.uniform(-jitter, jitter, pos.shape)
    return pos.astype(np.float32)


def fill_block(grid, x0, y0, z0, x1, y1, z1, rng: np.random.Generator, capacity=None) -> np.ndarray:
    """fillBlock with normalised bounds (fluid-gpu.js:107-128), seeded jitter
    (rand - 0.5) * 0.05; the loop bounds are computed in double like JS."""
    gx, gy, gz = grid
    lo = (x0 * gx, y0 * gy, z0 * gz)
    hi = (x1 * gx, y1 * gy, z1 * gz)
    pts = []
    cap = capacity if capacity is not None else 1 << 62
    zz = lo[2] + 0.25
    while zz < hi[2]:
        yy = lo[1] + 0.25
        while yy < hi[1]:
            xx = lo[0] + 0.25
            while xx < hi[0]:
                if len(pts) >= cap:
                    return np.asarray(pts, np.float32)
                pts.append((xx, yy, zz))
                xx += 0.5
            yy += 0.5
        zz += 0.5
    pos = np.asarray(pts, np.float64).reshape(-1, 3)
    pos = pos + (rng.random(pos.shape) - 0.5) * 0.05
    return pos.astype(np.float32)


def dam_break_small(seed: int = 0):
    """64^3 grid, block [2,30]x[2,40]x[2,30] cells, jitter +/-0.1:
    56*76*56 = 238,336 particles."""
    rng = np.random.default_rng(seed)
    return (64, 64, 64), lattice_block((2, 2, 2), (30, 40, 30), 0.1, rng), None


def dam_break_large(seed: int = 0):
    """256x128x128 grid, dam-break column x in [2,104), y in [2,43.5),
    z in [2,126): 204*83*248 = 4,199,136 particles, jitter +/-0.1."""
    rng = np.random.default_rng(seed)
    return (256, 128, 128), lattice_block((2, 2, 2), (104, 43.5, 126), 0.1, rng), None


def two_block_collision(seed: int = 0, grid=(256, 256, 256), speed_cells_per_step=2.0, dt=0.3):
    """256^3 grid, two 64x64x128-cell blocks moving toward each other at
    +/- 2 cells per step: 2 * 4,194,304 = 8,388,608 particles."""
    rng = np.random.default_rng(seed)
    a = lattice_block((8, 2, 64), (72, 66, 192), 0.1, rng)
    b = lattice_block((184, 2, 64), (248, 66, 192), 0.1, rng)
    pos = np.concatenate([a, b])
    v = speed_cells_per_step / dt
    vel = np.zeros_like(pos)
    vel[: len(a), 0] = v
    vel[len(a):, 0] = -v
    return grid, pos, vel


def scaled_dam_break(grid, n_target: int, rng: np.random.Generator, kind: str = "dam"):
    """A dam-break column (or two approaching blocks, kind='two') sized to
    about n_target particles inside a grid; capped at 8 particles per cell
    over half of the interior."""
    gx, gy, gz = grid
    inner = [(2.0, gx - 2.0), (2.0, gy - 2.0), (2.0, gz - 2.0)]
    vol_cap = 0.5 * (gx - 4) * (gy - 4) * (gz - 4)
    n_cells = min(n_target / 8.0, vol_cap)
    # column spanning the full z interior, x fraction and height sized to n
    zlen = inner[2][1] - inner[2][0]
    ylen_max = inner[1][1] - inner[1][0]
    xlen_max = inner[0][1] - inner[0][0]
    area = n_cells / zlen
    h = min(ylen_max, max(1.0, math.sqrt(area * ylen_max / xlen_max)))
    w = min(xlen_max, max(1.0, area / h))
    h = min(ylen_max, max(1.0, area / w))
    w = math.floor(w * 2) / 2
    h = math.floor(h * 2) / 2
    if kind == "two":
        half = max(0.5, math.floor(w / 4) * 2 / 2)
        a = lattice_block((2.0, 2.0, 2.0), (2.0 + half, 2.0 + h, 2.0 + zlen), 0.1, rng)
        b = lattice_block((gx - 2.0 - half, 2.0, 2.0), (gx - 2.0, 2.0 + h, 2.0 + zlen), 0.1, rng)
        pos = np.concatenate([a, b])
        vel = np.zeros_like(pos)
        vel[: len(a), 0] = 2.0
        vel[len(a):, 0] = -2.0
        return np.clip(pos, 1.0, None), vel
    pos = lattice_block((2.0, 2.0, 2.0), (2.0 + w, 2.0 + h, 2.0 + zlen), 0.1, rng)
    return pos, None


def scene_mix(seed: int, count: int, edge=(32, 128), n_range=(10_000, 2_000_000),
              kinds=("dam",)):
    """Seeded mix of independent scenes: every grid edge drawn log-uniform
    (integer) in `edge` per axis, particle target log-uniform in `n_range`,
    scene kind drawn from `kinds` ('dam' column or 'two' approaching blocks),
    sized by scaled_dam_break (so capped at 8 particles per cell over half
    the interior; the realized count is len(pos)).  Returns a list of
    (grid, pos, vel, kind)."""
    rng = np.random.default_rng(seed)
    out = []
    lo_e, hi_e = math.log(edge[0]), math.log(edge[1] + 1)
    lo_n, hi_n = math.log(n_range[0]), math.log(n_range[1])
    for _ in range(count):
        g = tuple(int(min(edge[1], max(edge[0], math.floor(math.exp(rng.uniform(lo_e, hi_e))))))
                  for _ in range(3))
        n = int(math.exp(rng.uniform(lo_n, hi_n)))
        kind = kinds[int(rng.integers(len(kinds)))]
        pos, vel = scaled_dam_break(g, n, rng, kind)
        out.append((g, pos, vel, kind))
    return out
