"""Writes the golden fixtures of tests/golden/ with the CPU restatement
(oracle/): inputs and per-step outputs of two small scenes.

    python tests/golden/make_golden.py

  tiny8.npz      8^3 grid, 64 particles on a 4x4x4 half-cell lattice with a
                 seeded jitter and seeded initial velocities, 3 steps:
                 i32 grids, f32 grids and particles after every step
  dam16_d.npz    16^3 grid, 2304-particle dam-break block, diffuse pool of
                 1024 slots, 4 steps: positions and velocities after every
                 step, final C, pool, free list and top

The fixtures pin the oracle and the HIP library against silent change; the
oracle itself is pinned by tests/test_oracle_pins.py.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "mls-mpm_amd")]

from oracle import oracle  # noqa: E402
from mlsmpm import scenes  # noqa: E402


def tiny8():
    rng = np.random.default_rng(8)
    pos = scenes.lattice_block((3.0, 3.0, 3.0), (5.0, 5.0, 5.0), 0.1, rng)
    vel = rng.uniform(-0.5, 0.5, pos.shape).astype(np.float32)
    o = oracle.OracleScene(grid=(8, 8, 8))
    o.reset(pos, vel)
    out = {"grid": np.array([8, 8, 8], np.int32), "pos3": pos, "vel3": vel}
    for k in range(3):
        o.step(1)
        out[f"cell_mass_{k}"] = o.cell_mass.copy()
        out[f"cell_mv_{k}"] = o.cell_mv.copy()
        out[f"cell_mass_f_{k}"] = o.cell_mass_f.copy()
        out[f"cell_v_{k}"] = o.cell_v.copy()
        out[f"pos4_{k}"] = o.pos4.copy()
        out[f"vel4_{k}"] = o.vel4.copy()
        out[f"C9_{k}"] = o.affine.copy()
        out[f"density_{k}"] = o.density.copy()
    return out


def dam16_d():
    rng = np.random.default_rng(16)
    pos = scenes.lattice_block((2.0, 2.0, 2.0), (8.0, 10.0, 8.0), 0.1, rng)
    D = 1024
    o = oracle.OracleScene(grid=(16, 16, 16), diffuse_max=D, seed=1)
    o.reset(pos)
    out = {"grid": np.array([16, 16, 16], np.int32), "pos3": pos,
           "diffuse_max": np.array([D], np.int32)}
    for k in range(4):
        o.step(1)
        out[f"pos4_{k}"] = o.pos4.copy()
        out[f"vel4_{k}"] = o.vel4.copy()
    out["C9_3"] = o.affine.copy()
    rec = o.pool
    for f in ("pos", "vel", "ptype", "alive", "lifetime", "pad"):
        out[f"pool_{f}"] = np.ascontiguousarray(rec[f])
    out["free_list"] = o.free_list.copy()
    out["top"] = np.array([int(o.top[0])], np.int32)
    return out


def main():
    for name, fn in (("tiny8", tiny8), ("dam16_d", dam16_d)):
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **fn())
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
