"""Writes the golden fixtures of tests/golden/ with the CPU restatement
(oracle/): inputs and per-step outputs of two small scenes.

    python tests/golden/make_golden.py

  tiny8.npz      8^3 grid, 64 particles on a 4x4x4 half-cell lattice with a
                 seeded jitter and seeded initial velocities, 3 steps:
                 i32 grids, f32 grids and particles after every step
  dam16_d.npz    16^3 grid, 2304-particle dam-break block, diffuse pool of
                 1024 slots, 4 steps: positions and velocities after every
                 step, final C, pool, free list and top

  interact16.npz 16^3 grid, 1568-particle block, 6 steps with the
                 interaction operations between them: a hand push (steps
                 0-2, moving), deleteNear before step 3, 200 appended
                 particles before step 4, the push off from step 5, and a
                 raycast at the end: positions and velocities after steps
                 2, 3 and 5, final grids, C and density, the ray's hit
                 distance

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


# The interaction script of interact16: (before step k) -> call.
HAND_EYE = (8.0, 30.0, -12.0)
HANDS = {0: ((5.0, 5.0, 5.0), (1.0, 0.5, 0.0)), 1: ((5.5, 5.0, 5.5), (1.0, 0.5, 0.0)),
         2: ((6.0, 5.5, 6.0), (0.5, 1.0, 0.5))}
DELETE = ((4.0, 0.0, 4.0), (0.0, 1.0, 0.0), 1.5)   # before step 3
RAY = ((5.0, 15.0, 5.0), (0.0, -1.0, 0.0), 1.0)    # after the last step
KEPT_STEPS = (2, 3, 5)  # after the push, after the delete, the last (appended) step


def interact16_inputs():
    rng = np.random.default_rng(161)
    pos = scenes.lattice_block((2.0, 2.0, 2.0), (9.0, 9.0, 9.0), 0.1, rng)
    add = rng.uniform((9.0, 3.0, 9.0), (13.0, 8.0, 13.0), (200, 3)).astype(np.float32)
    addv = rng.uniform(-1.0, 1.0, (200, 3)).astype(np.float32)
    return pos, add, addv


def run_interact16(scene, pos, add, addv, on_step):
    """Drives an OracleScene or an mlsmpm.Scene through the script."""
    scene.reset(pos)
    for k in range(6):
        if k in HANDS:
            hp, hv = HANDS[k]
            scene.set_hand(hp, hv, 2.5, 0.7, True, HAND_EYE)
        if k == 3:
            scene.delete_near(*DELETE)
        if k == 4:
            scene.append(add, addv)
        if k == 5:
            scene.set_hand((0, 0, 0), (0, 0, 0), 1.0, 1.0, False, HAND_EYE)
        scene.step(1)
        on_step(k)
    return scene.raycast(*RAY)


def interact16():
    pos, add, addv = interact16_inputs()
    o = oracle.OracleScene(grid=(16, 16, 16))
    out = {"grid": np.array([16, 16, 16], np.int32), "pos3": pos, "add3": add, "addv3": addv}

    def keep(k):
        if k in KEPT_STEPS:
            out[f"pos4_{k}"] = o.pos4.copy()
            out[f"vel4_{k}"] = o.vel4.copy()

    hit, t = run_interact16(o, pos, add, addv, keep)
    out["cell_mass"] = o.cell_mass.copy()
    out["cell_mv"] = o.cell_mv.copy()
    out["cell_v"] = o.cell_v.copy()
    out["C9"] = o.affine.copy()
    out["density"] = o.density.copy()
    out["ray"] = np.array([1.0 if hit else 0.0, t], np.float32)
    return out


def main():
    for name, fn in (("tiny8", tiny8), ("dam16_d", dam16_d), ("interact16", interact16)):
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **fn())
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
