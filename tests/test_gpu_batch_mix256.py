"""The batched form at its workload scale: 256 dam-break scenes (grid edges
log-uniform in 32..128 per axis, 10k..2M particles log-uniform, capped at 8
per cell over half the interior; about 63M particles in all) stepped as one
batch with captured-graph replay.  Every member must end bit-identical to
the same scene stepped alone, and one member of each particle-count class
must match the CPU restatement after 3 steps."""
import hashlib

import numpy as np
import pytest

import mlsmpm
from mlsmpm import autotune, scenes
from oracle import oracle

pytestmark = pytest.mark.gpu

SEED = 3          # the mix bench.py measures (scene_mix(seed + 3, 256), seed 0)
STEPS_A, STEPS_B = 3, 16


def digest(s):
    p = s.read_particles()
    h = hashlib.blake2b(digest_size=16)
    for k in ("pos4", "vel4", "C9", "density"):
        h.update(np.ascontiguousarray(p[k]).tobytes())
    return h.hexdigest()


def test_256_scene_mix_batched_equals_alone_and_oracle():
    mix = scenes.scene_mix(SEED, 256)
    assert len(mix) == 256
    ntot = sum(len(p) for _, p, _, _ in mix)
    assert 30_000_000 < ntot < 100_000_000
    print(f"[mix256] {len(mix)} scenes, {ntot} particles", flush=True)
    # one member per particle-count class, the median of its class
    classes = {}
    for i, (_, pos, _, _) in enumerate(mix):
        classes.setdefault(autotune.size_class(len(pos)), []).append(i)
    sample = [sorted(v, key=lambda i: len(mix[i][1]))[len(v) // 2] for v in classes.values()]
    assert len(sample) >= 4
    sc = []
    for grid, pos, vel, _ in mix:
        s = mlsmpm.Scene(mlsmpm.default_config(*grid, len(pos), sort_period=1))
        s.reset(pos, vel)
        sc.append(s)
    try:
        with mlsmpm.Batch() as b:
            for s in sc:
                b.add(s)
            b.step(1)              # first step of every member: the full sort
            b.step(STEPS_A - 1)
            b.sync()
            for i in sample:
                grid, pos, vel, _ = mix[i]
                o = oracle.OracleScene(grid=grid)
                o.reset(pos, vel)
                o.step(STEPS_A, threads=16)
                p = sc[i].read_particles()
                for k, ref in (("pos4", o.pos4), ("vel4", o.vel4), ("C9", o.affine),
                               ("density", o.density)):
                    np.testing.assert_array_equal(p[k], ref, err_msg=f"member {i} {k}")
            print(f"[mix256] oracle members {sample} bit-exact after {STEPS_A} steps", flush=True)
            b.step(STEPS_B)        # one captured 16-step graph, replayed
            b.sync()
            batched = [digest(s) for s in sc]
        print("[mix256] batched run read back", flush=True)
        steps = STEPS_A + STEPS_B
        for i, ((grid, pos, vel, _), s) in enumerate(zip(mix, sc)):
            s.reset(pos, vel)
            s.step(steps)
            assert digest(s) == batched[i], f"member {i} ({len(pos)} particles, grid {grid})"
        print(f"[mix256] all 256 members equal their alone runs after {steps} steps", flush=True)
    finally:
        for s in sc:
            s.close()
