"""FluidHIP + DiffuseHIP driven from Node through the N-API addon on the GPU:
a seeded fillBlock scene with an attached diffuse pool and a wall that moves
and turns between frames.  Every frame's positions and velocities, and the
final pool, must equal the CPU restatement replaying the same inputs and the
same wall uniforms."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from mlsmpm.scene import DREC_DTYPE
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS = os.path.join(ROOT, "mls-mpm_amd", "js")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not shutil.which("node"), reason="node is missing")]


def test_fluidhip_scene_matches_oracle(tmp_path):
    frames = 6
    r = subprocess.run(["node", os.path.join(JS, "test", "run_scene.mjs"), str(tmp_path),
                        str(frames)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    n, grid = info["n"], tuple(info["grid"])
    assert grid == (20, 20, 30)
    pos3 = np.fromfile(tmp_path / "pos3.f32", np.float32).reshape(n, 3)
    walls = np.fromfile(tmp_path / "walls.f32", np.float32).reshape(frames, 84)
    assert walls[0, 0] == 1 and walls[-1, 4 + 16 + 1] != 0  # one wall, turning at the end
    o = oracle.OracleScene(grid=grid, diffuse_max=8192, seed=1)
    o.reset(pos3)
    for k in range(frames):
        o.walls = np.ascontiguousarray(walls[k])
        o.step(1)
        p4 = np.fromfile(tmp_path / f"pos4_{k}.f32", np.float32).reshape(n, 4)
        v4 = np.fromfile(tmp_path / f"vel4_{k}.f32", np.float32).reshape(n, 4)
        np.testing.assert_array_equal(p4, o.pos4, err_msg=f"pos frame {k}")
        np.testing.assert_array_equal(v4, o.vel4, err_msg=f"vel frame {k}")
    pool = np.fromfile(tmp_path / "pool.bin", DREC_DTYPE)
    for f in ("pos", "vel", "ptype", "alive", "lifetime", "pad"):
        np.testing.assert_array_equal(pool[f], o.pool[f], err_msg=f"pool {f}")
    assert info["alive"] == 8192 - int(o.top[0])
