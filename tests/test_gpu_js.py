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


@pytest.mark.parametrize("substeps", [1, 3])
def test_fluidhip_scene_matches_oracle(tmp_path, substeps):
    """SUBSTEPS steps of DT / SUBSTEPS per simFrame (fluid-gpu.js:648-684),
    the wall uniform packed once per frame."""
    frames = 6
    r = subprocess.run(["node", os.path.join(JS, "test", "run_scene.mjs"), str(tmp_path),
                        str(frames), str(substeps)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    n, grid = info["n"], tuple(info["grid"])
    assert grid == (20, 20, 30)
    pos3 = np.fromfile(tmp_path / "pos3.f32", np.float32).reshape(n, 3)
    walls = np.fromfile(tmp_path / "walls.f32", np.float32).reshape(frames, 84)
    assert walls[0, 0] == 1 and walls[-1, 4 + 16 + 1] != 0  # one wall, turning at the end
    o = oracle.OracleScene(grid=grid, diffuse_max=8192, seed=1)
    o.params.dt = float(np.float32(0.3 / substeps))
    o.reset(pos3)
    for k in range(frames):
        o.walls = np.ascontiguousarray(walls[k])
        o.step(substeps)
        p4 = np.fromfile(tmp_path / f"pos4_{k}.f32", np.float32).reshape(n, 4)
        v4 = np.fromfile(tmp_path / f"vel4_{k}.f32", np.float32).reshape(n, 4)
        np.testing.assert_array_equal(p4, o.pos4, err_msg=f"pos frame {k}")
        np.testing.assert_array_equal(v4, o.vel4, err_msg=f"vel frame {k}")
    pool = np.fromfile(tmp_path / "pool.bin", DREC_DTYPE)
    for f in ("pos", "vel", "ptype", "alive", "lifetime", "pad"):
        np.testing.assert_array_equal(pool[f], o.pool[f], err_msg=f"pool {f}")
    assert info["alive"] == 8192 - int(o.top[0])
    # the pinned frame handoff holds the same final state
    assert info["frame"] == {"n": n, "top": int(o.top[0]), "steps": frames * substeps}
    np.testing.assert_array_equal(np.fromfile(tmp_path / "frame_pos4.f32", np.float32).reshape(n, 4),
                                  o.pos4)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "frame_vel4.f32", np.float32).reshape(n, 4),
                                  o.vel4)
    assert (tmp_path / "frame_pool.bin").read_bytes() == (tmp_path / "pool.bin").read_bytes()


def test_fluidhip_interaction_matches_oracle(tmp_path):
    """updateHand / deleteNear / raycastFluid / spawnParticles through the
    addon, replayed on the CPU restatement with the same inputs."""
    r = subprocess.run(["node", os.path.join(JS, "test", "interact_scene.mjs"), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    grid, eye = tuple(info["grid"]), info["eye"]
    pos3 = np.fromfile(tmp_path / "pos3.f32", np.float32).reshape(info["n0"], 3)
    o = oracle.OracleScene(grid=grid)
    o.reset(pos3)

    def frame(k):
        o.step(1)
        p4 = np.fromfile(tmp_path / f"pos4_{k}.f32", np.float32).reshape(-1, 4)
        v4 = np.fromfile(tmp_path / f"vel4_{k}.f32", np.float32).reshape(-1, 4)
        assert p4.shape[0] == o.n, f"count frame {k}"
        np.testing.assert_array_equal(p4, o.pos4, err_msg=f"pos frame {k}")
        np.testing.assert_array_equal(v4, o.vel4, err_msg=f"vel frame {k}")

    o.set_hand((6.0, 6.0, 6.0), (1.5, 0.0, -0.5), 2.5, 0.8, True, eye)
    frame(0)
    frame(1)
    left = o.delete_near((7.0, 8.0, -5.0), info["dirD"], 2.0)
    assert info["log"][2]["deleted"] == left < info["n0"]
    o.set_hand((8.0, 10.0, 7.0), (0.0, 1.0, 1.0), 3.0, 0.5, True, eye)
    frame(2)
    hit, t = o.raycast((6.0, 30.0, 6.0), (0.0, -1.0, 0.0), 1.0)
    assert hit and info["hit"] is not None
    np.testing.assert_allclose(info["hit"], [6.0, 30.0 - t, 6.0], rtol=0, atol=1e-5)
    sp = np.fromfile(tmp_path / "spawn_pos4.f32", np.float32).reshape(-1, 4)
    sv = np.fromfile(tmp_path / "spawn_vel4.f32", np.float32).reshape(-1, 4)
    assert sp.shape[0] == info["added"] == 500
    c = np.array(info["hit"]) - [0.0, 2.0, 0.0]
    assert np.all(np.linalg.norm(sp[:, :3] - c, axis=1) <= 1.5 + 1e-4)  # inside the spray ball
    o.append(sp[:, :3], sv[:, :3])
    o.set_hand((0, 0, 0), (0, 0, 0), 1.0, 1.0, False, eye)
    frame(3)
    frame(4)
