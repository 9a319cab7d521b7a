"""The JavaScript host (FluidHIP / DiffuseHIP over the Node-API addon) keeps
the reference's construction, tunables and fillBlock; without a GPU its
initGPU fails loudly and simFrame reports the error (no fallback)."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from mlsmpm import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS = os.path.join(ROOT, "mls-mpm_amd", "js")
ADDON = os.path.join(JS, "addon", "mlsmpm.node")

pytestmark = pytest.mark.skipif(not shutil.which("node") or not os.path.exists(ADDON),
                                reason="node or the built addon is missing")


def run_surface():
    r = subprocess.run(["node", os.path.join(JS, "test", "surface.mjs")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_fluidhip_surface_matches_fluidgpu():
    d = run_surface()
    # new FluidGPU(2, 2, 3, 0.0125, 400000) (main.js:21) -> 40 x 40 x 60
    assert d["grid"] == [40, 40, 60]
    df = d["defaults"]
    assert (df["DT"], df["GRAVITY"], df["REST_DENSITY"], df["STIFFNESS"]) == (0.3, -0.98, 8, 100)
    assert (df["EOS_POWER"], df["VISCOSITY"], df["HARD_MIN"]) == (1, 0.01, 2)
    assert df["HARD_MAX"] == [38, 38, 58] and df["SUBSTEPS"] == 1
    assert (df["WALL_MIN"], df["WALL_STIFFNESS"], df["LOOKAHEAD_K"]) == (3, 1, 2)
    # fillBlock(s, s, s, 1-s, 1-s, 0.4-s) count = the lattice of the same bounds
    s = 0.05
    pos = scenes.fill_block((40, 40, 60), s, s, s, 1 - s, 1 - s, 0.4 - s,
                            np.random.default_rng(0))
    assert d["filled"] == len(pos) and d["mass1"]
    assert d["diffuse"] == {"maxCount": 65536, "KE_THRESHOLD": 0.5, "LIFETIME_FOAM": 3,
                            "BUOYANCY": 1.2, "DRAG": 4, "seed": 1}


def test_fluidhip_fails_loudly_without_device():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a HIP device is present")
    except Exception:
        pass
    d = run_surface()
    assert d["init"]["code"] == "-4" and "device" in d["init"]["message"]
    assert d["simFrame"] is False and d["lastError"] == "-4"


def test_mobile_preset_upload_and_host_spawn():
    """The mobile preset's fill (main.js:19-21) is uploaded inside
    [1, grid - 1); spawnParticles before initGPU (fluid-gpu.js:145-180) fills
    the host arrays inside the hard bounds with unit mass, up to capacity."""
    m = run_surface()["mobile"]
    assert m["grid"] == [29, 29, 43] and m["filled"] == 80000
    assert m["moved"] > 0
    assert all(hi < g - 1 for hi, g in zip(m["max"], m["grid"]))
    assert all(lo >= 1 for lo in m["min"])
    assert m["spawned"] == 400 and m["active"] == 400
    assert m["spawnedInside"] and m["massOne"]
