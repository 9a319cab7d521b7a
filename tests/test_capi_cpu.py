"""C ABI checks that need no GPU: the library loads, exports every entry
point include/mlsmpm.h declares, fills the reference defaults, validates
configurations and fails loudly (no CPU fallback) when no device exists."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import mlsmpm
from mlsmpm import capi
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "mlsmpm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = capi.load()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/mlsmpm.h but not exported"
    # and the Python binding declares exactly the header's entry points
    assert sorted(capi.declared_symbols()) == names


def test_config_struct_size_matches_header():
    # 8 u32 + 22 f32 + seed + 4 tile fields + lifetime_mode + 1 reserved = 37 words
    assert C.sizeof(capi.Config) == 37 * 4


def test_defaults_are_the_reference_constants():
    c = mlsmpm.default_config(256, 128, 128, 1000)
    p = O.default_params(256, 128, 128)
    for f in ("dt", "gravity", "rest_density", "stiffness", "eos_power", "viscosity", "hard_min",
              "hard_max_x", "hard_max_y", "hard_max_z", "wall_min", "wall_stiffness",
              "lookahead_k"):
        assert np.float32(getattr(c, f)) == np.float32(getattr(p, f)), f
    g = O.default_gen()
    for f in ("ke_threshold", "spawn_rate_k", "density_spray_max", "density_bubble_min",
              "lifetime_foam", "pressure_threshold", "crest_dot_threshold"):
        assert np.float32(getattr(c, f)) == np.float32(getattr(g, f)), f
    assert c.seed == 1 and c.buoyancy == np.float32(1.2) and c.drag == 4.0
    assert c.sort_period == 1


def test_config_from_aspect_matches_fluidgpu_constructor():
    # new FluidGPU(2, 2, 3, 0.0125, 400000) -> 40 x 40 x 60 (main.js:21)
    c = mlsmpm.config_from_aspect(2, 2, 3, 0.0125, 400000)
    assert (c.grid_x, c.grid_y, c.grid_z) == (40, 40, 60)
    assert c.hard_max_x == 38.0 and c.hard_max_z == 58.0
    # mobile preset radius 0.0175 -> ceil(2/0.07) = 29, ceil(3/0.07) = 43
    c = mlsmpm.config_from_aspect(2, 2, 3, 0.0175, 80000)
    assert (c.grid_x, c.grid_y, c.grid_z) == (29, 29, 43)


@pytest.mark.parametrize("field,value,code", [
    ("grid_x", 4, capi.MPM_ERR_INVALID),
    ("eos_power", 7.0, capi.MPM_ERR_UNSUPPORTED),
    ("hard_min", 0.5, capi.MPM_ERR_INVALID),
    ("hard_max_x", 63.5, capi.MPM_ERR_INVALID),
    ("dt", float("nan"), capi.MPM_ERR_INVALID),
    ("tile_ppt", 12, capi.MPM_ERR_INVALID),
    ("tile_margin", 3, capi.MPM_ERR_INVALID),
    ("tile_threads", 128, capi.MPM_ERR_INVALID),
])
def test_invalid_configs_are_rejected_before_any_device_work(field, value, code):
    c = mlsmpm.default_config(64, 64, 64, 100)
    setattr(c, field, value)
    with pytest.raises(mlsmpm.MpmError) as e:
        mlsmpm.Scene(c)
    assert e.value.code == code


def test_create_without_device_fails_loudly():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a HIP device is present")
    with pytest.raises(mlsmpm.MpmError) as e:
        mlsmpm.Scene(mlsmpm.default_config(32, 32, 32, 100))
    assert e.value.code == capi.MPM_ERR_HIP


def test_null_handles_are_rejected():
    L = capi.load()
    assert L.mpm_step(None, 1) == capi.MPM_ERR_INVALID
    assert b"NULL" in L.mpm_last_error()
    assert L.mpm_destroy(None) == capi.MPM_OK


def test_batch_and_server_fail_loudly_without_device():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a HIP device is present")
    with pytest.raises(mlsmpm.MpmError) as e:
        mlsmpm.Batch()
    assert e.value.code == capi.MPM_ERR_HIP
    with pytest.raises(mlsmpm.MpmError) as e:
        mlsmpm.Server(hbm_budget=1 << 30)
    assert e.value.code == capi.MPM_ERR_HIP


def test_abi_struct_sizes():
    assert C.sizeof(capi.BatchTuning) == 8 * 4
    assert C.sizeof(capi.ServerConfig) == 56
    assert C.sizeof(capi.ServerStats) == 4 * 8 + 2 * 4 + 4 * 8 + 2 * 8 + 6 * 8 + 4 * 8 + 4 * 8
    assert C.sizeof(capi.RequestInfo) == 4 * 4 + 8 + 3 * 8 + 2 * 4


# ctypes structure -> C type of include/mlsmpm.h
ABI_STRUCTS = {"Config": "mpm_config", "BatchTuning": "mpm_batch_tuning",
               "ServerConfig": "mpm_server_config", "ServerStats": "mpm_server_stats",
               "RequestInfo": "mpm_request_info", "Hand": "mpm_hand", "Export": "mpm_export",
               "Frame": "mpm_frame", "StepRecord": "mpm_step_record"}


def test_ctypes_structs_match_the_header(tmp_path):
    """Every field offset and size of the ctypes structures equals what a C
    compiler makes of include/mlsmpm.h."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc is missing")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mlsmpm.h"', 'int main(void) {']
    want = []
    for py, c in ABI_STRUCTS.items():
        S = getattr(capi, py)
        lines.append(f'  printf("%zu\\n", sizeof({c}));')
        want.append(C.sizeof(S))
        for f in S._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({c}, {f[0]}));')
            want.append(getattr(S, f[0]).offset)
    lines.append('  return 0;\n}')
    src, exe = tmp_path / "abi.c", tmp_path / "abi"
    src.write_text("\n".join(lines))
    r = subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert got == want


def test_interaction_and_handoff_reject_null_handles():
    L = capi.load()
    v3 = (C.c_float * 3)(0, 0, 1)
    t = C.c_uint64()
    for rc in (L.mpm_set_hand(None, None), L.mpm_delete_near(None, v3, v3, 1.0, None),
               L.mpm_raycast(None, v3, v3, 1.0, None, None),
               L.mpm_append_particles(None, 1, v3, None), L.mpm_export_device(None, None, 0),
               L.mpm_frame_begin(None, capi.MPM_EXPORT_POS, C.byref(t)),
               L.mpm_frame_poll(None, 1, 0, None, None)):
        assert rc == capi.MPM_ERR_INVALID
    assert b"NULL" in L.mpm_last_error()
