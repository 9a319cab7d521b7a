"""The slab sub-allocator behind the served scenes' device memory and the
pinned result blocks (mls-mpm_amd/csrc/mpm_arena.h), checked on the host with
slabs from aligned_alloc: alignment, no overlap, best fit, merging back to
whole slabs, large requests, the cap and source failures
(tests/arena_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "mls-mpm_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_slab_arena(tmp_path):
    exe = str(tmp_path / "arena_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-pthread",
                    "-fsanitize=address,undefined", f"-I{CSRC}",
                    os.path.join(HERE, "arena_check.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ok"
