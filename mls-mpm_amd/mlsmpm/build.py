"""Build recipe of libmlsmpm.so (gfx950) and the Node-API addon.

Every HIP source is compiled with:
  -ffp-contract=off                        one rounding per operator, as the
                                           reference expressions are written
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE division and sqrt
  -fno-gpu-flush-denormals-to-zero         IEEE denormals
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                           no packed FP32 instructions
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))  # mls-mpm_amd/
REPO_ROOT = os.path.dirname(PKG_ROOT)
CSRC = os.path.join(PKG_ROOT, "csrc")
OBJ = os.path.join(PKG_ROOT, "_build")
LIBDIR = os.path.join(PKG_ROOT, "lib")
LIB = os.path.join(LIBDIR, "libmlsmpm.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["mpm_kernels.hip", "mpm_tiled.hip", "mpm_scene.hip", "mpm_capi.hip",
               "mpm_batch.hip", "mpm_server.hip", "mpm_interact.hip"]

HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
    f"-I{os.path.join(REPO_ROOT, 'include')}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature.
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{err}")


def build_library(verbose: bool = False) -> str:
    """Compile every HIP source for gfx950 and link libmlsmpm.so in-tree."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    cmds = []
    for s in srcs:
        o = os.path.join(OBJ, s.replace(".hip", ".o"))
        objs.append(o)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=min(len(cmds), 6)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_addon(verbose: bool = False) -> str | None:
    """Compile the Node-API addon (C, N-API v8) against libmlsmpm.so."""
    node_inc = "/usr/include/node"
    if not os.path.exists(os.path.join(node_inc, "node_api.h")):
        return None
    src = os.path.join(PKG_ROOT, "js", "addon", "mlsmpm_napi.c")
    if not os.path.exists(src):
        return None
    out = os.path.join(PKG_ROOT, "js", "addon", "mlsmpm.node")
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-Wextra", "-Wno-unused-parameter",
          "-DNODE_GYP_MODULE_NAME=mlsmpm",
          f"-I{node_inc}", f"-I{os.path.join(REPO_ROOT, 'include')}", "-o", out, src,
          f"-L{LIBDIR}", "-lmlsmpm", f"-Wl,-rpath,$ORIGIN/../../lib"])
    if verbose:
        print("built", out)
    return out


def build_all(verbose: bool = False) -> None:
    build_library(verbose)
    build_addon(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
