"""Summaries of rocprofv3 result databases (SQLite, views `kernels` and
`counters_collection`) for the kernels of this library.

    python -m mlsmpm.rocprof_summary kt  <kt_results.db> [--last N]
        per-kernel dispatch count, mean/min/max/total duration (us), LDS,
        VGPRs, workgroups (--last N: only each kernel's last N dispatches,
        e.g. the timed region of a run)
    python -m mlsmpm.rocprof_summary pmc <pmc_results.db> [kernel-substring ...]
        mean value per dispatch of every collected counter
    python -m mlsmpm.rocprof_summary traffic --fetch F.db --write W.db \\
        --copy-fetch CF.db --copy-write CW.db --copy-bytes B --workload NAME \\
        --stage g2p --kernel k_g2p_brick [--out profiles/traffic.json]
        HBM bytes per launch of one kernel from FETCH_SIZE and WRITE_SIZE
        (KiB per dispatch), each scaled by known/measured for a plain copy
        kernel of B bytes read and B bytes written.
    python -m mlsmpm.rocprof_summary valu --db A.db --kernel K --stage S \
        --workload NAME [--out profiles/valu.json]
        the kernel's floor at the exact-FP32 VALU issue rate from
        SQ_INSTS_VALU per dispatch.
"""
from __future__ import annotations

import argparse
import collections
import json
import re
import sqlite3
import sys


def short_name(name: str) -> str:
    """Kernel name without namespace, arguments and template noise, keeping
    the template arguments (k_g2p_brick<0,true>)."""
    base = re.sub(r"\(.*", "", name)
    base = base.replace("mpm::", "").replace("void ", "")
    return base.strip()[:60]


def kernel_table(path: str, last: int = 0):
    db = sqlite3.connect(path)
    d = collections.defaultdict(list)
    info = {}
    for name, dur, lds, vgpr, agpr, sgpr, gx, wgx in db.execute(
            "select name, duration, lds_size, vgpr_count, accum_vgpr_count, "
            "sgpr_count, grid_x, workgroup_x from kernels order by start"):
        k = short_name(name)
        d[k].append(dur)
        info[k] = (lds, vgpr, agpr, sgpr, gx // max(wgx, 1))
    if last:
        d = {k: v[-last:] for k, v in d.items()}
    rows = []
    for k, v in sorted(d.items(), key=lambda kv: -sum(kv[1])):
        rows.append((k, len(v), sum(v) / len(v) / 1e3, min(v) / 1e3, max(v) / 1e3, sum(v) / 1e3,
                     *info[k]))
    return rows


def print_kernels(path: str, last: int = 0) -> None:
    print(f"{'kernel':60s} {'calls':>5s} {'mean_us':>9s} {'min_us':>9s} {'max_us':>9s} "
          f"{'total_us':>10s} {'lds':>6s} {'vgpr':>4s} {'agpr':>4s} {'sgpr':>4s} {'wgs':>6s}")
    for r in kernel_table(path, last):
        print(f"{r[0]:60s} {r[1]:5d} {r[2]:9.2f} {r[3]:9.2f} {r[4]:9.2f} {r[5]:10.1f} "
              f"{r[6]:6d} {r[7]:4d} {r[8]:4d} {r[9]:4d} {r[10]:6d}")


def counter_means(path: str, match=None):
    db = sqlite3.connect(path)
    d = collections.defaultdict(lambda: collections.defaultdict(list))
    for name, cname, val in db.execute(
            "select kernel_name, counter_name, value from counters_collection"):
        k = short_name(name)
        if match and not any(m in k for m in match):
            continue
        d[k][cname].append(val)
    return {k: {c: (sum(v) / len(v), len(v)) for c, v in cs.items()} for k, cs in d.items()}


def print_counters(path: str, match=None) -> None:
    for k, cs in sorted(counter_means(path, match).items()):
        vals = ", ".join(f"{c}={m:.6g}" for c, (m, _) in sorted(cs.items()))
        n = max(c[1] for c in cs.values())
        print(f"{k:60s} dispatches={n}  {vals}")


def _one(path: str, kernel: str, counter: str) -> float:
    means = counter_means(path, [kernel])
    vals = [cs[counter][0] for k, cs in means.items() if counter in cs]
    if not vals:
        raise SystemExit(f"{counter} for {kernel} not found in {path}")
    if len(vals) > 1:
        raise SystemExit(f"{kernel} matches several kernels in {path}: {list(means)}")
    return vals[0]


def traffic(a) -> dict:
    kib = 1024.0
    fetch = _one(a.fetch, a.kernel, "FETCH_SIZE") * kib
    write = _one(a.write, a.kernel, "WRITE_SIZE") * kib
    cf = _one(a.copy_fetch, a.copy_kernel, "FETCH_SIZE") * kib
    cw = _one(a.copy_write, a.copy_kernel, "WRITE_SIZE") * kib
    sf, sw = a.copy_bytes / cf, a.copy_bytes / cw
    ent = {
        "kernel": a.kernel,
        "fetch_bytes_raw": fetch,
        "write_bytes_raw": write,
        "copy_bytes": a.copy_bytes,
        "copy_fetch_raw": cf,
        "copy_write_raw": cw,
        "scale_fetch": sf,
        "scale_write": sw,
        "bytes_per_launch": fetch * sf + write * sw,
    }
    if a.out:
        _merge(a.out, a.workload, a.stage, ent)
    return ent


# Exact FP32 retires one wave64 VALU instruction per SIMD every 4 clocks:
# the VALU floor of a launch is INSTS_VALU * 4 / (SIMDs * clock).
MI355X_SIMDS = 1024
MI355X_CLOCK_HZ = 2.4e9


def valu(a) -> dict:
    insts = _one(a.db, a.kernel, "SQ_INSTS_VALU")
    ent = {
        "kernel": a.kernel,
        "insts_valu": insts,
        "simds": MI355X_SIMDS,
        "clock_hz": MI355X_CLOCK_HZ,
        "floor_ms": insts * 4.0 / (MI355X_SIMDS * MI355X_CLOCK_HZ) * 1e3,
    }
    if a.out:
        _merge(a.out, a.workload, a.stage, ent)
    return ent


def _merge(path: str, workload: str, stage: str, ent: dict) -> None:
    try:
        with open(path) as f:
            out = json.load(f)
    except FileNotFoundError:
        out = {}
    out.setdefault(workload, {})[stage] = ent
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kt")
    k.add_argument("db", nargs="+")
    k.add_argument("--last", type=int, default=0)
    p = sub.add_parser("pmc")
    p.add_argument("db")
    p.add_argument("match", nargs="*")
    t = sub.add_parser("traffic")
    for f in ("fetch", "write", "copy_fetch", "copy_write"):
        t.add_argument("--" + f.replace("_", "-"), dest=f, required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--copy-kernel", default="elementwise")
    t.add_argument("--kernel", required=True)
    t.add_argument("--stage", required=True)
    t.add_argument("--workload", required=True)
    t.add_argument("--out")
    v = sub.add_parser("valu")
    v.add_argument("--db", required=True)
    v.add_argument("--kernel", required=True)
    v.add_argument("--stage", required=True)
    v.add_argument("--workload", required=True)
    v.add_argument("--out")
    a = ap.parse_args(argv)
    if a.cmd == "kt":
        for db in a.db:
            print(f"== {db}")
            print_kernels(db, a.last)
    elif a.cmd == "pmc":
        print_counters(a.db, a.match or None)
    elif a.cmd == "valu":
        print(json.dumps(valu(a), indent=1))
    else:
        print(json.dumps(traffic(a), indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
