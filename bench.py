#!/usr/bin/env python3
"""MLS-MPM fluid step throughput on one MI355X.

Workload (default): 3-D dam break on a 256x128x128 grid, 4,199,136 FP32
particles on a 0.5-cell lattice (x in [2,104), y in [2,43.5), z in [2,126),
seeded jitter +/-0.1 cell), i32 fixed-point grids, reference physics
constants, no diffuse particles.  Particles are re-laid out by brick inside
every G2P (--sort-period 1, the fastest period; --sort-period 8 re-sorts every
8 steps as the workload is specified, with identical grids).  A "step" is one
whole MLS-MPM step (P2G mass, P2G momentum, grid update, G2P with the
re-layout) over all particles, state resident in HBM.  The timed run is
--steps steps (default 1000, the workload's run length) from the initial
state: --warmup untimed steps run, the scene is reset and the graphs the timed
steps replay are captured (mpm_prepare).

Prints one JSON line:
  value            particle-steps/s over the timed region (wall clock,
                   stream-synchronised on both sides, no events recorded)
  roofline         the dominant kernel's algorithmic bytes per launch over its
                   mean launch time, from HIP events recorded on the scene's
                   stream around that kernel alone in a second timed region of
                   the same steps from the same reset (less the events' own
                   cost, an empty interval between two marks of the same
                   steps), against the 8 TB/s HBM
                   peak (active cells: the mean of ten samples over that run); `traffic` is
                   the HBM bytes per launch that rocprofv3 FETCH_SIZE +
                   WRITE_SIZE read for that kernel (profiles/traffic.json),
                   scaled by the same counters' reading for a plain copy;
                   `valu` is that kernel's floor at the exact-FP32 VALU issue
                   rate (SQ_INSTS_VALU, profiles/valu.json) over the same
                   launch time: the bound that holds it, as HBM does not
  cpu_baseline     the CPU restatement of the reference step on this host's
                   cores, on a bounded sample of the same workload
Extra keys: per-stage mean ms, whole-step roofline fraction, p50/p99 step
latency (from the event-timed region), the renderer handoff (device export
kernel against HBM, pinned frames alone and overlapped with the steps).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "mls-mpm_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

# Algorithmic bytes (information moved, 12 B per vec3) per particle and per
# active cell for each timed stage of the step.  "g2p", "p2g_mom" and
# "update" bracket one kernel each (KERNEL names it).
#   P2G mass     : pos 12 + mass 4 per particle; mass write 4 per cell
#   P2G momentum : pos 12 + vel 12 + C 36 + mass 4 + density 4 per particle;
#                  mass read 4 + momentum write 12 per cell
#   grid update  : read mass+momentum 16, write velocity 12, clear 16 per cell
#   G2P          : pos 12 read, pos 12 + vel 12 + C 36 written per particle;
#                  velocity read 12 per cell
ALG_BYTES = {
    "p2g_mass": (16, 4),
    "p2g_mom": (68, 16),
    "update": (0, 44),
    "g2p": (72, 12),
}
STAGES = ["p2g_mass", "p2g_mom", "update", "g2p", "diffuse", "sort", "step", "relayout"]
KERNEL = {"g2p": "k_g2p_brick", "p2g_mom": "k_p2g_mom_brick", "update": "k_resolve_update",
          "p2g_mass": "k_p2g_mass_brick + k_brick_list + k_resolve_mass"}


def workload(name: str, seed: int):
    from mlsmpm import scenes

    if name == "dam-large":
        grid, pos, vel = scenes.dam_break_large(seed)
        desc = "dam-break 256x128x128, column x[2,104) y[2,43.5) z[2,126), 0.5-cell lattice"
    elif name == "dam-small":
        grid, pos, vel = scenes.dam_break_small(seed)
        desc = "dam-break 64^3, block [2,30]x[2,40]x[2,30], 0.5-cell lattice"
    elif name == "two-block":
        grid, pos, vel = scenes.two_block_collision(seed)
        desc = "two-block collision 256^3, blocks at +/-2 cells/step"
    else:
        raise SystemExit(f"unknown workload {name}")
    return grid, pos, vel, desc


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(grid, pos, vel, seconds: float, max_threads: int = 0,
                 max_steps: int = 200) -> dict:
    """The oracle (C restatement of the reference kernels, OpenMP) on a
    bounded number of steps of the same scene, on every host core this
    process may use: the cores of its affinity mask, capped at the host's
    share for one GPU when OMP_NUM_THREADS states one (max_threads > 0 caps
    it further)."""
    from oracle import oracle

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(cores, share) if share > 0 else cores
    if max_threads > 0:
        threads = min(threads, max_threads)
    threads = max(1, threads)
    o = oracle.OracleScene(grid=tuple(grid))
    o.reset(pos, vel)
    t0 = time.perf_counter()
    o.step(1, threads=threads)
    t1 = time.perf_counter()
    per = t1 - t0
    # about `seconds` of CPU work in all (bounded at max_steps steps)
    more = max(0, min(max_steps - 1, int(seconds / max(per, 1e-6)) - 1))
    if more:
        o.step(more, threads=threads)
    t2 = time.perf_counter()
    steps = 1 + more
    n = len(pos)
    return {
        "value": n * steps / (t2 - t0),
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {steps} steps of the same scene ({n} particles), "
                  f"{threads} OpenMP threads, {t2 - t0:.1f} s",
        "affinity_cores": cores,
        "omp_num_threads": share or None,
        "cpu_model": cpu_model(),
    }


def batch_bench(args) -> dict:
    """Batched form against sequential plain steps of the same scene mix:
    `batch_scenes` dam-break scenes (grid edges log-uniform 32..128 per axis,
    10k..2M particles log-uniform, capped at 8 per cell over half the
    interior), every scene re-sorting every step.  Batched: all scenes in one
    mpm_batch (one persistent launch per stage, captured graphs, work-item
    sizes from the shipped autotuner table, persistent block counts tuned on
    the batch).  Sequential: each scene stepped alone (its own captured
    graphs), one scene after another, once with the single-scene default
    work-item size and once with the sizes the autotuner chose for scenes
    alone.  Every form runs the same `batch_steps` steps from the same
    initial state (a warm-up captures the graphs, then every scene is reset);
    rates are particle-steps/s."""
    import mlsmpm
    from mlsmpm import autotune, scenes

    mix = scenes.scene_mix(args.seed + 3, args.batch_scenes)
    ntot = sum(len(p) for _, p, _, _ in mix)
    print(f"[bench] batch: {len(mix)} scenes, {ntot} particles", file=sys.stderr, flush=True)
    # work-item sizes from the shipped autotuner table: per particle-count
    # class, tuned for batched steps (the batch) and for scenes alone
    tb_batch = autotune.load_table(key="batch_ppt_by_class")
    tb_alone = autotune.load_table(key="alone_ppt_by_class")

    def make(table):
        out = []
        for grid, pos, vel, _ in mix:
            cfg = autotune.apply_tiles(
                mlsmpm.default_config(*grid, len(pos), sort_period=1, device=0), table)
            s = mlsmpm.Scene(cfg)
            s.reset(pos, vel)
            out.append(s)
        return out

    def sequential(scs):
        for s in scs:
            s.step(W)
        for s in scs:
            s.sync()
        for s, (_, pos, vel, _) in zip(scs, mix):
            s.reset(pos, vel)
        t0 = time.perf_counter()
        for s in scs:
            s.step(K)
            s.sync()
        return time.perf_counter() - t0

    sc = make(tb_batch)

    def reset_all():
        for s, (_, pos, vel, _) in zip(sc, mix):
            s.reset(pos, vel)

    K, W = args.batch_steps, 17
    b = mlsmpm.Batch()
    for s in sc:
        b.add(s)
    tune = autotune.tune_batch(b)
    reset_all()
    b.step(W)
    b.sync()
    reset_all()
    t0 = time.perf_counter()
    b.step(K)
    b.sync()
    tb = time.perf_counter() - t0
    b.enable_timing(True)
    b.step(min(K, 32))
    lat = b.step_times()
    b.close()
    for s in sc:
        s.close()
    # sequential plain steps: the single-scene default work-item size, and the
    # sizes tuned for scenes alone
    sc = make(None)
    tseq = sequential(sc)
    for s in sc:
        s.close()
    sc = make(tb_alone)
    tseq_t = sequential(sc)
    for s in sc:
        s.close()
    batched = ntot * K / tb
    seq = ntot * K / tseq
    seq_t = ntot * K / tseq_t
    return {
        "scenes": len(mix),
        "particles": ntot,
        "steps": K,
        "value": batched,
        "unit": "particle-steps/s",
        "ms_per_step": tb / K * 1e3,
        "sequential_plain": seq,
        "sequential_ms_per_step": tseq / K * 1e3,
        "vs_sequential": batched / seq,
        "sequential_plain_tuned": seq_t,
        "vs_sequential_tuned": batched / seq_t,
        "tile_ppt_by_class": {"batched": tb_batch, "sequential_tuned": tb_alone},
        "step_latency_ms": {"p50": float(np.percentile(lat, 50)),
                            "p99": float(np.percentile(lat, 99))},
        "tuning": {"item_blocks": tune.item_blocks, "resolve_blocks": tune.resolve_blocks,
                   "aux_blocks": tune.aux_blocks, "graph_steps": tune.graph_steps},
    }


def serve_stream(args, mix, steps, rate: float) -> dict:
    """One seeded Poisson stream of the requests `mix` (with `steps` each) at
    `rate` per second into a fresh server; waits for every result."""
    import mlsmpm
    from mlsmpm import autotune

    nreq = len(mix)
    arrivals = np.cumsum(np.random.default_rng(args.seed + 17).exponential(1.0 / rate, nreq))
    tb_batch = autotune.load_table(key="batch_ppt_by_class")
    srv = mlsmpm.Server(hbm_budget=200 << 30, chunk_steps=args.serve_chunk,
                        tile_ppt_by_class=tb_batch, batch_flags=args.serve_batch_flags)
    # every request's client fetches its result as soon as it is done: a scan
    # over the outstanding requests between arrivals and after the last one,
    # not one wait after another in arrival order (which would hold the early
    # finishers' results behind slower requests)
    from mlsmpm import capi

    final = (capi.MPM_REQ_DONE, capi.MPM_REQ_FAILED, capi.MPM_REQ_CANCELLED)
    lat = []
    pending = []

    def drain():
        nonlocal pending
        rest = []
        for rid in pending:
            info = srv.poll(rid)
            if info["state"] in final:
                lat.append(info["t_done"] - info["t_submit"])
                srv.result(rid)
            else:
                rest.append(rid)
        done_any = len(rest) < len(pending)
        pending = rest
        return done_any

    t0 = time.perf_counter()
    for (grid, pos, vel, _), k, ta in zip(mix, steps, arrivals):
        while True:
            dt = ta - (time.perf_counter() - t0)
            if dt <= 0:
                break
            if not drain():
                time.sleep(min(dt, 0.001))
        pending.append(srv.submit(mlsmpm.default_config(*grid, len(pos)), pos, int(k), vel=vel))
    deadline = time.perf_counter() + 600.0
    while pending and time.perf_counter() < deadline:
        if not drain():
            time.sleep(0.001)
    wall = time.perf_counter() - t0
    st = srv.stats()
    step_ms = srv.step_times()
    rec = srv.step_records()
    srv.close()
    return {
        "rate_per_s": rate,
        "arrival_seconds": float(arrivals[-1]),
        "completed": st["completed"],
        "unfinished": len(pending),
        "particle_steps": st["particle_steps"],
        "value": st["particle_steps"] / st["busy_seconds"],
        "wall_rate": st["particle_steps"] / wall,
        "wall_seconds": wall,
        # the device's rate over its own batched-step time: what the loop
        # reaches if its host work never leaves the device idle
        "device_rate": st["particle_steps"] / max(st["gpu_step_seconds"], 1e-9),
        "mean_resident": st["resident_steps"] / max(1, st["batched_steps"]),
        "step_latency_ms": {"p50": float(np.percentile(step_ms, 50)),
                            "p99": float(np.percentile(step_ms, 99))},
        "request_latency_s": {"p50": float(np.percentile(lat, 50)),
                              "p99": float(np.percentile(lat, 99))},
        "batched_steps": st["batched_steps"],
        "chunks": st["chunks"],
        # batched steps replayed from captured graphs (member sets that last
        # more than one chunk), and the captures they needed
        "graph_steps": st["graph_steps"],
        "graph_captures": st["graph_captures"],
        # where the loop thread's time went, and how much of the busy time the
        # device spent in batched steps
        "loop_seconds": {k: st[k + "_seconds"]
                         for k in ("enqueue", "admit", "wait", "retire", "enroll",
                                   "admit_create", "admit_load", "retire_read",
                                   "retire_release")},
        "gpu_step_seconds": st["gpu_step_seconds"],
        "gpu_busy_frac": st["gpu_step_seconds"] / max(st["busy_seconds"], 1e-9),
        "tail": tail_attribution(rec),
    }


def serve_bench(args) -> dict:
    """Serving form: a seeded Poisson stream of requests (dam-break or
    two-block scene, grid edges 32..128, 10k..2M particles log-uniform, steps
    uniform in 100..2000), `serve_rate * serve_seconds` requests admitted by an
    HBM budget of 200 GB into one continuously batched loop.  The stream runs
    at `serve_rate` per second (the headline: latency at that load), then the
    same requests at four times the rate (`saturated`: arrivals faster than
    the loop finishes them, so the resident mix grows to what the budget
    admits).  Each reports aggregate particle-steps/s (over the loop's busy
    time, over the wall time from first arrival to last completion, and over
    the device's batched-step time), p50/p99 batched-step latency and p50/p99
    request latency; `vs_plain` divides the busy-time rate by the same
    requests' plain single-scene rate (each scene's time measured alone on a
    seeded sample of 48 requests: its full step count from the initial
    state)."""
    import mlsmpm
    from mlsmpm import autotune, scenes

    rng = np.random.default_rng(args.seed + 11)
    nreq = max(1, int(round(args.serve_rate * args.serve_seconds)))
    mix = scenes.scene_mix(args.seed + 5, nreq, kinds=("dam", "two"))
    steps = rng.integers(100, 2001, nreq)
    work = sum(len(p) * int(k) for (_, p, _, _), k in zip(mix, steps))
    print(f"[bench] serve: {nreq} requests, {work:.3e} particle-steps", file=sys.stderr,
          flush=True)
    out = serve_stream(args, mix, steps, args.serve_rate)
    sat = serve_stream(args, mix, steps, 4.0 * args.serve_rate)
    tb_alone = autotune.load_table(key="alone_ppt_by_class")
    # plain form: a seeded sample of the requests, each scene alone running
    # its request's steps from the initial state (after a warm-up that
    # captures its graphs and a reset)
    sample = np.random.default_rng(args.seed + 13).choice(nreq, min(nreq, 48), replace=False)

    def plain(table):
        plain_s, plain_work = 0.0, 0
        for i in sample:
            grid, pos, vel, _ = mix[i]
            cfg = autotune.apply_tiles(mlsmpm.default_config(*grid, len(pos), sort_period=1),
                                       table)
            with mlsmpm.Scene(cfg) as s:
                s.reset(pos, vel)
                s.step(17)
                s.reset(pos, vel)
                s.prepare()
                ta = time.perf_counter()
                s.step(int(steps[i]))
                s.sync()
                plain_s += time.perf_counter() - ta
                plain_work += len(pos) * int(steps[i])
        return plain_work / plain_s

    plain_rate = plain(None)
    plain_rate_t = plain(tb_alone)
    for o in (out, sat):
        o["vs_plain"] = o["value"] / plain_rate
        o["vs_plain_tuned"] = o["value"] / plain_rate_t
        o["device_vs_plain"] = o["device_rate"] / plain_rate
    res = {"requests": nreq, "hbm_budget_gb": 200, "unit": "particle-steps/s",
           "offered_rate": work / float(out["arrival_seconds"]),
           "plain_rate": plain_rate, "plain_sample": int(len(sample)),
           "plain_rate_tuned": plain_rate_t}
    res.update(out)
    res["saturated"] = sat
    return res


def tail_attribution(rec) -> dict:
    """What the slowest batched steps had resident: the steps at or above the
    99th percentile of latency against the rest (resident particles and
    scenes, admissions running beside them), and the latency per million
    resident particles, whose spread is what the mix itself does not explain."""
    if len(rec) == 0:
        return {}
    ms = rec["ms"].astype(np.float64)
    part = rec["resident_particles"].astype(np.float64)
    slow = ms >= np.percentile(ms, 99)
    per_m = ms / np.maximum(part / 1e6, 1e-9)

    def summ(m):
        return {"steps": int(m.sum()), "ms_mean": float(ms[m].mean()),
                "resident_particles_mean": float(part[m].mean()),
                "resident_scenes_mean": float(rec["resident"][m].mean()),
                "admitting_frac": float((rec["admitting"][m] > 0).mean())}

    return {"p99_steps": summ(slow), "other_steps": summ(~slow),
            "ms_per_million_particles": {"p50": float(np.percentile(per_m, 50)),
                                         "p99": float(np.percentile(per_m, 99))}}


def handoff_bench(s, n: int, steps_ms: float) -> dict:
    """Renderer handoff at the workload's size: the device export kernel
    (HIP events on the scene's stream around back-to-back exports) and the
    pinned frames (export + device-to-host copy), alone and every 8 steps
    while the steps run."""
    from mlsmpm import capi
    from mlsmpm.hipmem import DeviceBuffer, StreamTimer

    bufs = [DeviceBuffer(n * 16), DeviceBuffer(n * 16), DeviceBuffer(n * 4)]
    tm = StreamTimer(s.stream())
    out = {}
    for name, reset in (("export_storage_order", False), ("export_reset_order", True)):
        s.export_device(*(b.ptr for b in bufs), reset_order=reset)
        K = 50
        tm.start()
        for _ in range(K):
            s.export_device(*(b.ptr for b in bufs), reset_order=reset)
        ms = tm.stop() / K
        # 28 B read (pos, vel, density; + 4 B id in reset order), 36 B written
        alg = (64 + (4 if reset else 0)) * n
        out[name] = {"ms": ms, "alg_bytes": alg, "GB/s": alg / (ms * 1e-3) / 1e9,
                     "frac": alg / (ms * 1e-3) / HBM_PEAK}
    tm.close()
    for b in bufs:
        b.close()
    what = capi.MPM_EXPORT_POS | capi.MPM_EXPORT_VEL | capi.MPM_EXPORT_DENSITY
    s.frame_poll(s.frame_begin(what), wait=True, copy=False)
    K = 10
    t0 = time.perf_counter()
    for _ in range(K):
        s.frame_poll(s.frame_begin(what), wait=True, copy=False)
    frame_ms = (time.perf_counter() - t0) / K * 1e3
    out["pinned_frame_ms"] = frame_ms
    out["pinned_GB/s"] = 36 * n / (frame_ms * 1e-3) / 1e9
    # a frame every 8 steps: begin, step, then the previous frame is waited for
    s.sync()
    F, rounds = 8, 25
    t0 = time.perf_counter()
    prev = None
    for _ in range(rounds):
        t = s.frame_begin(what)
        s.step(F)
        if prev is not None:
            s.frame_poll(prev, wait=True, copy=False)
        prev = t
    s.frame_poll(prev, wait=True, copy=False)
    s.sync()
    per_step = (time.perf_counter() - t0) / (F * rounds) * 1e3
    out["steps_with_frame_every_8"] = {"ms_per_step": per_step,
                                       "vs_steps_alone": per_step / steps_ms}
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--workload", default="dam-large", choices=["dam-large", "dam-small", "two-block"])
    ap.add_argument("--diffuse", type=int, default=-1,
                    help="diffuse slots (default: 2,097,152 for two-block, else 0)")
    ap.add_argument("--lifetime-mode", type=int, default=0,
                    help="foam lifetime: 0 reference, 1 uniform [0.5, 3] s")
    ap.add_argument("--sort-period", type=int, default=1)
    ap.add_argument("--flags", type=int, default=0)
    ap.add_argument("--tile-ppt", type=int, default=0)
    ap.add_argument("--tile-margin", type=int, default=0)
    ap.add_argument("--tile-threads", type=int, default=0)
    ap.add_argument("--g2p-threads", type=int, default=0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="cap on the CPU baseline's host threads (0: every core available, "
                         "at most OMP_NUM_THREADS)")
    ap.add_argument("--cpu-max-steps", type=int, default=200)
    ap.add_argument("--no-events", action="store_true",
                    help="skip the event-timed region (no stage times, no roofline)")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=40.0)
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-chunk", type=int, default=16)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-batch-flags", type=int, default=0,
                    help="batch tuning flags of the server (0: graphs for lasting member sets, "
                         "1: every step eager)")
    ap.add_argument("--no-handoff", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 3 device-to-device copies of BYTES each (the plain copy that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes) and exit")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the step is single-device "
              "(multi-GPU sharding is a design note, not built). Run with --gpus 1.",
              file=sys.stderr)
        return 2

    if args.copy_calibration:
        from mlsmpm.hipmem import DeviceBuffer

        nb = args.copy_calibration
        src, dst = DeviceBuffer(nb, fill=1), DeviceBuffer(nb)
        for _ in range(3):
            dst.copy_from(src)
        src.close()
        dst.close()
        print(json.dumps({"copy_bytes": nb, "copies": 3}), flush=True)
        return 0

    import mlsmpm

    grid, pos, vel, desc = workload(args.workload, args.seed)
    n = len(pos)
    D = args.diffuse if args.diffuse >= 0 else (2_097_152 if args.workload == "two-block" else 0)
    cfg = mlsmpm.default_config(*grid, n, sort_period=args.sort_period, flags=args.flags, device=0,
                                tile_ppt=args.tile_ppt, tile_margin=args.tile_margin,
                                tile_threads=args.tile_threads, g2p_threads=args.g2p_threads,
                                diffuse_capacity=D, lifetime_mode=args.lifetime_mode)
    s = mlsmpm.Scene(cfg)
    s.reset(pos, vel)
    print(f"[bench] {desc}: {n} particles, warmup {args.warmup}", file=sys.stderr, flush=True)
    # Warm-up captures the step graphs; each region then runs the workload's
    # steps from the initial state (a reset to the same particle count keeps
    # the graphs).
    s.step(args.warmup)
    s.sync()

    # Region 1: throughput, wall clock only.  The graphs the run replays are
    # captured beforehand (mpm_prepare), whatever the warm-up length.
    s.reset(pos, vel)
    s.enable_timing(False)
    s.prepare()
    s.sync()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    # Region 2: the same steps from the same start with HIP events around
    # every stage; active cells are counted every tenth of the run.
    stage = np.zeros(len(STAGES), np.float32)
    step_ms = np.zeros(0, np.float32)
    events_wall = None
    g_samples, a_samples = [], []
    if not args.no_events:
        s.reset(pos, vel)
        s.enable_timing(True)
        s.prepare()
        chunk = max(1, args.steps // 10)
        done = 0
        t_ev = 0.0
        while done < args.steps:
            k = min(chunk, args.steps - done)
            t2 = time.perf_counter()
            s.step(k)
            s.sync()
            t_ev += time.perf_counter() - t2
            done += k
            g_samples.append(int(np.count_nonzero(s.read_grid(ints=False)["cell_mass_f"])))
            if D:
                a_samples.append(s.diffuse_alive())
        events_wall = t_ev / args.steps * 1e3
        stage = s.stage_times()
        step_ms = s.step_times()
        s.enable_timing(False)
    else:
        g_samples.append(int(np.count_nonzero(s.read_grid(ints=False)["cell_mass_f"])))
    tstats = s.tile_stats()
    g_act = float(np.mean(g_samples))
    ms = {k: float(v) for k, v in zip(STAGES, stage)}
    # dominant kernel and its roofline
    dom = max(("g2p", "p2g_mom", "update"), key=lambda k: ms[k])
    bp, bc = ALG_BYTES[dom]
    alg = bp * n + bc * g_act
    # The kernel's launch time: its event interval less the events' own cost,
    # which the "relayout" interval measures (two marks with nothing between
    # them, in the same captured steps).
    launch_ms = ms[dom] - max(0.0, min(ms["relayout"], 0.5 * ms[dom]))
    achieved = alg / (launch_ms * 1e-3) if launch_ms > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        ent = tr.get(args.workload, {}).get(dom)
        if ent:
            traffic = ent.get("bytes_per_launch")
    # the same kernel's VALU floor (exact FP32 issue rate) from the committed
    # counter summary: what bounds it, beside the HBM fraction
    valu = None
    vpath = os.path.join(ROOT, "profiles", "valu.json")
    if os.path.exists(vpath):
        with open(vpath) as f:
            ve = json.load(f).get(args.workload, {}).get(dom)
        if ve and launch_ms > 0:
            valu = {"floor_ms": ve["floor_ms"], "frac": ve["floor_ms"] / launch_ms,
                    "insts_per_launch": ve["insts_valu"]}
    d_alive = s.diffuse_alive() if D else 0
    d_alive_mean = float(np.mean(a_samples)) if a_samples else float(d_alive)
    # diffuse term: density read per particle, alive flag per slot, 72 B per
    # alive record at the run's mean alive count (spawn and kill traffic left out)
    step_alg = 156 * n + 76 * g_act + (4 * n + 4 * D + 72 * d_alive_mean if D else 0)
    step_ms_mean = elapsed / args.steps * 1e3
    line = {
        "metric": "particle-steps/s + % HBM roofline; batched/served vs plain; p99 step latency",
        "value": n * args.steps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms_mean,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "particles": n,
            "grid": list(grid),
            "grid_dtype": "i32 fixed point (x1e5)",
            "sort_period": args.sort_period,
            "diffuse": D,
            "diffuse_alive_end": d_alive,
            "diffuse_alive_mean": d_alive_mean,
            "lifetime_mode": args.lifetime_mode,
            "active_cells_mean": g_act,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": KERNEL[dom],
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "alg_bytes_per_launch": alg,
            "launch_ms": launch_ms,
            "event_interval_ms": ms[dom],
            "valu": valu,
        },
        "step_roofline_frac": step_alg / (step_ms_mean * 1e-3) / HBM_PEAK,
        "stage_ms": ms,
        "tile_stats": tstats,
        "events_ms_per_step": float(ms["step"]),
        "events_region_wall_ms_per_step": events_wall,
        "step_latency_ms": {
            "p50": float(np.percentile(step_ms, 50)) if len(step_ms) else None,
            "p99": float(np.percentile(step_ms, 99)) if len(step_ms) else None,
        },
    }
    if not args.no_handoff:
        print("[bench] renderer handoff ...", file=sys.stderr, flush=True)
        line["handoff"] = handoff_bench(s, n, step_ms_mean)
    s.close()
    if not args.no_batch:
        line["batched"] = batch_bench(args)
    if not args.no_serve:
        line["served"] = serve_bench(args)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(grid, pos, vel, args.cpu_seconds, args.cpu_threads,
                                            args.cpu_max_steps)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
