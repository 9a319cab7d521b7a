"""The serving form at its workload scale: a seeded Poisson stream of
requests at 40 per second (dam-break or two-block scenes, grid edges 32..128,
10k..2M particles log-uniform, 100..2000 steps) admitted by a 200 GB HBM
budget into one continuously batched loop, the arrivals shortened to two
seconds.  Every request must complete, and each request of a seeded sample
must equal its scene stepped alone for the same steps."""
import time

import numpy as np
import pytest

import mlsmpm
from mlsmpm import capi, scenes

pytestmark = pytest.mark.gpu

RATE, SECONDS, SAMPLE = 40.0, 2.0, 10


def test_poisson_stream_at_200gb_budget_matches_alone_runs():
    rng = np.random.default_rng(11)
    nreq = int(RATE * SECONDS)
    mix = scenes.scene_mix(5, nreq, kinds=("dam", "two"))
    steps = rng.integers(100, 2001, nreq)
    arrivals = np.cumsum(rng.exponential(1.0 / RATE, nreq))
    assert {k for *_, k in mix} == {"dam", "two"}
    pick = set(np.random.default_rng(13).choice(nreq, SAMPLE, replace=False).tolist())
    print(f"[poisson] {nreq} requests, {sum(len(p) * int(k) for (_, p, _, _), k in zip(mix, steps)):.3e}"
          " particle-steps", flush=True)
    results = {}
    with mlsmpm.Server(hbm_budget=200 << 30, chunk_steps=16) as srv:
        ids = []
        t0 = time.perf_counter()
        for (grid, pos, vel, _), k, ta in zip(mix, steps, arrivals):
            dt = ta - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(mlsmpm.default_config(*grid, len(pos)), pos, int(k), vel=vel))
        for i, rid in enumerate(ids):
            assert srv.wait(rid, 100.0) == capi.MPM_REQ_DONE, f"request {i}"
            info = srv.poll(rid)
            assert info["steps_done"] == int(steps[i])
            r = srv.result(rid)
            if i in pick:
                results[i] = r
        st = srv.stats()
        assert st["completed"] == nreq and st["failed"] == 0
        assert st["bytes_resident"] == 0
        # requests overlapped on the device: more than one resident on average
        assert st["resident_steps"] > 2 * st["batched_steps"]
    print(f"[poisson] all {nreq} done; mean resident "
          f"{st['resident_steps'] / max(1, st['batched_steps']):.1f}", flush=True)
    for i in sorted(results):
        grid, pos, vel, _ = mix[i]
        with mlsmpm.Scene(mlsmpm.default_config(*grid, len(pos), sort_period=1)) as s:
            s.reset(pos, vel)
            s.step(int(steps[i]))
            alone = s.read_particles()
        for k in ("pos4", "vel4", "C9", "density"):
            np.testing.assert_array_equal(results[i][k], alone[k], err_msg=f"request {i} {k}")
    print(f"[poisson] sample {sorted(results)} equal their alone runs", flush=True)
