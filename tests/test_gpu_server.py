"""Serving form: requests admitted by HBM budget into one continuously
batched loop.  Each result must be bit-identical to stepping the scene alone
(and so to the CPU restatement), whether the request ran alone or in a mix
whose membership changed around it."""
import time

import numpy as np
import pytest

import mlsmpm
from mlsmpm import capi, scenes

pytestmark = pytest.mark.gpu

GB = 1 << 30


def scene_inputs(i):
    rng = np.random.default_rng(100 + i)
    grid = [(24, 28, 24), (40, 32, 36), (33, 45, 29), (56, 40, 48), (20, 20, 20)][i % 5]
    pos, vel = scenes.scaled_dam_break(grid, [9000, 60000, 30000, 150000, 4000][i % 5], rng,
                                       "two" if i % 2 else "dam")
    return grid, pos, vel


def alone(grid, pos, vel, steps):
    cfg = mlsmpm.default_config(*grid, len(pos), sort_period=1)
    with mlsmpm.Scene(cfg) as s:
        s.reset(pos, vel)
        s.step(steps)
        return s.read_particles()


def assert_same(a, b, where):
    for k in ("pos4", "vel4", "C9", "density"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{where}: {k}")


def test_server_results_equal_scenes_stepped_alone():
    reqs = []
    with mlsmpm.Server(hbm_budget=64 * GB, chunk_steps=8) as srv:
        for i in range(5):
            grid, pos, vel = scene_inputs(i)
            steps = [13, 40, 8, 27, 1][i]
            cfg = mlsmpm.default_config(*grid, len(pos))
            reqs.append((srv.submit(cfg, pos, steps, vel=vel), grid, pos, vel, steps))
        for rid, grid, pos, vel, steps in reqs:
            assert srv.wait(rid, 120.0) == capi.MPM_REQ_DONE
            info = srv.poll(rid)
            assert info["steps_done"] == steps and info["t_done"] >= info["t_admit"]
            assert_same(srv.result(rid), alone(grid, pos, vel, steps), f"request {rid}")
        st = srv.stats()
        assert st["completed"] == 5 and st["failed"] == 0 and st["active"] == 0
        assert st["bytes_resident"] == 0
        ms = srv.step_times()
        assert len(ms) > 0
        rec = srv.step_records()
        assert len(rec) == len(ms) and np.array_equal(rec["ms"], ms)
        assert np.all(rec["resident"] >= 1) and np.all(rec["resident_particles"] >= 4000)
        assert st["chunks"] > 0 and st["admitted"] == 5
        assert 0 < st["gpu_step_seconds"] <= st["busy_seconds"] * 1.5


def test_same_request_alone_and_in_a_changing_mix():
    grid, pos, vel = scene_inputs(1)
    steps = 45
    cfg = mlsmpm.default_config(*grid, len(pos))
    with mlsmpm.Server(hbm_budget=64 * GB, chunk_steps=4) as srv:
        rid = srv.submit(cfg, pos, steps, vel=vel)
        assert srv.wait(rid, 120.0) == capi.MPM_REQ_DONE
        solo = srv.result(rid)
    with mlsmpm.Server(hbm_budget=64 * GB, chunk_steps=4) as srv:
        others = []
        for i in (0, 2):
            g, p, v = scene_inputs(i)
            others.append(srv.submit(mlsmpm.default_config(*g, len(p)), p, 10 + 7 * i, vel=v))
        time.sleep(0.05)
        rid = srv.submit(cfg, pos, steps, vel=vel)
        for i in (3, 4):
            g, p, v = scene_inputs(i)
            others.append(srv.submit(mlsmpm.default_config(*g, len(p)), p, 30 + i, vel=v))
        assert srv.wait(rid, 120.0) == capi.MPM_REQ_DONE
        mixed = srv.result(rid)
        for o in others:
            assert srv.wait(o, 120.0) == capi.MPM_REQ_DONE
            srv.result(o)
    assert_same(solo, mixed, "alone vs mix")
    assert_same(solo, alone(grid, pos, vel, steps), "served vs plain")


@pytest.mark.parametrize("flags", [0, mlsmpm.MPM_BATCH_NO_GRAPH])
def test_graph_replays_and_eager_chunks_give_the_alone_results(flags):
    """A member set that lasts several chunks replays captured graphs (the
    default); forced eager steps replay none.  Either way every result equals
    the scene stepped alone, also for scenes whose memory came back to the
    server's arena from retired requests."""
    reqs = []
    with mlsmpm.Server(hbm_budget=64 * GB, chunk_steps=16, batch_flags=flags) as srv:
        for i in range(3):
            grid, pos, vel = scene_inputs(i)
            cfg = mlsmpm.default_config(*grid, len(pos))
            reqs.append((srv.submit(cfg, pos, 97, vel=vel), grid, pos, vel, 97))
        for rid, *_ in reqs[:3]:
            assert srv.wait(rid, 120.0) == capi.MPM_REQ_DONE
        # later requests reuse the retired scenes' arena blocks
        for i in (3, 4, 5):
            grid, pos, vel = scene_inputs(i)
            cfg = mlsmpm.default_config(*grid, len(pos))
            reqs.append((srv.submit(cfg, pos, 40 + i, vel=vel), grid, pos, vel, 40 + i))
        for rid, grid, pos, vel, steps in reqs:
            assert srv.wait(rid, 120.0) == capi.MPM_REQ_DONE
            assert_same(srv.result(rid), alone(grid, pos, vel, steps), f"request {rid}")
        st = srv.stats()
        assert st["completed"] == 6 and st["failed"] == 0
        assert len(srv.step_times()) == st["batched_steps"]
        if flags == mlsmpm.MPM_BATCH_NO_GRAPH:
            assert st["graph_steps"] == 0 and st["graph_captures"] == 0
        else:
            assert st["graph_steps"] >= 32 and st["graph_captures"] >= 1


def test_budget_admission_and_limits():
    grid, pos, vel = scene_inputs(3)
    cfg = mlsmpm.default_config(*grid, len(pos))
    foot = None
    with mlsmpm.Server(hbm_budget=8 * GB) as srv:
        rid = srv.submit(cfg, pos, 2, vel=vel)
        foot = srv.poll(rid)["footprint"]
        srv.wait(rid, 60.0)
        srv.result(rid)
    # room for one such scene at a time: the second waits for the first
    with mlsmpm.Server(hbm_budget=int(foot * 1.5), chunk_steps=4) as srv:
        a = srv.submit(cfg, pos, 12, vel=vel)
        b = srv.submit(cfg, pos, 12, vel=vel)
        assert srv.wait(a, 60.0) == capi.MPM_REQ_DONE
        assert srv.wait(b, 60.0) == capi.MPM_REQ_DONE
        ia, ib = srv.poll(a), srv.poll(b)
        assert ib["t_admit"] >= ia["t_done"]
        ra, rb = srv.result(a), srv.result(b)
        assert_same(ra, rb, "same request twice")
        # a scene larger than the whole budget is refused at submission
        big = mlsmpm.default_config(128, 128, 128, 4_000_000)
        with pytest.raises(mlsmpm.MpmError) as e:
            srv.submit(big, pos, 1)
        assert e.value.code == capi.MPM_ERR_NOMEM


def test_cancel_queued_and_running():
    grid, pos, vel = scene_inputs(3)
    cfg = mlsmpm.default_config(*grid, len(pos))
    with mlsmpm.Server(hbm_budget=64 * GB, max_active=1, chunk_steps=2) as srv:
        a = srv.submit(cfg, pos, 100000, vel=vel)
        b = srv.submit(cfg, pos, 5, vel=vel)
        srv.cancel(b)  # queued behind a (one slot)
        t0 = time.time()
        while srv.poll(a)["steps_done"] == 0 and time.time() - t0 < 30:
            time.sleep(0.01)
        srv.cancel(a)  # running
        assert srv.wait(a, 60.0) == capi.MPM_REQ_CANCELLED
        assert srv.wait(b, 60.0) == capi.MPM_REQ_CANCELLED
        assert 0 < srv.poll(a)["steps_done"] < 100000
        with pytest.raises(mlsmpm.MpmError):
            srv.result(a)
        c = srv.submit(cfg, pos, 3, vel=vel)
        assert srv.wait(c, 60.0) == capi.MPM_REQ_DONE
        assert_same(srv.result(c), alone(grid, pos, vel, 3), "after cancellations")


def test_invalid_request_fails_cleanly():
    grid = (24, 24, 24)
    pos = np.full((10, 3), 30.0, np.float32)  # outside the grid
    with mlsmpm.Server(hbm_budget=8 * GB) as srv:
        rid = srv.submit(mlsmpm.default_config(*grid, 10), pos, 3)
        assert srv.wait(rid, 60.0) == capi.MPM_REQ_FAILED
        with pytest.raises(mlsmpm.MpmError):
            srv.result(rid)
        g, p, v = scene_inputs(0)
        ok = srv.submit(mlsmpm.default_config(*g, len(p)), p, 2, vel=v)
        assert srv.wait(ok, 60.0) == capi.MPM_REQ_DONE
        srv.result(ok)


def test_submit_normalises_the_drift_margin():
    """Every served scene joins one batch, whose members share one drift
    margin: margin 0 and 2 are accepted and served together, any other is
    refused at submit, before anything runs."""
    with mlsmpm.Server(hbm_budget=16 * GB, chunk_steps=4) as srv:
        g, p, v = scene_inputs(0)
        with pytest.raises(mlsmpm.MpmError) as e:
            srv.submit(mlsmpm.default_config(*g, len(p), tile_margin=1), p, 5, vel=v)
        assert e.value.code == capi.MPM_ERR_UNSUPPORTED
        ids = []
        for i, margin in ((0, 2), (2, 0), (4, 2)):
            g, p, v = scene_inputs(i)
            cfg = mlsmpm.default_config(*g, len(p), tile_margin=margin)
            ids.append((srv.submit(cfg, p, 9 + i, vel=v), g, p, v, 9 + i))
        for rid, g, p, v, k in ids:
            assert srv.wait(rid, 60.0) == capi.MPM_REQ_DONE
            assert_same(srv.result(rid), alone(g, p, v, k), f"request {rid}")
        assert srv.stats()["failed"] == 0


def test_tile_table_applies_per_class_and_keeps_results():
    """A per-class work-item table given to the server sets tile_ppt of the
    requests that leave it 0 (class floor(log2 n) / 2); each request equals
    its scene stepped alone with that size."""
    from mlsmpm import autotune

    inputs = [scene_inputs(i) for i in (0, 3, 1)]
    cls = [autotune.size_class(len(p)) for _, p, _ in inputs]
    assert len(set(cls[:2])) == 2
    table = {cls[0]: 4, cls[1]: 16, cls[2]: 16}
    with mlsmpm.Server(hbm_budget=16 * GB, chunk_steps=4, tile_ppt_by_class=table) as srv:
        reqs = []
        for i, want in ((0, 4), (1, 16), (2, 8)):
            grid, pos, vel = inputs[i]
            ppt = 8 if i == 2 else 0                # an explicit size is kept
            cfg = mlsmpm.default_config(*grid, len(pos), tile_ppt=ppt)
            reqs.append((srv.submit(cfg, pos, 11, vel=vel), grid, pos, vel, want))
        for rid, grid, pos, vel, want in reqs:
            assert srv.poll(rid)["tile_ppt"] == want
            assert srv.wait(rid, 60.0) == capi.MPM_REQ_DONE
            cfg = mlsmpm.default_config(*grid, len(pos), sort_period=1, tile_ppt=want)
            with mlsmpm.Scene(cfg) as s:
                s.reset(pos, vel)
                s.step(11)
                assert_same(srv.result(rid), s.read_particles(), f"request {rid}")
    with pytest.raises(mlsmpm.MpmError):
        mlsmpm.Server(hbm_budget=GB, tile_ppt_by_class={3: 5})
