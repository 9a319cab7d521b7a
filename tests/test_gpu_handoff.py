"""Renderer handoff (fluid-renderer.js:232, 256, 315-316, 850-856): the
device export into caller buffers and the pinned, double-buffered frames
must hold exactly the state that mpm_read_particles / mpm_read_diffuse give
at the moment the export was enqueued, in the renderer's layouts."""
import time

import numpy as np
import pytest

import mlsmpm
from mlsmpm import capi, scenes
from mlsmpm.hipmem import DeviceBuffer
from mlsmpm.scene import DREC_DTYPE

pytestmark = pytest.mark.gpu

GRID = (32, 32, 32)
D = 4096


def make(sort_period=8, D=D):
    pos = scenes.lattice_block((2, 2, 2), (14, 20, 14), 0.1, np.random.default_rng(21))
    cfg = mlsmpm.default_config(*GRID, len(pos), diffuse_capacity=D, sort_period=sort_period)
    s = mlsmpm.Scene(cfg)
    s.reset(pos)
    return s


def rows(p4, v4, d):
    """The particle rows sorted, to compare two orders of the same particles."""
    a = np.concatenate([p4, v4, d[:, None]], axis=1)
    return a[np.lexsort(a.T[::-1])]


def check_frame(f, ref, rec, top, where):
    np.testing.assert_array_equal(f["pos4"], ref["pos4"], err_msg=f"pos {where}")
    np.testing.assert_array_equal(f["vel4"], ref["vel4"], err_msg=f"vel {where}")
    np.testing.assert_array_equal(f["density"], ref["density"], err_msg=f"density {where}")
    assert f["top"] == top, where
    assert f["diffuse"].tobytes() == rec.tobytes(), f"diffuse {where}"


@pytest.mark.parametrize("sort_period", [0, 8])
def test_export_device_matches_read_back(sort_period):
    s = make(sort_period)
    s.step(30)  # a captured run with the sorted form: storage order != reset order
    ref = s.read_particles()
    rec, _, top = s.read_diffuse()
    assert D - top > 0, "no diffuse particles alive"
    n = s.n
    bufs = [DeviceBuffer(n * 16, fill=0xff), DeviceBuffer(n * 16, fill=0xff),
            DeviceBuffer(n * 4, fill=0xff), DeviceBuffer(D * 48, fill=0xff)]
    s.export_device(*(b.ptr for b in bufs), reset_order=True)
    s.sync()
    p4 = bufs[0].to_numpy(np.float32, (n, 4))
    v4 = bufs[1].to_numpy(np.float32, (n, 4))
    dn = bufs[2].to_numpy(np.float32, (n,))
    np.testing.assert_array_equal(p4, ref["pos4"])
    np.testing.assert_array_equal(v4, ref["vel4"])
    np.testing.assert_array_equal(dn, ref["density"])
    assert bufs[3].to_numpy(np.uint8).tobytes() == rec.tobytes()
    # storage order: the same rows, pos / vel / density kept together
    s.export_device(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, reset_order=False)
    s.sync()
    q4 = bufs[0].to_numpy(np.float32, (n, 4))
    w4 = bufs[1].to_numpy(np.float32, (n, 4))
    en = bufs[2].to_numpy(np.float32, (n,))
    np.testing.assert_array_equal(rows(q4, w4, en), rows(ref["pos4"], ref["vel4"], ref["density"]))
    if sort_period:
        assert not np.array_equal(q4, p4), "the sorted form should store another order"
    assert np.all(q4[:, 3] == 0) and np.all(w4[:, 3] == 0)
    for b in bufs:
        b.close()
    s.close()


def test_frames_hold_the_state_at_begin_while_steps_run():
    s = make()
    s.step(3)
    ref_a = s.read_particles()
    rec_a, _, top_a = s.read_diffuse()
    t1 = s.frame_begin(capi.MPM_EXPORT_ALL, reset_order=True)
    s.step(40)  # enqueued behind the frame; runs while its copy does
    t2 = s.frame_begin(capi.MPM_EXPORT_ALL, reset_order=True)
    f1 = s.frame_poll(t1, wait=True)
    check_frame(f1, ref_a, rec_a, top_a, "frame 1")
    assert f1["steps"] == 3 and f1["n"] == s.n
    ref_b = s.read_particles()
    rec_b, _, top_b = s.read_diffuse()
    f2 = s.frame_poll(t2, wait=True)
    check_frame(f2, ref_b, rec_b, top_b, "frame 2")
    assert f2["steps"] == 43
    assert not np.array_equal(f1["pos4"], f2["pos4"])
    # a third frame reuses the first one's memory
    t3 = s.frame_begin(capi.MPM_EXPORT_POS)
    with pytest.raises(mlsmpm.MpmError) as e:
        s.frame_poll(t1)
    assert e.value.code == capi.MPM_ERR_NOT_FOUND
    # polling without waiting comes back ready within a bounded time
    deadline = time.time() + 20.0
    f3 = None
    while f3 is None and time.time() < deadline:
        f3 = s.frame_poll(t3, wait=False)
        if f3 is None:
            time.sleep(0.001)
    assert f3 is not None
    assert f3["vel4"] is None and f3["density"] is None and f3["diffuse"] is None
    np.testing.assert_array_equal(rows(f3["pos4"], np.zeros_like(f3["pos4"]),
                                       np.zeros(f3["n"], np.float32)),
                                  rows(ref_b["pos4"], np.zeros_like(ref_b["pos4"]),
                                       np.zeros(f3["n"], np.float32)))
    s.close()


def test_frames_follow_count_changes_and_errors():
    s = make(D=0)
    s.step(2)
    t1 = s.frame_begin(capi.MPM_EXPORT_POS | capi.MPM_EXPORT_DENSITY, reset_order=True)
    left = s.delete_near((0.0, 8.0, 8.0), (1.0, 0.0, 0.0), 2.0)
    t2 = s.frame_begin(capi.MPM_EXPORT_POS | capi.MPM_EXPORT_DENSITY, reset_order=True)
    f1, f2 = s.frame_poll(t1), s.frame_poll(t2)
    assert f1["n"] > left == f2["n"]
    np.testing.assert_array_equal(f2["pos4"], s.read_particles()["pos4"])
    with pytest.raises(mlsmpm.MpmError) as e:
        s.frame_begin(capi.MPM_EXPORT_DIFFUSE)
    assert e.value.code == capi.MPM_ERR_STATE
    with pytest.raises(mlsmpm.MpmError) as e:
        s.frame_begin(0)
    assert e.value.code == capi.MPM_ERR_INVALID
    with pytest.raises(mlsmpm.MpmError) as e:
        s.frame_poll(12345)
    assert e.value.code == capi.MPM_ERR_NOT_FOUND
    s.close()


def test_export_from_a_batch_member():
    s = make(sort_period=1, D=0)
    other = make(sort_period=1, D=0)
    with mlsmpm.Batch() as b:
        b.add(s)
        b.add(other)
        b.step(20)
        t = s.frame_begin(capi.MPM_EXPORT_POS | capi.MPM_EXPORT_VEL, reset_order=True)
        b.step(5)
        f = s.frame_poll(t)
        b.sync()
        assert f["steps"] == 20
        alone = make(sort_period=1, D=0)
        alone.step(20)
        ref = alone.read_particles()
        np.testing.assert_array_equal(f["pos4"], ref["pos4"])
        np.testing.assert_array_equal(f["vel4"], ref["vel4"])
        alone.close()
        b.remove(s)
        b.remove(other)
    s.close()
    other.close()
