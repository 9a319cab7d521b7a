"""Batched form: many scenes of uneven sizes stepped with one persistent launch
per stage must give every scene exactly the bits of stepping it alone (and so
of the CPU restatement), through graph replay, eager steps, membership
changes and steps taken outside the batch."""
import numpy as np
import pytest

import mlsmpm
from mlsmpm import scenes
from oracle import oracle

pytestmark = pytest.mark.gpu

KEEP = mlsmpm.MPM_FLAG_KEEP_GRID

# (grid, block lo, block hi, jitter seed, mass)
SHAPES = [
    ((24, 24, 24), (2, 2, 2), (10, 14, 12), 1, None),
    ((40, 32, 28), (3, 2, 2.5), (30, 17, 22), 2, None),
    ((32, 48, 36), (2, 2, 2), (16, 30, 20), 3, "ramp"),
    ((56, 40, 48), (2, 2, 2), (40, 24, 30), 4, None),
    ((17, 21, 19), (2, 2, 2), (12, 12, 14), 5, None),
]


def make_scene(shape, margin=2, ppt=8, flags=KEEP):
    grid, lo, hi, seed, mass = shape
    pos = scenes.lattice_block(lo, hi, 0.1, np.random.default_rng(seed))
    m = None
    if mass == "ramp":
        m = (0.5 + (np.arange(len(pos)) % 7) * 0.25).astype(np.float32)
    cfg = mlsmpm.default_config(*grid, len(pos), flags=flags, sort_period=1, tile_margin=margin,
                                tile_ppt=ppt)
    s = mlsmpm.Scene(cfg)
    s.reset(pos, mass=m)
    return s, pos, m


def assert_same(a, b, where):
    pa, pb = a.read_particles(), b.read_particles()
    for k in pa:
        np.testing.assert_array_equal(pa[k], pb[k], err_msg=f"{where}: {k}")
    ga, gb = a.read_grid(), b.read_grid()
    for k in ga:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=f"{where}: grid {k}")
    assert a.counts() == b.counts(), where


@pytest.mark.parametrize("margin,ppt", [(2, 8), (1, 4)])
def test_batch_equals_scenes_stepped_alone(margin, ppt):
    shapes = SHAPES if margin == 2 else [s for s in SHAPES if s[4] is None]
    batched = [make_scene(sh, margin, ppt)[0] for sh in shapes]
    alone = [make_scene(sh, margin, ppt)[0] for sh in shapes]
    with mlsmpm.Batch() as b:
        for s in batched:
            b.add(s)
        assert len(b) == len(batched)
        b.step(1)    # first step after reset: the full sort, per scene
        b.step(37)   # two captured 16-step graphs and 5 eager batched steps
        b.step(3)
        b.sync()
        for s in alone:
            s.step(41)
        for i, (x, y) in enumerate(zip(batched, alone)):
            assert_same(x, y, f"scene {i} after 41 steps")
    for s in batched + alone:
        s.close()


def test_batch_member_matches_oracle():
    shape = SHAPES[1]
    s, pos, _ = make_scene(shape)
    o = oracle.OracleScene(grid=shape[0])
    o.reset(pos)
    other, _, _ = make_scene(SHAPES[0])
    with mlsmpm.Batch() as b:
        b.add(other)
        b.add(s)
        for step in range(6):
            b.step(1)
            o.step(1)
            g = s.read_grid()
            np.testing.assert_array_equal(g["cell_mass"], o.cell_mass, err_msg=f"mass {step}")
            np.testing.assert_array_equal(g["cell_mv"], o.cell_mv, err_msg=f"mv {step}")
            p = s.read_particles()
            np.testing.assert_array_equal(p["pos4"], o.pos4, err_msg=f"pos {step}")
            np.testing.assert_array_equal(p["vel4"], o.vel4, err_msg=f"vel {step}")
            np.testing.assert_array_equal(p["C9"], o.affine, err_msg=f"C {step}")
    s.close()
    other.close()


def test_batch_membership_changes_and_outside_steps():
    batched = [make_scene(sh)[0] for sh in SHAPES]
    alone = [make_scene(sh)[0] for sh in SHAPES]
    with mlsmpm.Batch(tuning={"graph_steps": 4, "item_blocks": 512, "resolve_blocks": 64,
                              "aux_blocks": 32}) as b:
        for s in batched[:3]:
            b.add(s)
        b.step(9)
        for s in alone[:3]:
            s.step(9)
        # a scene already stepped alone joins
        batched[3].step(5)
        alone[3].step(5)
        b.add(batched[3])
        # a member steps on its own between batched steps
        batched[1].step(2)
        alone[1].step(2)
        b.step(10)
        for i in range(4):
            alone[i].step(10)
        # one leaves, another joins fresh
        b.remove(batched[0])
        b.add(batched[4])
        b.step(7)
        for i in (1, 2, 3, 4):
            alone[i].step(7)
        b.sync()
        for i in range(5):
            assert_same(batched[i], alone[i], f"scene {i}")
        # a member reset inside the batch
        pos = scenes.lattice_block((2, 2, 2), (20, 12, 18), 0.1, np.random.default_rng(9))
        batched[2].reset(pos)
        alone[2].reset(pos)
        b.step(5)
        for i in (1, 2, 3, 4):
            alone[i].step(5)
        for i in (1, 2, 3, 4):
            assert_same(batched[i], alone[i], f"scene {i} after a reset in the batch")
    for s in batched + alone:
        s.close()


def test_batch_timing_records_every_step():
    batched = [make_scene(sh)[0] for sh in SHAPES[:3]]
    with mlsmpm.Batch() as b:
        for s in batched:
            b.add(s)
        b.step(2)
        b.enable_timing(True)
        b.step(20)
        st = b.step_times()
        assert len(st) == 20 and np.all(st > 0)
    for s in batched:
        s.close()


@pytest.mark.parametrize("flags", [0, mlsmpm.MPM_BATCH_GRAPH_REUSED])
def test_batch_timed_graph_replays_match_alone(flags):
    """Timed batches replay graphs that record each step's events as nodes:
    every step is timed, in order, through replays mixed with eager steps and
    a membership change, and every member keeps the bits of running alone."""
    batched = [make_scene(sh)[0] for sh in SHAPES[:3]]
    alone = [make_scene(sh)[0] for sh in SHAPES[:3]]
    with mlsmpm.Batch(tuning={"graph_steps": 4, "flags": flags}) as b:
        for s in batched:
            b.add(s)
        b.step(1)
        b.enable_timing(True)
        total = 0
        for n in (9, 8, 3, 12, 4):  # 4-step timed replays and eager steps
            b.step(n)
            total += n
        st = b.step_times()
        assert len(st) == total and np.all(st > 0)
        b.remove(batched[2])
        b.step(8)  # a new table: eager with GRAPH_REUSED, else replays
        b.step(8)
        st2 = b.step_times()
        assert len(st2) == total + 16 and np.all(st2 > 0)
        np.testing.assert_array_equal(st2[:total], st)
        b.sync()
    alone[2].step(1 + total)
    for s in alone[:2]:
        s.step(1 + total + 16)
    for i, (x, y) in enumerate(zip(batched, alone)):
        assert_same(x, y, f"scene {i}")
    for s in batched + alone:
        s.close()


def test_batch_rejects_unsupported_members():
    cfg = mlsmpm.default_config(24, 24, 24, 1000, sort_period=8)
    with mlsmpm.Scene(cfg) as s, mlsmpm.Batch() as b:
        with pytest.raises(mlsmpm.MpmError) as e:
            b.add(s)
        assert e.value.code == mlsmpm.capi.MPM_ERR_UNSUPPORTED
    cfg = mlsmpm.default_config(24, 24, 24, 1000, diffuse_capacity=100)
    with mlsmpm.Scene(cfg) as s, mlsmpm.Batch() as b:
        with pytest.raises(mlsmpm.MpmError):
            b.add(s)
