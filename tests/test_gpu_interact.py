"""GPU vs CPU-restatement parity of the interaction operations
(fluid-gpu.js:145-180, 479-638): the hand push inside the step (updateHand +
WGSL_APPLY_HAND), deleteNear's compaction, raycastFluid and the upload of
spawnParticles.  Same bar as the step parity: every grid and particle array
bit-exact after every call and every later step, in each execution form and
across captured multi-step replays."""
import numpy as np
import pytest

import mlsmpm
from mlsmpm import capi, scenes
from oracle import oracle
from test_gpu_parity import KEEP, MODES, check_equal

pytestmark = pytest.mark.gpu

FORMS = ["direct", "sorted8", "sorted1", "spill"]
GRID = (32, 32, 32)


def block(seed=1):
    return scenes.lattice_block((2, 2, 2), (14, 20, 14), 0.1, np.random.default_rng(seed))


def pair(pos, capacity=None, **kw):
    cfg = mlsmpm.default_config(*GRID, capacity or len(pos), **kw)
    s = mlsmpm.Scene(cfg)
    s.reset(pos)
    o = oracle.OracleScene(grid=GRID, seed=cfg.seed)
    o.reset(pos)
    return s, o


def both(s, o, fn, *a):
    getattr(s, fn)(*a)
    return getattr(o, fn)(*a)


@pytest.mark.parametrize("mode", FORMS)
def test_hand_push_every_step(mode):
    s, o = pair(block(), **MODES[mode])
    eye = (16.0, 40.0, 60.0)
    # through the fluid, then across empty bricks (cells with no mass get the
    # push and must read back as the dense update leaves them next step),
    # then switched off
    hands = [((8.0, 8.0, 8.0), (2.0, 0.5, -1.0), 3.0, 0.5),
             ((8.5, 8.0, 8.2), (2.0, 0.5, -1.0), 3.0, 0.5),
             ((25.0, 26.0, 24.0), (-1.0, -1.0, 0.0), 4.0, 1.0),
             ((26.0, 25.0, 20.0), (0.0, -2.0, 0.5), 2.5, 0.75)]
    for step in range(10):
        if step < 8:
            pos, vel, rad, k = hands[step // 2]
            both(s, o, "set_hand", pos, vel, rad, k, True, eye)
        elif step == 8:
            both(s, o, "set_hand", (0, 0, 0), (0, 0, 0), 1.0, 1.0, False, eye)
        s.step(1)
        o.step(1)
        check_equal(s, o, f"{mode} step {step}")


@pytest.mark.parametrize("mode", ["sorted8", "sorted1"])
def test_hand_push_in_captured_replays(mode):
    s, o = pair(block(2), **MODES[mode])
    eye = (-20.0, 10.0, 10.0)
    both(s, o, "set_hand", (9.0, 9.0, 9.0), (0.0, 1.0, 1.5), 3.5, 0.6, True, eye)
    s.step(40)
    o.step(40)
    check_equal(s, o, "40 steps, hand on")
    # a replay reads the push from device memory: a moved hand applies to
    # the next replay without a recapture
    both(s, o, "set_hand", (6.0, 12.0, 7.0), (1.0, 0.0, -1.0), 2.0, 1.2, True, eye)
    s.step(40)
    o.step(40)
    check_equal(s, o, "40 more steps, hand moved")
    both(s, o, "set_hand", (0, 0, 0), (0, 0, 0), 1.0, 1.0, False, eye)
    s.step(24)
    o.step(24)
    check_equal(s, o, "24 more steps, hand off")


def test_hand_outside_the_grid_and_degenerate():
    s, o = pair(block(3), **MODES["sorted8"])
    # the line misses the grid grown by the radius: empty box, no push
    both(s, o, "set_hand", (80.0, 80.0, 80.0), (1.0, 1.0, 1.0), 2.0, 1.0, True, (80.0, 90.0, 80.0))
    assert o.hand.bbox_max[0] < o.hand.bbox_min[0]
    s.step(3)
    o.step(3)
    check_equal(s, o, "hand outside")
    # eye on an axis-parallel line through the fluid (two zero direction
    # components)
    both(s, o, "set_hand", (8.0, 8.0, 8.0), (0.0, 0.0, 3.0), 2.0, 1.0, True, (8.0, 8.0, -5.0))
    s.step(3)
    o.step(3)
    check_equal(s, o, "axis-parallel hand")


@pytest.mark.parametrize("mode", FORMS)
def test_delete_near_then_steps(mode):
    s, o = pair(block(4), **MODES[mode])
    s.step(5)
    o.step(5)
    d = np.array([0.3, 0.2, 1.0]) / np.linalg.norm([0.3, 0.2, 1.0])
    n0 = o.n
    left = s.delete_near((8.0, 9.0, -10.0), d, 3.0)
    assert left == o.delete_near((8.0, 9.0, -10.0), d, 3.0)
    assert 0 < left < n0
    assert s.n == left
    p = s.read_particles()
    np.testing.assert_array_equal(p["pos4"], o.pos4)
    np.testing.assert_array_equal(p["vel4"], o.vel4)
    np.testing.assert_array_equal(p["C9"], o.affine)
    for step in range(4):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"{mode} step {step} after delete")
    # a second cut, then a captured run
    assert s.delete_near((0.0, 5.0, 5.0), (1.0, 0.0, 0.0), 2.0) == \
        o.delete_near((0.0, 5.0, 5.0), (1.0, 0.0, 0.0), 2.0)
    s.step(20)
    o.step(20)
    check_equal(s, o, f"{mode} 20 steps after the second delete")


def test_delete_everything_and_nothing():
    s, o = pair(block(5), **MODES["sorted8"])
    s.step(2)
    o.step(2)
    assert s.delete_near((50.0, 50.0, 50.0), (1.0, 0.0, 0.0), 1.0) == o.n  # misses
    check_equal(s, o, "nothing deleted")
    assert s.delete_near((0.0, 10.0, 10.0), (1.0, 0.0, 0.0), 100.0) == 0
    o.delete_near((0.0, 10.0, 10.0), (1.0, 0.0, 0.0), 100.0)
    s.step(3)
    o.step(3)
    check_equal(s, o, "everything deleted")
    assert s.delete_near((0.0, 10.0, 10.0), (1.0, 0.0, 0.0), 100.0) == 0


@pytest.mark.parametrize("mode", FORMS)
def test_append_then_steps(mode):
    pos = block(6)
    s, o = pair(pos, capacity=len(pos) + 3000, **MODES[mode])
    s.step(4)
    o.step(4)
    rng = np.random.default_rng(7)
    add = rng.uniform((18, 10, 18), (26, 20, 26), (2000, 3)).astype(np.float32)
    vel = rng.normal(0, 1.0, (2000, 3)).astype(np.float32)
    s.append(add, vel)
    o.append(add, vel)
    assert s.n == o.n == len(pos) + 2000
    for step in range(4):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"{mode} step {step} after append")
    with pytest.raises(mlsmpm.MpmError) as e:
        s.append(np.array([[0.5, 10.0, 10.0]], np.float32))
    assert e.value.code == capi.MPM_ERR_RANGE
    with pytest.raises(mlsmpm.MpmError) as e:
        s.append(add)  # 2000 more do not fit
    assert e.value.code == capi.MPM_ERR_CAPACITY
    assert s.n == o.n
    s.append(add[:1000] - 8.0)
    o.append(add[:1000] - 8.0)
    s.step(20)
    o.step(20)
    check_equal(s, o, f"{mode} 20 steps after the second append")


def test_delete_then_append_with_mass():
    pos = block(8)
    rng = np.random.default_rng(9)
    mass = rng.uniform(0.5, 2.0, len(pos)).astype(np.float32)
    cfg = mlsmpm.default_config(*GRID, len(pos) + 500, **MODES["sorted8"])
    s = mlsmpm.Scene(cfg)
    s.reset(pos, mass=mass)
    o = oracle.OracleScene(grid=GRID, seed=cfg.seed)
    o.reset(pos, mass=mass)
    s.step(3)
    o.step(3)
    both(s, o, "delete_near", (6.0, 6.0, 0.0), (0.0, 0.0, 1.0), 2.5)
    add = rng.uniform((3, 3, 3), (12, 12, 12), (500, 3)).astype(np.float32)
    both(s, o, "append", add)
    for step in range(10):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"step {step}")


def test_raycast_matches_oracle():
    s, o = pair(block(10), **MODES["sorted8"])
    s.step(6)
    o.step(6)
    rays = [((8.0, 40.0, 8.0), (0.0, -1.0, 0.0), 1.0),
            ((-5.0, 5.0, 7.0), (1.0, 0.0, 0.0), 0.5),
            ((30.0, 30.0, 30.0), tuple(np.array([-1.0, -1.2, -0.9]) / np.linalg.norm([1, 1.2, 0.9])),
             2.0),
            ((8.0, 8.0, 8.0), (0.0, 1.0, 0.0), 1.0),   # origin inside the fluid
            ((8.0, 40.0, 8.0), (0.0, 1.0, 0.0), 1.0),  # pointing away: miss
            ((8.0, 40.0, 8.0), (0.0, -1.0, 0.0), 0.0)]  # zero radius: miss
    hits = 0
    for r in rays:
        hg, tg = s.raycast(*r)
        ho, to = o.raycast(*r)
        assert hg == ho, r
        if ho:
            assert np.float32(tg) == np.float32(to), (r, tg, to)
            hits += 1
    assert hits >= 3


def test_interaction_errors_and_batches():
    pos = block(11)
    cfg = mlsmpm.default_config(*GRID, len(pos) + 400, flags=KEEP, sort_period=1)
    s = mlsmpm.Scene(cfg)
    s.reset(pos)
    o = oracle.OracleScene(grid=GRID, seed=cfg.seed)
    o.reset(pos)
    with pytest.raises(mlsmpm.MpmError) as e:
        s.set_hand((np.nan, 0, 0), (0, 0, 0), 1.0, 1.0, True, (0, 0, 0))
    assert e.value.code == capi.MPM_ERR_INVALID
    with pytest.raises(mlsmpm.MpmError) as e:
        s.delete_near((0, 0, 0), (1, 0, 0), -1.0)
    assert e.value.code == capi.MPM_ERR_INVALID
    with pytest.raises(mlsmpm.MpmError) as e:
        s.raycast((0, np.inf, 0), (1, 0, 0), 1.0)
    assert e.value.code == capi.MPM_ERR_INVALID
    s.set_hand((8, 8, 8), (1, 0, 0), 2.0, 1.0, True, (0, 0, 0))
    with mlsmpm.Batch() as b:
        with pytest.raises(mlsmpm.MpmError) as e:
            b.add(s)
        assert e.value.code == capi.MPM_ERR_UNSUPPORTED
        s.clear_hand()
        b.add(s)
        with pytest.raises(mlsmpm.MpmError) as e:
            s.set_hand((8, 8, 8), (1, 0, 0), 2.0, 1.0, True, (0, 0, 0))
        assert e.value.code == capi.MPM_ERR_STATE
        # delete and append on a member: the batch's tables follow the count
        b.step(2)
        o.step(2)
        n = both(s, o, "delete_near", (0.0, 8.0, 8.0), (1.0, 0.0, 0.0), 2.0)
        assert s.n == n and n < len(pos)
        b.step(3)
        o.step(3)
        check_equal(s, o, "batched after delete")
        add = np.random.default_rng(12).uniform((4, 4, 4), (12, 12, 12), (400, 3)).astype(np.float32)
        both(s, o, "append", add)
        b.step(20)
        o.step(20)
        check_equal(s, o, "batched after append")
        b.remove(s)


@pytest.mark.parametrize("mode", ["direct", "sorted8", "sorted1"])
def test_hand_with_walls_and_diffuse_pool(mode):
    """The hand push runs after the wall-aware grid update and before G2P and
    the diffuse step: with a wall and the pool on, every grid, particle and
    pool value stays bit-exact."""
    from test_gpu_parity import make_pair, wall_block
    pos = block(22)
    walls = wall_block(2)
    s, _ = make_pair(GRID, pos, D=8192, **MODES[mode])
    s.set_walls(walls)
    o = oracle.OracleScene(grid=GRID, diffuse_max=8192, seed=1, walls=walls)
    o.reset(pos)
    eye = (12.0, 30.0, -20.0)
    for step in range(8):
        both(s, o, "set_hand", (7.0 + 0.5 * step, 9.0, 8.0), (1.5, 2.0, 0.0), 3.0, 0.8, True,
             eye)
        s.step(1)
        o.step(1)
        check_equal(s, o, f"{mode} step {step}", diffuse=True)


@pytest.mark.parametrize("mode", ["direct", "sorted1"])
def test_delete_and_append_with_diffuse_pool(mode):
    """The pool's spawners are found by reset-order id each step: after a
    delete renumbers the survivors and an append adds ids, the pool, free
    list and top stay bit-exact."""
    pos = block(23)
    cfg = mlsmpm.default_config(*GRID, len(pos) + 600, diffuse_capacity=8192, seed=1,
                                **MODES[mode])
    s = mlsmpm.Scene(cfg)
    s.reset(pos)
    o = oracle.OracleScene(grid=GRID, diffuse_max=8192, seed=1)
    o.reset(pos)
    s.step(4)
    o.step(4)
    check_equal(s, o, "before", diffuse=True)
    both(s, o, "delete_near", (6.0, 0.0, 6.0), (0.0, 1.0, 0.0), 2.5)
    for step in range(3):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"{mode} step {step} after delete", diffuse=True)
    add = np.random.default_rng(24).uniform((3, 12, 3), (12, 18, 12), (600, 3)).astype(np.float32)
    both(s, o, "append", add, np.tile(np.float32([0.0, -2.0, 0.0]), (600, 1)))
    for step in range(4):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"{mode} step {step} after append", diffuse=True)
