"""Full-size two-block collision (256^3 grid, 8,388,608 particles at
+/- 2 cells/step, 2,097,152 diffuse slots), checked through properties that
do not need the CPU restatement at this size: the diffuse pool conserves its
slots (alive + free = capacity, the free list holds exactly the dead slots),
two identical scenes stay bit-identical (determinism), and graph replay
equals eager enqueue."""
import numpy as np
import pytest

import mlsmpm
from mlsmpm import scenes

pytestmark = pytest.mark.gpu

D = 2_097_152


def scene(flags=0, lifetime_mode=0):
    grid, pos, vel = scenes.two_block_collision(seed=0)
    cfg = mlsmpm.default_config(*grid, len(pos), diffuse_capacity=D, sort_period=1,
                                flags=flags, lifetime_mode=lifetime_mode)
    s = mlsmpm.Scene(cfg)
    s.reset(pos, vel)
    return s


def check_pool(s):
    rec, fl, top = s.read_diffuse()
    alive = rec["alive"] == 1
    assert int(alive.sum()) + top == D
    free = np.sort(fl[:top])
    np.testing.assert_array_equal(free, np.flatnonzero(~alive).astype(np.uint32))
    return int(alive.sum())


@pytest.mark.parametrize("lifetime_mode", [0, 1])
def test_two_block_collision_full_size(lifetime_mode):
    a = scene(lifetime_mode=lifetime_mode)
    b = scene(flags=mlsmpm.MPM_FLAG_NO_GRAPH, lifetime_mode=lifetime_mode)
    a.step(1)
    a.step(32)           # two captured 16-step graphs
    for _ in range(33):
        b.step(1)
    alive = check_pool(a)
    assert alive > 0     # the collision sprays
    check_pool(b)
    pa, pb = a.read_particles(), b.read_particles()
    for k in pa:
        np.testing.assert_array_equal(pa[k], pb[k], err_msg=k)
    ra, fa, ta = a.read_diffuse()
    rb, fb, tb = b.read_diffuse()
    assert ta == tb
    np.testing.assert_array_equal(fa, fb)
    for f in ("pos", "vel", "ptype", "alive", "lifetime"):
        np.testing.assert_array_equal(ra[f], rb[f], err_msg=f)
    a.close()
    b.close()
