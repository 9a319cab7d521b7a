"""Full-size workloads, checked through properties that need no CPU
restatement at this size.

Two-block collision (256^3 grid, 8,388,608 particles at +/- 2 cells/step,
2,097,152 diffuse slots): the diffuse pool conserves its slots (alive + free
= capacity, the free list holds exactly the dead slots), two identical scenes
stay bit-identical (determinism), and graph replay equals eager enqueue.

Dam break (256x128x128 grid, 4,199,136 particles): the sorted brick forms
equal the direct form bit for bit."""
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


# Full-size dam break (256x128x128 grid, 4,199,136 particles): the i32 grids
# are sums of integers, so they do not depend on the order in which particles
# add into them. The sorted brick forms (re-sort every step, or every 8 steps
# with drift margins and the overflow grid) must therefore leave every grid
# and every particle bit-identical to the direct form, whose particles add
# straight into the global grid in reset order.
DAM_STEPS = 24


def dam_run(flags, sort_period):
    grid, pos, vel = scenes.dam_break_large(seed=0)
    cfg = mlsmpm.default_config(*grid, len(pos), sort_period=sort_period,
                                flags=flags | mlsmpm.MPM_FLAG_KEEP_GRID)
    with mlsmpm.Scene(cfg) as s:
        s.reset(pos, vel)
        s.step(DAM_STEPS)
        return s.read_grid(), s.read_particles()


@pytest.fixture(scope="module")
def dam_direct():
    return dam_run(mlsmpm.MPM_FLAG_NO_LDS, 0)


@pytest.mark.parametrize("sort_period", [1, 8])
def test_dam_break_full_size_sorted_equals_direct(dam_direct, sort_period):
    gd, pd = dam_direct
    assert len(pd["pos4"]) == 4_199_136
    # the flow has started and fills a large part of the grid
    assert int(np.count_nonzero(gd["cell_mass"])) > 200_000
    assert float(np.abs(pd["vel4"][:, 1]).max()) > 1.0
    gs, ps = dam_run(0, sort_period)
    for k in ("cell_mass", "cell_mv", "cell_mass_f", "cell_v"):
        np.testing.assert_array_equal(gs[k], gd[k], err_msg=k)
    for k in ("pos4", "vel4", "C9", "density"):
        np.testing.assert_array_equal(ps[k], pd[k], err_msg=k)
