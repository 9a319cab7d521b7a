"""The allocation-free entry point (mpm_raycast_noalloc) against the entry
point it varies (mpm_raycast), through the C ABI: every ray's hit flag and t
bit-identical, on two calls in a row, the second with fewer rays (and more
particles, after appending), so a stale hit word of the first call would
show.  Its size limit is refused with a clear error."""
import numpy as np
import pytest

import mlsmpm
from mlsmpm import capi, scenes

pytestmark = pytest.mark.gpu


def rays(rng, k, grid):
    g = np.array(grid, np.float64)
    o = rng.uniform(0.1, 0.9, (k, 3)) * g
    o[: k // 2, 1] = g[1] + 5.0                 # half from above the fluid, looking down
    d = rng.normal(size=(k, 3))
    d[: k // 2] = [0.0, -1.0, 0.0] + 0.2 * rng.normal(size=(k // 2, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(0.2, 2.5, k)
    r[-3:] = 0.0                                 # radius 0: never a hit
    return o.astype(np.float32), d.astype(np.float32), r.astype(np.float32)


def check_against_single(s, o, d, r):
    hit, t = s.raycast_noalloc(o, d, r)
    single = [s.raycast(o[i], d[i], float(r[i])) for i in range(len(o))]
    want_hit = np.array([h for h, _ in single])
    want_t = np.array([x for _, x in single], np.float32)
    np.testing.assert_array_equal(hit, want_hit)
    np.testing.assert_array_equal(t.view(np.uint32), want_t.view(np.uint32))
    return int(hit.sum()), int((~hit).sum())


def test_noalloc_raycast_matches_raycast_on_two_calls():
    grid = (48, 40, 44)
    rng = np.random.default_rng(3)
    pos = scenes.lattice_block((2, 2, 2), (30, 24, 28), 0.1, rng)
    with mlsmpm.Scene(mlsmpm.default_config(*grid, len(pos) + 5000)) as s:
        s.reset(pos)
        s.step(5)
        o, d, r = rays(rng, 700, grid)
        hits, misses = check_against_single(s, o, d, r)
        assert hits > 50 and misses > 3
        # second call: fewer rays, after more steps and appended particles;
        # the first rays that hit now point away, so old hit words must be cleared
        s.append(rng.uniform(20, 40, (5000, 3)).astype(np.float32) * [1, 0.9, 1])
        s.step(3)
        o2, d2, r2 = rays(rng, 61, grid)
        o2[:10], d2[:10], r2[:10] = o[:10], -d[:10], 0.5
        hits2, misses2 = check_against_single(s, o2, d2, r2)
        assert hits2 > 0 and misses2 > 10


def test_noalloc_raycast_limits_and_empty_calls():
    grid = (24, 24, 24)
    pos = scenes.lattice_block((2, 2, 2), (10, 10, 10), 0.1, np.random.default_rng(1))
    with mlsmpm.Scene(mlsmpm.default_config(*grid, len(pos))) as s:
        s.reset(pos)
        hit, t = s.raycast_noalloc(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0))
        assert len(hit) == 0
        k = capi.MPM_RAYCAST_NOALLOC_MAX
        o = np.tile([[5.0, 30.0, 5.0]], (k + 1, 1))
        dd = np.tile([[0.0, -1.0, 0.0]], (k + 1, 1))
        with pytest.raises(mlsmpm.MpmError) as e:
            s.raycast_noalloc(o, dd, 1.0)
        assert e.value.code == capi.MPM_ERR_INVALID
        assert str(k) in str(e.value)
        hit, t = s.raycast_noalloc(o[:k], dd[:k], 1.0)          # exactly the limit
        h1, t1 = s.raycast(o[0], dd[0], 1.0)
        assert hit.all() == h1 and np.all(t == np.float32(t1))
        with pytest.raises(mlsmpm.MpmError):
            s.raycast_noalloc(o[:2], dd[:2], float("nan"))
        s.reset(np.zeros((0, 3), np.float32))                   # no particles: all miss
        hit, t = s.raycast_noalloc(o[:5], dd[:5], 1.0)
        assert not hit.any() and np.all(t == 0)
