"""GPU vs CPU-restatement parity of the whole step, through the C ABI.

Bar: i32 grids bit-exact per step (integer sums are order independent given
identical particles); particle position, velocity, C and density, the f32
grids and the diffuse pool bit-exact as well, because every kernel evaluates
the reference expressions in the same order with one rounding per operator.
(The stated tolerance for particle state is 1e-5 relative; these tests hold
the stronger bit-exact bar.)
"""
import numpy as np
import pytest

import mlsmpm
from mlsmpm import scenes
from oracle import oracle

pytestmark = pytest.mark.gpu

KEEP = mlsmpm.MPM_FLAG_KEEP_GRID
NO_LDS = mlsmpm.MPM_FLAG_NO_LDS

# Execution forms that must all give the same bits:
#   direct  : per-particle global atomics, particles in reset order
#   tiled   : LDS windows, particles in reset order
#   sorted8 : LDS windows, counting sort by cell every 8 steps (the default)
#   sorted1 : sort every step, 4 particles per thread (1024-particle items)
#   sorted32: sort every 3 steps, 16 particles per thread, drift margin 2
#   spill   : sort every 8 steps, margin 1: many particles drift out of their
#             LDS window and take the overflow-grid path
MODES = {
    "direct": dict(flags=KEEP | NO_LDS, sort_period=0),
    "tiled": dict(flags=KEEP, sort_period=0),
    "sorted8": dict(flags=KEEP, sort_period=8),
    "sorted1": dict(flags=KEEP, sort_period=1, tile_ppt=4),
    "sorted32": dict(flags=KEEP, sort_period=3, tile_ppt=16, tile_margin=2),
    "spill": dict(flags=KEEP, sort_period=8, tile_ppt=8, tile_margin=1),
}


def make_pair(grid, pos, vel=None, D=0, flags=KEEP, sort_period=8, seed=1, **kw):
    cfg = mlsmpm.default_config(*grid, len(pos), diffuse_capacity=D, flags=flags,
                                sort_period=sort_period, seed=seed, **kw)
    s = mlsmpm.Scene(cfg)
    s.reset(pos, vel)
    o = oracle.OracleScene(grid=grid, diffuse_max=D, seed=seed)
    o.reset(pos, vel)
    return s, o


def check_equal(s, o, where, ints=True, diffuse=False):
    g = s.read_grid(ints=ints)
    if ints:
        np.testing.assert_array_equal(g["cell_mass"], o.cell_mass, err_msg=f"cell_mass {where}")
        np.testing.assert_array_equal(g["cell_mv"], o.cell_mv, err_msg=f"cell_mv {where}")
    np.testing.assert_array_equal(g["cell_mass_f"], o.cell_mass_f, err_msg=f"cell_mass_f {where}")
    np.testing.assert_array_equal(g["cell_v"], o.cell_v, err_msg=f"cell_v {where}")
    p = s.read_particles()
    np.testing.assert_array_equal(p["pos4"], o.pos4, err_msg=f"pos {where}")
    np.testing.assert_array_equal(p["vel4"], o.vel4, err_msg=f"vel {where}")
    np.testing.assert_array_equal(p["C9"], o.affine, err_msg=f"C {where}")
    np.testing.assert_array_equal(p["density"], o.density, err_msg=f"density {where}")
    if diffuse:
        rec, fl, top = s.read_diffuse()
        assert top == int(o.top[0]), f"free-list top {where}: {top} vs {int(o.top[0])}"
        np.testing.assert_array_equal(fl, o.free_list, err_msg=f"free list {where}")
        for f in ("pos", "vel", "ptype", "alive", "lifetime", "pad"):
            np.testing.assert_array_equal(rec[f], o.pool[f], err_msg=f"diffuse {f} {where}")


@pytest.mark.parametrize("mode", list(MODES))
def test_small_dam_break_every_step(mode):
    grid = (32, 32, 32)
    pos = scenes.lattice_block((2, 2, 2), (14, 20, 14), 0.1, np.random.default_rng(1))
    s, o = make_pair(grid, pos, **MODES[mode])
    for step in range(12):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"step {step}")


@pytest.mark.parametrize("mode", ["direct", "sorted8", "sorted32"])
def test_multi_step_call_matches(mode):
    grid = (32, 24, 28)
    pos = scenes.lattice_block((3, 2, 2.5), (20, 15, 17), 0.1, np.random.default_rng(2))
    s, o = make_pair(grid, pos, **MODES[mode])
    s.step(25)
    o.step(25)
    check_equal(s, o, "after 25 steps")


@pytest.mark.parametrize("mode", ["direct", "sorted8"])
def test_dam_break_64_sample(mode):
    """64^3, block [2,30]x[2,40]x[2,30] (238,336 particles), first 3 steps."""
    grid, pos, _ = scenes.dam_break_small(seed=0)
    s, o = make_pair(grid, pos, **MODES[mode])
    for step in range(3):
        s.step(1)
        o.step(1, threads=8)
        check_equal(s, o, f"step {step}")


def test_fused_clear_equals_separate_clear():
    grid = (32, 32, 32)
    pos = scenes.lattice_block((2, 2, 2), (14, 20, 14), 0.1, np.random.default_rng(3))
    cfg_a = mlsmpm.default_config(*grid, len(pos), flags=KEEP, sort_period=0)
    cfg_b = mlsmpm.default_config(*grid, len(pos), flags=0, sort_period=0)
    with mlsmpm.Scene(cfg_a) as a, mlsmpm.Scene(cfg_b) as b:
        a.reset(pos)
        b.reset(pos)
        a.step(20)
        b.step(20)
        pa, pb = a.read_particles(), b.read_particles()
        for k in pa:
            np.testing.assert_array_equal(pa[k], pb[k], err_msg=k)
        ga, gb = a.read_grid(), b.read_grid(ints=False)
        np.testing.assert_array_equal(ga["cell_v"], gb["cell_v"])


@pytest.mark.parametrize("mode", ["direct", "sorted8", "sorted1"])
def test_diffuse_every_step(mode):
    grid = (32, 32, 32)
    pos = scenes.lattice_block((2, 2, 2), (14, 22, 14), 0.1, np.random.default_rng(4))
    s, o = make_pair(grid, pos, D=20000, **MODES[mode])
    for step in range(15):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"step {step}", diffuse=True)


@pytest.mark.parametrize("mode", ["direct", "sorted1"])
def test_diffuse_uniform_lifetime_mode(mode):
    """lifetime_mode 1 (foam lifetime uniform in [0.5, 3] s, seeded): pool
    bit-exact against the CPU restatement in the same mode."""
    grid = (32, 32, 32)
    pos = scenes.lattice_block((2, 2, 2), (14, 22, 14), 0.1, np.random.default_rng(4))
    s, o = make_pair(grid, pos, D=20000, lifetime_mode=1, **MODES[mode])
    o.gen.lifetime_mode = 1
    for step in range(12):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"step {step}", diffuse=True)
    rec, _, _ = s.read_diffuse()
    foam = rec[(rec["alive"] == 1) & (rec["ptype"] == 1)]
    assert len(foam) > 0 and np.all(foam["lifetime"] <= 3.0)


@pytest.mark.parametrize("mode", ["direct", "sorted8"])
def test_diffuse_pool_saturation(mode):
    """A pool far smaller than the spawn rate: lowest particle index wins."""
    grid = (32, 32, 32)
    pos = scenes.lattice_block((2, 2, 2), (14, 22, 14), 0.1, np.random.default_rng(5))
    s, o = make_pair(grid, pos, D=300, **MODES[mode])
    for step in range(10):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"step {step}", diffuse=True)


@pytest.mark.parametrize("mode", ["direct", "sorted1"])
def test_diffuse_trapped_air_spawns_match_oracle(mode):
    """density_bubble_min raised above the rest density: particles of the
    compressed dam-break core pass the density window, so the trapped-air
    term of the spawn score (shaders.js:813-816) decides spawns too."""
    grid = (32, 32, 32)
    pos = scenes.lattice_block((2, 2, 2), (14, 22, 14), 0.1, np.random.default_rng(6))
    s, o = make_pair(grid, pos, D=40000, density_bubble_min=12.0, **MODES[mode])
    o.gen.density_bubble_min = 12.0
    dens_hi = 0
    for step in range(12):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"step {step}", diffuse=True)
        dens_hi = max(dens_hi, int(np.count_nonzero(o.density > 9.0)))
    assert dens_hi > 0          # the window admitted particles whose pressure exceeds 1
    s.close()


@pytest.mark.parametrize("mode", ["direct", "sorted8"])
def test_non_unit_mass(mode):
    grid = (28, 28, 28)
    rng = np.random.default_rng(6)
    pos = scenes.lattice_block((3, 3, 3), (12, 16, 12), 0.1, rng)
    mass = rng.uniform(0.5, 2.0, len(pos)).astype(np.float32)
    vel = rng.uniform(-1, 1, (len(pos), 3)).astype(np.float32)
    cfg = mlsmpm.default_config(*grid, len(pos), **MODES[mode])
    with mlsmpm.Scene(cfg) as s:
        s.reset(pos, vel, mass=mass)
        o = oracle.OracleScene(grid=grid)
        o.reset(pos, vel, mass=mass)
        for step in range(5):
            s.step(1)
            o.step(1)
            check_equal(s, o, f"step {step}")


@pytest.mark.parametrize("mode", ["direct", "sorted8", "spill"])
def test_random_order_scattered(mode):
    """Particles in random order with random velocities: before the first
    sort every chunk spans the whole block and takes the fallback path."""
    grid = (40, 36, 32)
    rng = np.random.default_rng(8)
    pos = scenes.lattice_block((3, 3, 3), (30, 20, 25), 0.1, rng)
    perm = rng.permutation(len(pos))
    pos = pos[perm]
    vel = rng.uniform(-3, 3, pos.shape).astype(np.float32)
    s, o = make_pair(grid, pos, vel, **MODES[mode])
    for step in range(10):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"step {step}")


@pytest.mark.parametrize("mode", list(MODES))
def test_empty_scene_then_single_particle(mode):
    """A scene with no particles steps and reads back empty (grids zero, the
    diffuse pool untouched); a reset to one particle then matches the oracle
    every step, diffuse pool included."""
    grid = (24, 20, 28)
    D = 256
    cfg = mlsmpm.default_config(*grid, 64, diffuse_capacity=D, seed=1, **MODES[mode])
    with mlsmpm.Scene(cfg) as s:
        s.reset(np.zeros((0, 3), np.float32))
        s.step(3)
        s.step(20)
        p = s.read_particles()
        assert p["pos4"].shape == (0, 4) and p["density"].shape == (0,)
        g = s.read_grid()
        assert not g["cell_mass"].any() and not g["cell_mv"].any() and not g["cell_v"].any()
        rec, fl, top = s.read_diffuse()
        assert top == D and not rec["alive"].any()
        one = np.array([[11.3, 14.7, 9.2]], np.float32)
        v1 = np.array([[1.5, -2.0, 0.7]], np.float32)
        s.reset(one, v1)
        o = oracle.OracleScene(grid=grid, diffuse_max=D, seed=1)
        o.reset(one, v1)
        for step in range(20):
            s.step(1)
            o.step(1)
            check_equal(s, o, f"one particle, step {step}", diffuse=True)


@pytest.mark.parametrize("mode", ["tiled", "sorted8", "sorted1", "wide_items"])
def test_ragged_items_in_one_brick(mode):
    """4097 particles packed into one 8^3 brick (work items of 1024 or 2048
    particles and a last item of one), a few more scattered elsewhere, random
    velocities: every step bit-exact against the oracle."""
    kw = dict(MODES.get(mode, {}))
    if mode == "wide_items":
        kw = dict(flags=KEEP, sort_period=1, tile_ppt=16)
    grid = (24, 24, 24)
    rng = np.random.default_rng(31)
    pos = np.concatenate([rng.uniform(8.0, 16.0, (4097, 3)),
                          rng.uniform(3.0, 20.0, (5, 3))]).astype(np.float32)
    vel = rng.uniform(-1.5, 1.5, pos.shape).astype(np.float32)
    s, o = make_pair(grid, pos, vel, **kw)
    for step in range(12):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"step {step}")
    s.close()


def test_reset_rejects_out_of_grid():
    cfg = mlsmpm.default_config(16, 16, 16, 10)
    with mlsmpm.Scene(cfg) as s:
        with pytest.raises(mlsmpm.MpmError) as e:
            s.reset(np.array([[0.5, 5, 5]], np.float32))
        assert e.value.code == -2
        with pytest.raises(mlsmpm.MpmError):
            s.reset(np.array([[5, 15.0, 5]], np.float32))
        with pytest.raises(mlsmpm.MpmError):
            s.reset(np.array([[5, np.nan, 5]], np.float32))
        with pytest.raises(mlsmpm.MpmError) as e:
            s.reset(np.full((11, 3), 5.0, np.float32))
        assert e.value.code == -3


NO_GRAPH = mlsmpm.MPM_FLAG_NO_GRAPH


@pytest.mark.parametrize("period,timed", [(1, False), (3, False), (8, False), (8, True), (1, True)])
def test_graph_replay_matches_eager_and_oracle(period, timed):
    """Multi-step calls replay a captured HIP graph of whole re-sort cycles;
    the result must equal step-by-step eager enqueue and the CPU restatement,
    with diffuse particles on, and timing (event nodes in the graph) must
    report every step."""
    grid = (40, 32, 36)
    pos = scenes.lattice_block((2, 2, 2), (22, 22, 20), 0.1, np.random.default_rng(11))
    D = 20000
    kw = dict(flags=KEEP, sort_period=period, tile_ppt=4)
    g, o = make_pair(grid, pos, D=D, **kw)
    e, _ = make_pair(grid, pos, D=D, **{**kw, "flags": KEEP | NO_GRAPH})
    if timed:
        g.enable_timing(True)
    g.step(1)
    g.step(50)
    g.step(19)
    for _ in range(70):
        e.step(1)
    o.step(70)
    if timed:
        st = g.step_times()
        assert len(st) == 70 and np.all(st > 0)
        assert g.stage_times()[6] > 0
    check_equal(g, o, "graph vs oracle after 70 steps", diffuse=True)
    pg, pe = g.read_particles(), e.read_particles()
    for k in pg:
        np.testing.assert_array_equal(pg[k], pe[k], err_msg=f"graph vs eager {k}")
    # parameter change drops the captured graph
    g.set_params(gravity=-4.0)
    e.set_params(gravity=-4.0)
    g.step(40)
    for _ in range(40):
        e.step(1)
    pg, pe = g.read_particles(), e.read_particles()
    for k in pg:
        np.testing.assert_array_equal(pg[k], pe[k], err_msg=f"after set_params {k}")
    rg, flg, tg = g.read_diffuse()
    re_, fle, te = e.read_diffuse()
    assert tg == te
    np.testing.assert_array_equal(flg, fle)
    g.close()
    e.close()


@pytest.mark.parametrize("period", [1, 8])
def test_reset_keeps_graphs_and_bits(period):
    """A reset to the same particle count keeps the captured graphs: the run
    after the reset must equal a fresh scene's (and the oracle's), with the
    diffuse pool on; a reset to another count drops them and still matches."""
    grid = (40, 32, 36)
    pos = scenes.lattice_block((2, 2, 2), (22, 22, 20), 0.1, np.random.default_rng(12))
    D = 20000
    kw = dict(flags=KEEP, sort_period=period, tile_ppt=4)
    g, _ = make_pair(grid, pos, D=D, **kw)
    g.step(40)
    g.reset(pos)
    g.step(40)
    f, o = make_pair(grid, pos, D=D, **kw)
    f.step(40)
    o.step(40)
    check_equal(g, o, "reset-and-replay vs oracle after 40 steps", diffuse=True)
    pg, pf = g.read_particles(), f.read_particles()
    for k in pg:
        np.testing.assert_array_equal(pg[k], pf[k], err_msg=f"reset-and-replay vs fresh {k}")
    # a smaller particle set: launch sizes change, the graphs are re-captured
    sub = pos[: len(pos) // 2 + 7]
    g.reset(sub)
    g.step(40)
    f2, o2 = make_pair(grid, sub, D=D, **kw)
    o2.step(40)
    check_equal(g, o2, "after reset to a smaller count", diffuse=True)
    g.close()
    f.close()
    f2.close()


@pytest.mark.parametrize("period,timed", [(1, False), (8, False), (1, True)])
def test_prepare_captures_graphs_the_run_replays(period, timed):
    """mpm_prepare right after a reset (and again mid-run) captures without
    stepping: the steps that follow equal eager steps and the oracle."""
    grid = (40, 32, 36)
    pos = scenes.lattice_block((2, 2, 2), (22, 22, 20), 0.1, np.random.default_rng(13))
    kw = dict(flags=KEEP, sort_period=period, tile_ppt=4)
    g, o = make_pair(grid, pos, D=20000, **kw)
    e, _ = make_pair(grid, pos, D=20000, flags=KEEP | mlsmpm.MPM_FLAG_NO_GRAPH,
                     sort_period=period, tile_ppt=4)
    if timed:
        g.enable_timing(True)
    g.prepare()
    assert g.counts()[1] == 0          # nothing ran
    check_equal(g, o, "after prepare", diffuse=True)
    g.step(20)
    g.prepare()
    g.step(19)
    e.step(39)
    o.step(39)
    check_equal(g, o, "prepared run vs oracle after 39 steps", diffuse=True)
    pg, pe = g.read_particles(), e.read_particles()
    for k in pg:
        np.testing.assert_array_equal(pg[k], pe[k], err_msg=f"prepared vs eager {k}")
    if timed:
        assert len(g.step_times()) == 39
    g.close()
    e.close()


def test_state_round_trip_with_diffuse_pool():
    """A scene rebuilt from another's read-back state (particles in reset
    order, the diffuse records, free list and top, the generator seed)
    continues bit-identically: mpm_write_diffuse restores which slots are
    alive, and the step depends on no other carried state."""
    grid = (40, 32, 36)
    pos = scenes.lattice_block((2, 2, 2), (22, 22, 20), 0.1, np.random.default_rng(21))
    D, seed0, n0 = 20000, 5, 30
    kw = dict(flags=KEEP, sort_period=1, tile_ppt=4, seed=seed0)
    a, _ = make_pair(grid, pos, D=D, **kw)
    a.step(n0)
    rec, fl, top = a.read_diffuse()
    assert int(np.count_nonzero(rec["alive"])) > 0 and top < D
    p = a.read_particles()
    b, _ = make_pair(grid, p["pos4"][:, :3], vel=p["vel4"][:, :3], D=D, **kw)
    b.reset(p["pos4"][:, :3], p["vel4"][:, :3], p["C9"])
    b.set_seed(seed0 + n0)
    b.write_diffuse(rec, fl, top)
    a.step(25)
    b.step(25)
    pa, pb = a.read_particles(), b.read_particles()
    for k in pa:
        np.testing.assert_array_equal(pa[k], pb[k], err_msg=f"round trip {k}")
    ra, fa, ta = a.read_diffuse()
    rb, fb, tb = b.read_diffuse()
    assert ta == tb
    np.testing.assert_array_equal(fa, fb)
    for f in ("pos", "vel", "ptype", "alive", "lifetime"):
        np.testing.assert_array_equal(ra[f], rb[f], err_msg=f"round trip diffuse {f}")
    a.close()
    b.close()


def wall_block(count):
    """WallParams (shaders.js:123-133): header (count) + 4 walls x 20 floats:
    centre, cos yaw, local min, sin yaw, local max, cos pitch, linear
    velocity, sin pitch, angular velocity, pad."""
    w = np.zeros(84, np.float32)
    w[0] = count
    walls = [
        # a yawed box moving into the fluid block's side, spinning about y
        dict(c=(16.5, 6.0, 9.0), yaw=0.5, pitch=0.0, mn=(-2.0, -4.0, -5.0), mx=(2.0, 4.0, 5.0),
             v=(-0.4, 0.0, 0.1), w=(0.0, 0.15, 0.0)),
        # a pitched plate over the block, sinking
        dict(c=(8.0, 23.0, 8.0), yaw=0.0, pitch=0.3, mn=(-5.0, -0.5, -5.0), mx=(5.0, 0.5, 5.0),
             v=(0.0, -0.6, 0.0), w=(0.05, 0.0, 0.0)),
    ]
    for i, d in enumerate(walls[:count]):
        o = 4 + 20 * i
        w[o:o + 3] = d["c"]
        w[o + 3] = np.cos(d["yaw"])
        w[o + 4:o + 7] = d["mn"]
        w[o + 7] = np.sin(d["yaw"])
        w[o + 8:o + 11] = d["mx"]
        w[o + 11] = np.cos(d["pitch"])
        w[o + 12:o + 15] = d["v"]
        w[o + 15] = np.sin(d["pitch"])
        w[o + 16:o + 19] = d["w"]
    return w


@pytest.mark.parametrize("mode", ["direct", "sorted1", "sorted8"])
def test_moving_walls_match_oracle(mode):
    """Walls at count > 0 (UPDATE_GRID, G2P and ADVECT wall loops) with the
    diffuse pool on: every grid, particle and pool value bit-exact."""
    grid = (32, 32, 32)
    pos = scenes.lattice_block((2, 2, 2), (14, 20, 14), 0.1, np.random.default_rng(21))
    walls = wall_block(2)
    s, _ = make_pair(grid, pos, D=20000, **MODES[mode])
    s.set_walls(walls)
    o = oracle.OracleScene(grid=grid, diffuse_max=20000, seed=1, walls=walls)
    o.reset(pos)
    for step in range(10):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"step {step}", diffuse=True)
    s.close()


# The reference's two presets (main.js:19-21, config.js:1-7): aspect 2x2x3,
# particle radius 0.0125 / 0.0175, capacity 400,000 / 80,000, diffuse pool
# 30,000 / 12,000 slots, fillBlock(0.03, 0.03, 0.03, 0.97, 0.97, 0.37).
# WG_SIZE 64 / 32 shapes only the WebGPU dispatches and has no counterpart.
PRESETS = {"desktop": (0.0125, 400_000, 30_000), "mobile": (0.0175, 80_000, 12_000)}


@pytest.mark.parametrize("preset", list(PRESETS))
def test_reference_presets(preset):
    radius, cap, D = PRESETS[preset]
    cfg = mlsmpm.config_from_aspect(2, 2, 3, radius, cap)
    cfg.flags = KEEP
    cfg.diffuse_capacity = D
    grid = (cfg.grid_x, cfg.grid_y, cfg.grid_z)
    assert grid == {"desktop": (40, 40, 60), "mobile": (29, 29, 43)}[preset]
    pos = scenes.fill_block(grid, 0.03, 0.03, 0.03, 0.97, 0.97, 0.37,
                            np.random.default_rng(5), capacity=cap)
    # the scene takes coordinates in [1, grid - 1): the mobile fill's last
    # column (x, y near 28.12 of 29 cells) moves inside, as FluidHIP uploads it
    hi = np.nextafter(np.float32(np.array(grid) - 1), np.float32(0))
    moved = int(np.count_nonzero(pos > hi))
    assert (moved > 0) == (preset == "mobile")
    pos = np.minimum(pos, hi).astype(np.float32)
    s = mlsmpm.Scene(cfg)
    s.reset(pos)
    o = oracle.OracleScene(grid=grid, diffuse_max=D, seed=cfg.seed)
    o.reset(pos)
    for step in range(6):
        s.step(1)
        o.step(1)
        check_equal(s, o, f"{preset} step {step}", diffuse=True)
    s.close()


@pytest.mark.parametrize("period", [1, 8])
def test_long_run_with_diffuse_matches_oracle(period):
    """400 steps of the 64^3 dam break (238,336 particles) with a 65,536-slot
    pool, in 100-step captured replays: grids, particles and the whole pool
    stay bit-exact through hundreds of re-sorts and pool churn."""
    grid, pos, _ = scenes.dam_break_small(seed=3)
    D = 65536
    s, o = make_pair(grid, pos, D=D, flags=KEEP, sort_period=period)
    for chunk in range(4):
        s.step(100)
        o.step(100)
        check_equal(s, o, f"period {period} after {100 * (chunk + 1)} steps", diffuse=True)
    assert s.diffuse_alive() > 0
    s.close()


@pytest.mark.parametrize("lifetime_mode", [0, 1])
def test_two_block_collision_small_matches_oracle(lifetime_mode):
    """The two-block collision at 64^3 (two 16x16x32-cell blocks at +/- 2
    cells per step, 131,072 particles, 32,768 diffuse slots): particles
    drift past the LDS margin and collide, so the overflow path and the
    spray carry most of the work; bit-exact every 20 steps for 80 steps."""
    rng = np.random.default_rng(31)
    a = scenes.lattice_block((4, 2, 16), (20, 18, 48), 0.1, rng)
    b = scenes.lattice_block((44, 2, 16), (60, 18, 48), 0.1, rng)
    pos = np.concatenate([a, b])
    vel = np.zeros_like(pos)
    vel[: len(a), 0] = 2.0 / 0.3
    vel[len(a):, 0] = -2.0 / 0.3
    grid, D = (64, 64, 64), 32768
    cfg = mlsmpm.default_config(*grid, len(pos), diffuse_capacity=D, flags=KEEP, sort_period=1,
                                lifetime_mode=lifetime_mode, seed=1)
    s = mlsmpm.Scene(cfg)
    s.reset(pos, vel)
    o = oracle.OracleScene(grid=grid, diffuse_max=D, seed=1)
    o.reset(pos, vel)
    o.gen.lifetime_mode = lifetime_mode
    for chunk in range(4):
        s.step(20)
        o.step(20)
        check_equal(s, o, f"lifetime {lifetime_mode} after {20 * (chunk + 1)} steps",
                    diffuse=True)
    assert s.diffuse_alive() > 0
    s.close()
