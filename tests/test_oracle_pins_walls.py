"""Closed-form pins of the oracle's moving-wall geometry, the wall boundary
conditions of UPDATE_GRID / G2P / DIFFUSE_ADVECT, the density-gradient
normal and the spawn score (oracle/mpm_oracle.c).

Expectations come from the mathematics the reference states in its comments
(shaders.js:110-193: R = R_y(yaw) R_x(pitch), local = R^T (p - c), surface
velocity v + omega x r, outward normal and signed distance of a box), from
numpy float64 matrices and cross products, and from hand-derived values of
the boundary rules (shaders.js:489-507, 657-693, 766-787, 803-826, 959-979).
A sign, axis or transposition mistake in the restatement fails one of them.
"""
import numpy as np
import pytest

from oracle import oracle as O

F32 = np.float32
EPS = 2e-6


def rot_y(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def rot_x(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


ANGLES = [(0.0, 0.0), (np.pi / 2, 0.0), (0.0, np.pi / 2), (np.pi / 2, np.pi / 2),
          (0.37, -0.81), (-2.1, 1.3), (np.pi, 0.25)]


# ---- wall_to_local / wall_to_world_dir (shaders.js:137-157) -----------------
@pytest.mark.parametrize("yaw,pitch", ANGLES)
def test_wall_rotations_are_R_and_its_transpose(yaw, pitch):
    c = np.array([10.25, 7.5, 12.0])
    w = O.make_wall(c, (-1, -1, -1), (1, 1, 1), yaw=yaw, pitch=pitch)
    R = rot_y(yaw) @ rot_x(pitch)
    rng = np.random.default_rng(int(1000 * (yaw + 3) + 7 * (pitch + 3)))
    for _ in range(8):
        v = rng.uniform(-3, 3, 3)
        p = c + rng.uniform(-4, 4, 3)
        world = O.wall_helper("wall_to_world_dir", v, w).astype(np.float64)
        local = O.wall_helper("wall_to_local", p, w).astype(np.float64)
        assert np.allclose(world, R @ v, atol=1e-5)
        assert np.allclose(local, R.T @ (p - c), atol=1e-5)
        # the two are inverse maps: c + R (R^T (p - c)) = p
        back = O.wall_helper("wall_to_world_dir", local, w).astype(np.float64) + c
        assert np.allclose(back, p, atol=2e-5)


def test_wall_quarter_turns_map_axes_to_known_images():
    e = np.eye(3)
    yaw = O.make_wall((0, 0, 0), (-1, -1, -1), (1, 1, 1), yaw=np.pi / 2)
    # R_y(90): x -> -z, z -> +x, y fixed
    assert np.allclose(O.wall_helper("wall_to_world_dir", e[0], yaw), [0, 0, -1], atol=EPS)
    assert np.allclose(O.wall_helper("wall_to_world_dir", e[2], yaw), [1, 0, 0], atol=EPS)
    assert np.allclose(O.wall_helper("wall_to_world_dir", e[1], yaw), [0, 1, 0], atol=EPS)
    pitch = O.make_wall((0, 0, 0), (-1, -1, -1), (1, 1, 1), pitch=np.pi / 2)
    # R_x(90): y -> +z, z -> -y, x fixed
    assert np.allclose(O.wall_helper("wall_to_world_dir", e[1], pitch), [0, 0, 1], atol=EPS)
    assert np.allclose(O.wall_helper("wall_to_world_dir", e[2], pitch), [0, -1, 0], atol=EPS)
    assert np.allclose(O.wall_helper("wall_to_world_dir", e[0], pitch), [1, 0, 0], atol=EPS)
    both = O.make_wall((0, 0, 0), (-1, -1, -1), (1, 1, 1), yaw=np.pi / 2, pitch=np.pi / 2)
    # pitch first, then yaw: y -> z -> x
    assert np.allclose(O.wall_helper("wall_to_world_dir", e[1], both), [1, 0, 0], atol=EPS)
    # to_local undoes it: a point one unit along world +x from the centre is local +y
    assert np.allclose(O.wall_helper("wall_to_local", e[0], both), [0, 1, 0], atol=EPS)


# ---- wall_vel_at (shaders.js:161-164) ---------------------------------------
def test_wall_vel_at_is_v_plus_omega_cross_r():
    rng = np.random.default_rng(5)
    for _ in range(10):
        c, v, om = rng.uniform(4, 12, 3), rng.uniform(-1, 1, 3), rng.uniform(-0.5, 0.5, 3)
        w = O.make_wall(c, (-1, -1, -1), (1, 1, 1), yaw=0.3, vel=v, omega=om)
        p = c + rng.uniform(-3, 3, 3)
        got = O.wall_helper("wall_vel_at", p, w).astype(np.float64)
        assert np.allclose(got, v + np.cross(om, p - c), atol=1e-5)
    # a pure spin about y: at r = +x the surface moves along -z (right hand)
    w = O.make_wall((5, 5, 5), (-1, -1, -1), (1, 1, 1), omega=(0, 2, 0))
    assert np.allclose(O.wall_helper("wall_vel_at", (6, 5, 5), w), [0, 0, -2], atol=EPS)


# ---- wall_sdf_normal (shaders.js:174-193) -----------------------------------
MN, MX = (-2.0, -1.0, -3.0), (2.0, 1.0, 3.0)


@pytest.mark.parametrize("lp,depth,normal", [
    ((1.5, 0.0, 0.0), -0.5, (1, 0, 0)),      # nearest the +x face
    ((-1.75, 0.1, 0.5), -0.25, (-1, 0, 0)),  # nearest the -x face
    ((0.0, -0.8, 0.0), -0.2, (0, -1, 0)),    # -y face
    ((0.5, 0.3, 2.6), -0.4, (0, 0, 1)),      # +z face beats y (0.7) and x (1.5)
    ((0.0, 0.0, -2.95), -0.05, (0, 0, -1)),
])
def test_sdf_inside_takes_the_shallowest_face(lp, depth, normal):
    d, n = O.wall_sdf_normal(lp, MN, MX)
    assert abs(d - depth) < EPS
    assert np.array_equal(n, np.array(normal, F32))


@pytest.mark.parametrize("lp,dist,normal", [
    ((0.0, 3.0, 0.0), 2.0, (0, 1, 0)),                          # above a face
    ((-2.5, 0.0, 1.0), 0.5, (-1, 0, 0)),
    ((3.0, 4.0, 0.0), np.sqrt(1 + 9), (1 / np.sqrt(10), 3 / np.sqrt(10), 0)),  # past an edge
    ((3.0, 2.0, 4.0), np.sqrt(3), (1, 1, 1) / np.sqrt(3)),      # past a corner
    ((-4.0, -2.0, -5.0), 3.0, (-2 / 3, -1 / 3, -2 / 3)),
])
def test_sdf_outside_is_distance_to_the_nearest_surface_point(lp, dist, normal):
    d, n = O.wall_sdf_normal(lp, MN, MX)
    assert abs(d - dist) < 1e-5
    assert np.allclose(n, normal, atol=1e-6)
    # the normal points from the surface point back to the query point
    surf = np.clip(np.array(lp), MN, MX)
    assert np.allclose(surf + d * n.astype(np.float64), lp, atol=1e-5)


# ---- UPDATE_GRID wall condition (shaders.js:489-507) ------------------------
def _update_scene(walls, cells):
    """Cells given as {(x, y, z): velocity} with unit mass; gravity off."""
    g = (24, 24, 24)
    s = O.OracleScene(grid=g)
    s.params.gravity = 0.0
    s.walls = O.wall_params(*walls)
    gx, gy, _ = g
    idx = {}
    for (x, y, z), v in cells.items():
        c = (z * gy + y) * gx + x
        idx[(x, y, z)] = c
        s.cell_mass[c] = 100000
        s.cell_mv[3 * c:3 * c + 3] = [int(round(t * 1e5)) for t in v]
    s.k_update()
    return {k: s.cell_v3()[c].astype(np.float64) for k, c in idx.items()}


def test_update_inside_a_moving_box_takes_the_wall_velocity():
    wv = (0.25, 0.1, -0.5)
    w = O.make_wall((10, 10, 10), (-3, -3, -3), (3, 3, 3), vel=wv)
    out = _update_scene([w], {(10, 10, 10): (1.0, -2.0, 0.5), (8, 11, 12): (0.0, 0.0, 0.0)})
    assert np.array_equal(out[(10, 10, 10)], np.array(wv, F32).astype(np.float64))
    assert np.array_equal(out[(8, 11, 12)], np.array(wv, F32).astype(np.float64))


def test_update_band_removes_only_the_approaching_normal_component():
    # +y face at y = 13; cell (10, 13, 10) has its centre 0.5 above it
    wv = (0.25, 0.1, -0.5)
    w = O.make_wall((10, 10, 10), (-3, -3, -3), (3, 3, 3), vel=wv)
    toward, away, far = (0.3, -0.4, 0.2), (0.3, 0.4, 0.2), (0.3, -0.4, 0.2)
    out = _update_scene([w], {(10, 13, 10): toward, (11, 13, 11): away, (10, 15, 10): far,
                              (10, 14, 10): far})
    # relative normal speed -0.5 removed: vy becomes the wall's 0.1, x and z untouched
    assert np.allclose(out[(10, 13, 10)], [0.3, 0.1, 0.2], atol=EPS)
    assert np.allclose(out[(11, 13, 11)], away, atol=EPS)      # moving away: unchanged
    assert np.allclose(out[(10, 15, 10)], far, atol=EPS)       # 2.5 cells out: unchanged
    assert np.allclose(out[(10, 14, 10)], far, atol=EPS)       # 1.5 out: sdf >= 1, unchanged


def test_update_band_normal_follows_the_rotation():
    # yaw 90: the local +x face (x = 1) faces world -z.  The cell centre
    # (10.5, 10.5, 10.5) is 1.5 along world -z from the centre: local (1.5, 0, 0).
    w = O.make_wall((10.5, 10.5, 12.0), (-1, -3, -4), (1, 3, 4), yaw=np.pi / 2)
    out = _update_scene([w], {(10, 10, 10): (0.2, 0.1, 0.6), (11, 10, 10): (0.2, 0.1, -0.6)})
    assert np.allclose(out[(10, 10, 10)], [0.2, 0.1, 0.0], atol=1e-5)   # +z toward the wall
    # cell (11, 10, 10): local (1.5, 0, -1) -> same face; moving -z is away
    assert np.allclose(out[(11, 10, 10)], [0.2, 0.1, -0.6], atol=1e-5)


# ---- G2P lookahead spring and push-out (shaders.js:657-693) -----------------
def _g2p_scene(wall, pos, cell_v=(0.0, 0.0, 0.0)):
    g = (24, 24, 24)
    s = O.OracleScene(grid=g)
    s.walls = O.wall_params(wall)
    s.cell_v[:] = np.tile(np.array(cell_v, F32), g[0] * g[1] * g[2])
    s.reset(np.array(pos, dtype=np.float64))
    s.k_weights()
    s.k_g2p()
    return s.pos4[:, :3].astype(np.float64), s.vel4[:, :3].astype(np.float64)


def test_g2p_pushout_is_capped_at_half_a_cell():
    w = O.make_wall((12, 12, 12), (-3, -3, -3), (3, 3, 3))
    pos, vel = _g2p_scene(w, [(12.0, 12.0, 14.0), (12.0, 12.0, 14.8), (9.6, 12.0, 12.0)])
    # 1.0 deep below +z: pushed 0.5 (the cap); spring s = (1 * min(1/3, 1)) * 1
    assert np.allclose(pos[0], [12, 12, 14.5], atol=EPS)
    assert np.allclose(vel[0], [0, 0, 1 / 3], atol=EPS)
    # 0.2 deep: pushed onto the face; s = (0.2 / 3) * 0.2
    assert np.allclose(pos[1], [12, 12, 15.0], atol=EPS)
    assert np.allclose(vel[1], [0, 0, 0.04 / 3], atol=EPS)
    # 0.6 deep below -x: pushed 0.5 along -x
    assert np.allclose(pos[2], [9.1, 12, 12], atol=EPS)
    assert np.allclose(vel[2], [-0.12, 0, 0], atol=EPS)


def test_g2p_pushout_matches_a_moving_wall_normal_speed():
    # the wall moves along +z at 0.5; after the push the particle's relative
    # normal velocity 1/3 - 0.5 < 0 is removed: vz = 0.5
    w = O.make_wall((12, 12, 12), (-3, -3, -3), (3, 3, 3), vel=(0.1, 0.0, 0.5))
    pos, vel = _g2p_scene(w, [(12.0, 12.0, 14.0)])
    assert np.allclose(pos[0], [12, 12, 14.5], atol=EPS)
    assert np.allclose(vel[0], [0, 0, 0.5], atol=EPS)


def test_g2p_lookahead_spring_acts_before_contact():
    # grid velocity -1 along z: np = 15.5 - 0.3 = 15.2 (outside), the
    # lookahead np + v dt k = 14.6 is 0.4 inside: vz += (0.4 / 3) * 0.4
    w = O.make_wall((12, 12, 12), (-3, -3, -3), (3, 3, 3))
    pos, vel = _g2p_scene(w, [(12.0, 12.0, 15.5)], cell_v=(0.0, 0.0, -1.0))
    assert np.allclose(pos[0], [12, 12, 15.2], atol=1e-5)
    assert np.allclose(vel[0], [0, 0, -1 + 0.16 / 3], atol=1e-5)


# ---- DIFFUSE_ADVECT wall reflection (shaders.js:959-979) --------------------
def _advect_one(wall, pos, vel, ptype=0):
    s = O.OracleScene(grid=(24, 24, 24), diffuse_max=1)
    s.params.gravity = 0.0
    s.walls = O.wall_params(wall)
    rec = np.zeros(1, s.L.drec_dtype)
    rec["alive"], rec["ptype"], rec["lifetime"] = 1, ptype, 5.0
    rec[0]["pos"][:3], rec[0]["vel"][:3] = pos, vel
    s.set_pool(rec, np.zeros(1, np.uint32), 0)
    s.cell_v[:] = 0
    s.k_advect()
    return s.pool[0]


def test_advect_reflects_off_a_moving_wall_with_restitution():
    w = O.make_wall((10, 10, 10), (-2, -2, -2), (2, 2, 2), vel=(0.2, 0.0, 0.0))
    r = _advect_one(w, (11.7, 10.0, 10.0), (-1.0, 0.5, 0.0))
    # ballistic to (11.4, 10.15, 10): 0.6 inside the +x face -> onto the face
    assert np.allclose(r["pos"][:3], [12.0, 10.15, 10.0], atol=EPS)
    # relative normal velocity -1.2 becomes +0.36 = -0.3 * -1.2; tangent kept
    assert np.allclose(r["vel"][:3], [0.2 + 0.36, 0.5, 0.0], atol=EPS)
    assert r["alive"] == 1


def test_advect_reflection_uses_the_spinning_wall_surface_velocity():
    om = np.array([0.0, 0.0, 0.3])
    c = np.array([10.0, 10.0, 10.0])
    w = O.make_wall(c, (-2, -2, -2), (2, 2, 2), omega=om)
    r = _advect_one(w, (10.5, 11.9, 10.0), (0.0, 0.1, 0.0))
    p = np.array([10.5, 11.9 + 0.03, 10.0])        # after the ballistic step, 0.07 deep
    p_out = np.array([10.5, 12.0, 10.0])            # pushed onto the +y face
    assert np.allclose(r["pos"][:3], p_out, atol=EPS)
    wv = np.cross(om, p_out - c)                    # surface velocity at the pushed point
    vrel = np.array([0.0, 0.1, 0.0]) - wv
    vn = vrel[1]
    want = wv + vrel - 1.3 * vn * np.array([0, 1, 0]) if vn < 0 else np.array([0.0, 0.1, 0.0])
    assert np.allclose(r["vel"][:3], want, atol=EPS)
    del p


def test_advect_outside_every_wall_is_untouched_by_it():
    w = O.make_wall((10, 10, 10), (-2, -2, -2), (2, 2, 2), vel=(5.0, 5.0, 5.0))
    r = _advect_one(w, (14.5, 10.0, 10.0), (0.1, 0.0, 0.0))
    assert np.allclose(r["pos"][:3], [14.53, 10.0, 10.0], atol=EPS)
    assert np.allclose(r["vel"][:3], [0.1, 0.0, 0.0], atol=EPS)


# ---- estimate_normal (shaders.js:766-787) -----------------------------------
def _ramp(grid, grad, c0=50.0):
    gx, gy, gz = grid
    zz, yy, xx = np.meshgrid(np.arange(gz), np.arange(gy), np.arange(gx), indexing="ij")
    return (c0 + grad[0] * xx + grad[1] * yy + grad[2] * zz).reshape(-1).astype(F32)


def test_estimate_normal_points_down_the_mass_gradient():
    g = (12, 12, 12)
    P = O.default_params(*g)
    for grad in [(2.0, 3.0, -1.0), (0.0, -4.0, 0.0), (0.5, 0.0, 0.25)]:
        cm = _ramp(g, grad)
        ok, n = O.estimate_normal(P, cm, (5.3, 6.7, 4.1))
        gd = np.array(grad)
        assert ok
        assert np.allclose(n, -gd / np.linalg.norm(gd), atol=1e-6)


def test_estimate_normal_invalid_at_the_border_and_on_flat_mass():
    g = (12, 12, 12)
    P = O.default_params(*g)
    cm = _ramp(g, (1.0, 1.0, 1.0))
    for p in [(0.5, 5.5, 5.5), (5.5, 11.2, 5.5), (5.5, 5.5, 11.0)]:
        ok, n = O.estimate_normal(P, cm, p)
        assert not ok and np.array_equal(n, np.zeros(3, F32))
    ok, _ = O.estimate_normal(P, cm, (1.0, 5.5, 10.99))      # 1 <= floor(x) < X - 1
    assert ok
    ok, _ = O.estimate_normal(P, np.full(12 ** 3, 3.0, F32), (5.5, 5.5, 5.5))
    assert not ok


# ---- spawn score (shaders.js:803-826) ---------------------------------------
def _score(vel, density, mass=1.0, grad=(0.0, -2.0, 0.0), **gen):
    g = (12, 12, 12)
    P = O.default_params(*g)
    G = O.default_gen()
    for k, v in gen.items():
        setattr(G, k, v)
    return O.spawn_score(P, G, _ramp(g, grad), (5.5, 5.5, 5.5), vel, mass, density)


def test_spawn_score_kinetic_term():
    sc, t = _score((1.0, 2.0, 2.0), 3.0, mass=2.0, grad=(0.0, 0.0, 0.0))
    assert t["speed"] == 3.0 and t["ke"] == 9.0      # 0.5 * 2 * 9
    assert t["trapped"] == 0.0 and t["crest"] == 0.0
    assert sc == 9.0


def test_spawn_score_wave_crest_above_the_dot_threshold_only():
    # mass falls off along +y, so the surface normal is +y
    sc, t = _score((4.0, 3.0, 0.0), 3.0)             # dot(v/|v|, n) = 0.6 > 0.3
    assert abs(t["crest"] - 0.6 * 5.0) < 1e-5
    assert abs(sc - (12.5 + 1.5 * 3.0)) < 1e-4
    _, t = _score((0.96, 0.28, 0.0), 3.0)            # 0.28 <= 0.3: no crest
    assert t["crest"] == 0.0
    _, t = _score((4.0, -3.0, 0.0), 3.0)             # moving into the fluid
    assert t["crest"] == 0.0
    # strictly above: dot = 1 exactly at threshold 1 gives nothing, just below gives |v|
    _, t = _score((0.0, 2.0, 0.0), 3.0, crest_dot_threshold=1.0)
    assert t["crest"] == 0.0
    _, t = _score((0.0, 2.0, 0.0), 3.0, crest_dot_threshold=0.999)
    assert t["crest"] == 2.0


def test_spawn_score_trapped_air_above_rest_density():
    # rest 8, stiffness 100: rho = 8.4 gives p = 100 * (1.05 - 1) = 5, trapped 4
    sc, t = _score((0.0, 0.0, 0.0), 8.4, grad=(0.0, 0.0, 0.0))
    assert abs(t["trapped"] - 4.0) < 1e-4
    assert abs(sc - 8.0) < 2e-4                     # ke 0 + 2 * trapped
    sc, t = _score((1.0, 0.0, 0.0), 9.0, grad=(0.0, 0.0, 0.0))
    assert abs(t["trapped"] - (100 * (9 / 8 - 1) - 1)) < 1e-4
    assert abs(sc - (0.5 + 2 * t["trapped"])) < 1e-4
    # just below rest: the cohesion floor -0.1 minus the threshold 1 < 0
    _, t = _score((0.0, 0.0, 0.0), 7.9, grad=(0.0, 0.0, 0.0))
    assert t["trapped"] == 0.0
    # pressure_threshold shifts the term one for one
    _, t = _score((0.0, 0.0, 0.0), 8.4, grad=(0.0, 0.0, 0.0), pressure_threshold=2.0)
    assert abs(t["trapped"] - 3.0) < 1e-4


def test_generate_spawns_from_trapped_air_alone():
    """A particle at rest above rest density spawns foam through the trapped-
    air term once the density window admits it (density_bubble_min raised)."""
    s = O.OracleScene(grid=(16, 16, 16), diffuse_max=4)
    s.reset(np.array([[8.5, 8.5, 8.5], [6.5, 9.5, 7.5]]))
    s.density[:] = [8.4, 7.9]                        # score 8 (spawns), score 0 (not)
    s.gen.density_bubble_min = 10.0
    s.k_generate()
    assert s.top[0] == 3
    r = s.pool[3]
    assert r["alive"] == 1 and r["ptype"] == 1
    assert np.array_equal(r["vel"][:3], np.zeros(3, F32))   # jitter scales with speed 0
    assert s.pool["alive"].sum() == 1
