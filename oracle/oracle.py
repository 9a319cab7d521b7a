"""ctypes binding of the CPU restatement (oracle/mpm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg, always as the checker (or the timed CPU
baseline), never as the product path.

``OracleScene`` mirrors the reference host classes closely enough for tests:
``FluidGPU`` state and simFrame (fluid-gpu.js:9-70, 648-684) plus the optional
``DiffuseGPU`` pool (diffuse-gpu.js:15-166).  Arrays use the reference layouts
(pos/vel stride-4 vec4 with w = 0, C as 9 reals row-major, cell_mv 3
interleaved i32 per cell).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "build")


def build() -> None:
    """Compile liborc32.so / liborc64.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class Params(C.Structure):
    """Params uniform, 80 bytes (shaders.js:27-49)."""

    _fields_ = [
        ("grid_x", C.c_uint32), ("grid_y", C.c_uint32), ("grid_z", C.c_uint32),
        ("particle_num", C.c_uint32),
        ("dt", C.c_float), ("gravity", C.c_float), ("rest_density", C.c_float),
        ("stiffness", C.c_float), ("eos_power", C.c_float), ("viscosity", C.c_float),
        ("hard_min", C.c_float), ("hard_max_x", C.c_float), ("hard_max_y", C.c_float),
        ("hard_max_z", C.c_float), ("wall_min", C.c_float), ("wall_stiffness", C.c_float),
        ("lookahead_k", C.c_float), ("pad0", C.c_float), ("pad1", C.c_float), ("pad2", C.c_float),
    ]


class Gen(C.Structure):
    """GenParams (shaders.js:734-743)."""

    _fields_ = [
        ("ke_threshold", C.c_float), ("spawn_rate_k", C.c_float),
        ("density_spray_max", C.c_float), ("density_bubble_min", C.c_float),
        ("lifetime_foam", C.c_float), ("pressure_threshold", C.c_float),
        ("crest_dot_threshold", C.c_float), ("seed", C.c_uint32),
        ("lifetime_mode", C.c_uint32),
    ]


class Adv(C.Structure):
    """AdvectParams (shaders.js:872-877)."""

    _fields_ = [("diffuse_max", C.c_uint32), ("buoyancy", C.c_float), ("drag", C.c_float),
                ("pad", C.c_float)]


def default_params(gx: int, gy: int, gz: int, n: int = 0) -> Params:
    """FluidGPU constructor defaults (fluid-gpu.js:24-45), SUBSTEPS = 1."""
    p = Params()
    p.grid_x, p.grid_y, p.grid_z, p.particle_num = gx, gy, gz, n
    p.dt, p.gravity, p.rest_density, p.stiffness = 0.3, -0.98, 8.0, 100.0
    p.eos_power, p.viscosity = 1.0, 0.01
    p.hard_min = 2.0
    p.hard_max_x, p.hard_max_y, p.hard_max_z = gx - 2.0, gy - 2.0, gz - 2.0
    p.wall_min, p.wall_stiffness, p.lookahead_k = 3.0, 1.0, 2.0
    return p


def default_gen(seed: int = 1) -> Gen:
    """DiffuseGPU tunables (diffuse-gpu.js:20-31); _seed starts at 1."""
    g = Gen()
    g.ke_threshold, g.spawn_rate_k = 0.5, 3.0
    g.density_spray_max, g.density_bubble_min = 1.5, 6.0
    g.lifetime_foam, g.pressure_threshold, g.crest_dot_threshold = 3.0, 1.0, 0.3
    g.seed = seed
    return g


def default_adv(dmax: int) -> Adv:
    a = Adv()
    a.diffuse_max, a.buoyancy, a.drag = dmax, 1.2, 4.0
    return a


class _Lib:
    def __init__(self, path: str, real):
        self.lib = C.CDLL(path)
        self.real = np.dtype(real)
        rb = self.lib.orc_real_bytes()
        assert rb == self.real.itemsize, (path, rb)
        self.rp = np.ctypeslib.ndpointer(dtype=self.real, flags="C_CONTIGUOUS")
        self.ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        self.up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
        ct = C.c_float if self.real == np.float32 else C.c_double
        L = self.lib
        L.orc_encode.argtypes = [ct]
        L.orc_encode.restype = C.c_int32
        L.orc_decode.argtypes = [C.c_int32]
        L.orc_decode.restype = ct
        L.orc_diffuse_rand.argtypes = [C.c_uint32]
        L.orc_diffuse_rand.restype = C.c_float
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_hand_bbox.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_apply_hand.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_ray_perp_dist.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_ray_perp_dist.restype = ct
        L.orc_raycast.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
        L.orc_raycast.restype = C.c_int
        L.orc_delete_mask.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
        L.orc_delete_mask.restype = C.c_uint32
        for name in ("orc_wall_to_local", "orc_wall_to_world_dir", "orc_wall_vel_at"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_wall_sdf_normal.argtypes = [C.c_void_p] * 4
        L.orc_wall_sdf_normal.restype = ct
        L.orc_estimate_normal.argtypes = [C.c_void_p] * 4
        L.orc_estimate_normal.restype = C.c_int
        L.orc_spawn_score.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      ct, ct, C.c_void_p]
        L.orc_spawn_score.restype = ct
        self.bw_dtype = np.dtype([("b", np.int32, 3), ("f", self.real, 3), ("w", self.real, (3, 3))],
                                 align=True)
        self.drec_dtype = np.dtype([("pos", self.real, 4), ("vel", self.real, 4), ("ptype", np.uint32),
                                    ("alive", np.uint32), ("lifetime", self.real), ("pad", self.real)],
                                   align=True)


_LIBS: dict = {}


def lib(real: str = "f32") -> _Lib:
    """The f32 exact-semantics build or its f64 twin."""
    if real not in _LIBS:
        name = {"f32": "liborc32.so", "f64": "liborc64.so"}[real]
        path = os.path.join(_BUILD, name)
        if not os.path.exists(path):
            build()
        _LIBS[real] = _Lib(path, np.float32 if real == "f32" else np.float64)
    return _LIBS[real]


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Hand(C.Structure):
    """HandParams (shaders.js:394-402), box filled by orc_hand_bbox."""
    _fields_ = [("px", C.c_float), ("py", C.c_float), ("pz", C.c_float), ("radius", C.c_float),
                ("vx", C.c_float), ("vy", C.c_float), ("vz", C.c_float), ("is_on", C.c_uint32),
                ("strength", C.c_float), ("ex", C.c_float), ("ey", C.c_float), ("ez", C.c_float),
                ("bbox_min", C.c_int32 * 3), ("bbox_max", C.c_int32 * 3),
                ("pad0", C.c_int32), ("pad1", C.c_int32)]


class Ray(C.Structure):
    """RayParams (shaders.js:78)."""
    _fields_ = [("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("radius", C.c_float),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float), ("pad", C.c_float)]


def make_ray(origin, direction, radius) -> Ray:
    r = Ray()
    r.ox, r.oy, r.oz = (float(v) for v in origin)
    r.dx, r.dy, r.dz = (float(v) for v in direction)
    r.radius = float(radius)
    return r


class _State(C.Structure):
    _fields_ = [
        ("pos4", C.c_void_p), ("vel4", C.c_void_p), ("affine", C.c_void_p), ("mass", C.c_void_p),
        ("density", C.c_void_p), ("bw", C.c_void_p), ("cell_mass", C.c_void_p),
        ("cell_mv", C.c_void_p), ("cell_mass_f", C.c_void_p), ("cell_v", C.c_void_p),
        ("wallp", C.c_void_p), ("pool", C.c_void_p), ("free_list", C.c_void_p),
        ("top", C.c_void_p), ("gen", Gen), ("adv", Adv), ("hand", C.c_void_p),
    ]


@dataclass
class OracleScene:
    """One scene held in numpy arrays, stepped by the C restatement."""

    grid: tuple
    params: Params = None
    diffuse_max: int = 0
    real: str = "f32"
    seed: int = 1
    walls: np.ndarray | None = None
    gen: Gen = field(default=None)
    adv: Adv = field(default=None)
    hand: Hand = field(default=None)

    def __post_init__(self):
        gx, gy, gz = self.grid
        self.L = lib(self.real)
        R = self.L.real
        if self.params is None:
            self.params = default_params(gx, gy, gz)
        self.G = gx * gy * gz
        self.cell_mass = np.zeros(self.G, np.int32)
        self.cell_mv = np.zeros(3 * self.G, np.int32)
        self.cell_mass_f = np.zeros(self.G, R)
        self.cell_v = np.zeros(3 * self.G, R)
        if self.gen is None:
            self.gen = default_gen(self.seed)
        if self.adv is None:
            self.adv = default_adv(self.diffuse_max)
        D = self.diffuse_max
        if D:
            self.pool = np.zeros(D, self.L.drec_dtype)
            self.free_list = np.arange(D, dtype=np.uint32)
            self.top = np.array([D], np.int32)
        else:
            self.pool = self.free_list = self.top = None
        self.reset(np.zeros((0, 3), np.float32))

    def reset(self, pos, vel=None, C9=None, mass=None):
        """Load particles (host arrays copied), like fillBlock + initGPU upload."""
        R = self.L.real
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        self.n = n
        self.params.particle_num = n
        self.pos4 = np.zeros((n, 4), R)
        self.pos4[:, :3] = pos
        self.vel4 = np.zeros((n, 4), R)
        if vel is not None:
            self.vel4[:, :3] = np.asarray(vel, dtype=np.float64).reshape(-1, 3)
        self.affine = np.zeros((n, 9), R)
        if C9 is not None:
            self.affine[:] = np.asarray(C9, dtype=np.float64).reshape(-1, 9)
        self.mass = np.ones(n, R) if mass is None else np.asarray(mass, dtype=R).copy()
        self.density = np.zeros(n, R)
        self.bw = np.zeros(n, self.L.bw_dtype)

    def set_pool(self, pool: np.ndarray, free_list: np.ndarray, top: int):
        self.pool[:] = pool
        self.free_list[:] = free_list
        self.top[0] = top

    def _state(self) -> _State:
        s = _State()
        s.pos4, s.vel4, s.affine = _ptr(self.pos4), _ptr(self.vel4), _ptr(self.affine)
        s.mass, s.density, s.bw = _ptr(self.mass), _ptr(self.density), _ptr(self.bw)
        s.cell_mass, s.cell_mv = _ptr(self.cell_mass), _ptr(self.cell_mv)
        s.cell_mass_f, s.cell_v = _ptr(self.cell_mass_f), _ptr(self.cell_v)
        s.wallp = _ptr(self.walls)
        s.pool, s.free_list, s.top = _ptr(self.pool), _ptr(self.free_list), _ptr(self.top)
        s.gen, s.adv = self.gen, self.adv
        s.hand = C.cast(C.pointer(self.hand), C.c_void_p) if self.hand is not None else None
        return s

    # --- interaction (fluid-gpu.js:145-180, 479-638) --------------------------
    def set_hand(self, pos, vel, radius, strength, active, eye):
        """FluidGPU.updateHand: the push applies in every later step."""
        h = Hand()
        h.px, h.py, h.pz = (float(v) for v in pos)
        h.vx, h.vy, h.vz = (float(v) for v in vel)
        h.ex, h.ey, h.ez = (float(v) for v in eye)
        h.radius, h.strength, h.is_on = float(radius), float(strength), 1 if active else 0
        self.L.lib.orc_hand_bbox(C.byref(self.params), C.byref(h))
        self.hand = h

    def raycast(self, origin, direction, radius):
        """raycastFluid: (hit, t) of the nearest particle near the ray."""
        t = np.zeros(1, self.L.real)
        hit = self.L.lib.orc_raycast(_ptr(self.pos4), self.n,
                                     C.byref(make_ray(origin, direction, radius)), _ptr(t))
        return bool(hit), float(t[0])

    def delete_near(self, origin, direction, radius) -> int:
        """deleteNear: drop the particles within `radius` of the line; the
        survivors keep their order (and their C)."""
        keep = np.zeros(self.n, np.uint8)
        k = self.L.lib.orc_delete_mask(_ptr(self.pos4), self.n,
                                       C.byref(make_ray(origin, direction, radius)), _ptr(keep))
        m = keep.astype(bool)
        for name in ("pos4", "vel4", "affine", "mass", "density", "bw"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name)[m]))
        self.n = int(k)
        self.params.particle_num = self.n
        return self.n

    def append(self, pos, vel=None):
        """spawnParticles' upload: particles added after the current ones with
        unit mass and zero C."""
        R = self.L.real
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        k = pos.shape[0]
        p4 = np.zeros((k, 4), R)
        p4[:, :3] = pos
        v4 = np.zeros((k, 4), R)
        if vel is not None:
            v4[:, :3] = np.asarray(vel, dtype=np.float64).reshape(-1, 3)
        self.pos4 = np.concatenate([self.pos4, p4])
        self.vel4 = np.concatenate([self.vel4, v4])
        self.affine = np.concatenate([self.affine, np.zeros((k, 9), R)])
        self.mass = np.concatenate([self.mass, np.ones(k, R)])
        self.density = np.concatenate([self.density, np.zeros(k, R)])
        self.bw = np.concatenate([self.bw, np.zeros(k, self.L.bw_dtype)])
        self.n += k
        self.params.particle_num = self.n

    def step(self, nsteps: int = 1, threads: int = 0):
        if threads:
            self.L.lib.orc_set_threads(threads)
        st = self._state()
        self.L.lib.orc_step(C.byref(self.params), C.byref(st), C.c_uint32(nsteps))
        self.gen.seed = st.gen.seed

    # --- single kernels, for unit tests -------------------------------------
    def k_clear(self):
        self.L.lib.orc_clear(C.byref(self.params), _ptr(self.cell_mv), _ptr(self.cell_mass))

    def k_weights(self):
        self.L.lib.orc_compute_weights(C.byref(self.params), _ptr(self.pos4), _ptr(self.bw))

    def k_p2g_mass(self):
        self.L.lib.orc_p2g_mass(C.byref(self.params), _ptr(self.mass), _ptr(self.cell_mass), _ptr(self.bw))

    def k_decode_mass(self):
        self.L.lib.orc_decode_mass(C.byref(self.params), _ptr(self.cell_mass), _ptr(self.cell_mass_f))

    def k_p2g_mom(self):
        self.L.lib.orc_p2g_mom(C.byref(self.params), _ptr(self.vel4), _ptr(self.affine), _ptr(self.mass),
                               _ptr(self.cell_mass_f), _ptr(self.cell_mv), _ptr(self.density), _ptr(self.bw))

    def k_update(self):
        self.L.lib.orc_update_grid(C.byref(self.params), _ptr(self.cell_mv), _ptr(self.cell_mass),
                                   _ptr(self.cell_v), _ptr(self.walls))

    def k_g2p(self):
        self.L.lib.orc_g2p(C.byref(self.params), _ptr(self.pos4), _ptr(self.vel4), _ptr(self.affine),
                           _ptr(self.cell_v), _ptr(self.bw), _ptr(self.walls))

    def k_generate(self):
        self.L.lib.orc_diffuse_generate(C.byref(self.params), C.byref(self.gen), _ptr(self.pos4),
                                        _ptr(self.vel4), _ptr(self.mass), _ptr(self.density),
                                        _ptr(self.cell_mass_f), _ptr(self.pool), _ptr(self.free_list),
                                        _ptr(self.top), C.c_uint32(self.diffuse_max))

    def k_advect(self):
        self.L.lib.orc_diffuse_advect(C.byref(self.params), C.byref(self.adv), _ptr(self.pool),
                                      _ptr(self.cell_v), _ptr(self.free_list), _ptr(self.top),
                                      _ptr(self.walls))

    # --- views --------------------------------------------------------------
    def cell_v3(self) -> np.ndarray:
        return self.cell_v.reshape(-1, 3)

    def cell_mv3(self) -> np.ndarray:
        return self.cell_mv.reshape(-1, 3)


def encode(v: float, real: str = "f32") -> int:
    return lib(real).lib.orc_encode(v)


def decode(i: int, real: str = "f32") -> float:
    return lib(real).lib.orc_decode(i)


def diffuse_rand(seed: int) -> float:
    return lib("f32").lib.orc_diffuse_rand(seed)


# --- single helpers of the kernels, for the closed-form pins ---------------
def make_wall(center, lmin, lmax, yaw=0.0, pitch=0.0, vel=(0, 0, 0), omega=(0, 0, 0)) -> np.ndarray:
    """One wall's 20 floats (shaders.js:123-129): center + cos(yaw), local min
    + sin(yaw), local max + cos(pitch), velocity + sin(pitch), omega."""
    w = np.zeros(20, np.float32)
    w[0:3], w[3] = center, np.cos(yaw)
    w[4:7], w[7] = lmin, np.sin(yaw)
    w[8:11], w[11] = lmax, np.cos(pitch)
    w[12:15], w[15] = vel, np.sin(pitch)
    w[16:19] = omega
    return w


def wall_params(*walls) -> np.ndarray:
    """The WallParams uniform: 4 header floats (count in [0]) + 4 walls."""
    wp = np.zeros(4 + 4 * 20, np.float32)
    wp[0] = len(walls)
    for i, w in enumerate(walls):
        wp[4 + 20 * i:24 + 20 * i] = w
    return wp


def _vec(v, R):
    return np.ascontiguousarray(np.asarray(v, dtype=R).reshape(3))


def wall_helper(name: str, v, wall, real: str = "f32") -> np.ndarray:
    """wall_to_local / wall_to_world_dir / wall_vel_at of one wall."""
    L = lib(real)
    out = np.zeros(3, L.real)
    w = np.ascontiguousarray(wall, dtype=np.float32)
    getattr(L.lib, "orc_" + name)(_ptr(_vec(v, L.real)), _ptr(w), _ptr(out))
    return out


def wall_sdf_normal(lp, mn, mx, real: str = "f32"):
    L = lib(real)
    n = np.zeros(3, L.real)
    d = L.lib.orc_wall_sdf_normal(_ptr(_vec(lp, L.real)), _ptr(_vec(mn, L.real)),
                                  _ptr(_vec(mx, L.real)), _ptr(n))
    return float(d), n


def estimate_normal(params: Params, cell_mass_f: np.ndarray, p, real: str = "f32"):
    L = lib(real)
    n = np.zeros(3, L.real)
    cm = np.ascontiguousarray(cell_mass_f, dtype=L.real)
    ok = L.lib.orc_estimate_normal(C.byref(params), _ptr(cm), _ptr(_vec(p, L.real)), _ptr(n))
    return bool(ok), n


def spawn_score(params: Params, gen: Gen, cell_mass_f: np.ndarray, pos, vel, mass: float,
                density: float, real: str = "f32"):
    """(score, {ke, trapped, crest, speed}) of one fluid particle."""
    L = lib(real)
    t = np.zeros(4, L.real)
    cm = np.ascontiguousarray(cell_mass_f, dtype=L.real)
    sc = L.lib.orc_spawn_score(C.byref(params), C.byref(gen), _ptr(cm), _ptr(_vec(pos, L.real)),
                               _ptr(_vec(vel, L.real)), mass, density, _ptr(t))
    return float(sc), {"ke": float(t[0]), "trapped": float(t[1]), "crest": float(t[2]),
                       "speed": float(t[3])}
