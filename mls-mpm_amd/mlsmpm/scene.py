"""Python handle of one scene of libmlsmpm.so (ctypes).

Method names follow the reference host classes: the scene is created from
a config (FluidGPU / DiffuseGPU constructors, fluid-gpu.js:10-45,
diffuse-gpu.js:16-31), loaded with reset() (fillBlock + initGPU upload,
fluid-gpu.js:107-128, 278-281), advanced with step() (simFrame,
fluid-gpu.js:648-684) and read back in the renderer's layouts.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import Config, check, load

# DiffuseParticle, 48 bytes (shaders.js:725-732).
DREC_DTYPE = np.dtype([("pos", np.float32, 4), ("vel", np.float32, 4), ("ptype", np.uint32),
                       ("alive", np.uint32), ("lifetime", np.float32), ("pad", np.float32)])
assert DREC_DTYPE.itemsize == 48


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a, shape):
    if a is None:
        return None
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(shape))


class Scene:
    def __init__(self, cfg: Config):
        L = load()
        self.cfg = cfg.copy()
        self._h = C.c_void_p()
        check(L.mpm_create(C.byref(self.cfg), C.byref(self._h)))

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if self._h:
            load().mpm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # -- state ---------------------------------------------------------------
    def reset(self, pos, vel=None, C9=None, mass=None):
        pos = _f32(pos, (-1, 3))
        n = pos.shape[0]
        vel = _f32(vel, (n, 3))
        C9 = _f32(C9, (n, 9))
        mass = _f32(mass, (n,))
        check(load().mpm_reset(self._h, n, _p(pos), _p(vel), _p(C9), _p(mass)))

    def step(self, nsteps: int = 1):
        check(load().mpm_step(self._h, nsteps))

    def prepare(self):
        """Capture the graphs the next multi-step call replays (mpm_prepare)."""
        check(load().mpm_prepare(self._h))

    def sync(self):
        check(load().mpm_sync(self._h))

    def counts(self):
        n = C.c_uint32()
        st = C.c_uint64()
        check(load().mpm_get_counts(self._h, C.byref(n), C.byref(st)))
        return n.value, st.value

    @property
    def n(self) -> int:
        return self.counts()[0]

    def read_particles(self, pos=True, vel=True, C9=True, density=True) -> dict:
        n = self.n
        out = {
            "pos4": np.empty((n, 4), np.float32) if pos else None,
            "vel4": np.empty((n, 4), np.float32) if vel else None,
            "C9": np.empty((n, 9), np.float32) if C9 else None,
            "density": np.empty(n, np.float32) if density else None,
        }
        check(load().mpm_read_particles(self._h, _p(out["pos4"]), _p(out["vel4"]), _p(out["C9"]),
                                        _p(out["density"])))
        return {k: v for k, v in out.items() if v is not None}

    def read_grid(self, ints=True) -> dict:
        G = self.cfg.grid_x * self.cfg.grid_y * self.cfg.grid_z
        out = {
            "cell_mass": np.empty(G, np.int32) if ints else None,
            "cell_mv": np.empty(3 * G, np.int32) if ints else None,
            "cell_mass_f": np.empty(G, np.float32),
            "cell_v": np.empty(3 * G, np.float32),
        }
        check(load().mpm_read_grid(self._h, _p(out["cell_mass"]), _p(out["cell_mv"]),
                                   _p(out["cell_mass_f"]), _p(out["cell_v"])))
        return {k: v for k, v in out.items() if v is not None}

    def read_diffuse(self):
        D = self.cfg.diffuse_capacity
        rec = np.empty(D, DREC_DTYPE)
        fl = np.empty(D, np.uint32)
        top = C.c_int32()
        check(load().mpm_read_diffuse(self._h, _p(rec), _p(fl), C.byref(top)))
        return rec, fl, top.value

    def write_diffuse(self, rec, free_list, top: int):
        rec = np.ascontiguousarray(rec, dtype=DREC_DTYPE)
        fl = np.ascontiguousarray(free_list, dtype=np.uint32)
        check(load().mpm_write_diffuse(self._h, _p(rec), _p(fl), int(top)))

    def diffuse_alive(self) -> int:
        a = C.c_uint32()
        check(load().mpm_diffuse_alive(self._h, C.byref(a)))
        return a.value

    def set_params(self, **kw):
        c = self.get_config()
        for k, v in kw.items():
            setattr(c, k, v)
        check(load().mpm_set_params(self._h, C.byref(c)))
        self.cfg = c

    def get_config(self) -> Config:
        c = Config()
        check(load().mpm_get_config(self._h, C.byref(c)))
        return c

    def set_seed(self, seed: int):
        check(load().mpm_set_seed(self._h, seed))

    def set_walls(self, wallp):
        if wallp is None:
            check(load().mpm_set_walls(self._h, None))
            return
        w = np.ascontiguousarray(np.asarray(wallp, dtype=np.float32).reshape(84))
        check(load().mpm_set_walls(self._h, _p(w)))

    # -- interaction (fluid-gpu.js:145-180, 479-638) ----------------------
    def set_hand(self, pos, vel, radius: float, strength: float, active: bool = True, eye=None):
        """updateHand: push the fluid along the line from `eye` through `pos`
        (cells within `radius` of it get vel * (1 - d/radius)^2 * strength
        added after the grid update).  active=False switches it off."""
        h = capi.Hand()
        h.pos[:] = [float(x) for x in pos]
        h.vel[:] = [float(x) for x in vel]
        h.eye[:] = [float(x) for x in (eye if eye is not None else (0.0, 0.0, 0.0))]
        h.radius = float(radius)
        h.strength = float(strength)
        h.active = 1 if active else 0
        check(load().mpm_set_hand(self._h, C.byref(h)))

    def clear_hand(self):
        check(load().mpm_set_hand(self._h, None))

    @staticmethod
    def _ray(origin, direction):
        o = np.ascontiguousarray(np.asarray(origin, np.float32).reshape(3))
        d = np.ascontiguousarray(np.asarray(direction, np.float32).reshape(3))
        return o, d

    def delete_near(self, origin, direction, radius: float) -> int:
        """deleteNear: removes the particles within `radius` of the ray's
        axis (direction taken as given, normally unit length); returns the
        number left.  Survivors keep their relative order."""
        o, d = self._ray(origin, direction)
        left = C.c_uint32()
        check(load().mpm_delete_near(self._h, _p(o), _p(d), float(radius), C.byref(left)))
        return left.value

    def raycast(self, origin, direction, radius: float):
        """raycastFluid: (hit, t) with t the smallest positive distance along
        the ray among particles within `radius` of its axis."""
        o, d = self._ray(origin, direction)
        t = C.c_float()
        hit = C.c_int()
        check(load().mpm_raycast(self._h, _p(o), _p(d), float(radius), C.byref(t), C.byref(hit)))
        return bool(hit.value), float(t.value)

    def raycast_noalloc(self, origins, directions, radii):
        """mpm_raycast for many rays in one call, through the library's
        static scratch (no device allocation): (hit bool array, t array)."""
        o = np.ascontiguousarray(np.asarray(origins, np.float32).reshape(-1, 3))
        d = np.ascontiguousarray(np.asarray(directions, np.float32).reshape(-1, 3))
        k = o.shape[0]
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(radii, np.float32), (k,)))
        t = np.zeros(k, np.float32)
        hit = np.zeros(k, np.int32)
        check(load().mpm_raycast_noalloc(self._h, k, _p(o), _p(d), _p(r), _p(t), _p(hit)))
        return hit.astype(bool), t

    def append(self, pos, vel=None):
        """spawnParticles' upload: new particles after the current ones, unit
        mass, zero affine momentum."""
        pos = _f32(pos, (-1, 3))
        k = pos.shape[0]
        vel = _f32(vel, (k, 3))
        check(load().mpm_append_particles(self._h, k, _p(pos), _p(vel)))

    # -- renderer handoff (fluid-renderer.js:232, 256, 315-316, 850-856) ----
    def export_device(self, pos4=0, vel4=0, density=0, diffuse48=0, reset_order=False):
        """Enqueue the renderer layouts into device buffers (integer device
        addresses, 0 = skip) on the scene's stream; returns at once."""
        e = capi.Export(pos4 or None, vel4 or None, density or None, diffuse48 or None)
        check(load().mpm_export_device(self._h, C.byref(e),
                                       capi.MPM_EXPORT_RESET_ORDER if reset_order else 0))

    def frame_begin(self, what: int = capi.MPM_EXPORT_ALL, reset_order=False) -> int:
        """Take a frame into pinned host memory; the copy overlaps later steps."""
        t = C.c_uint64()
        if reset_order:
            what |= capi.MPM_EXPORT_RESET_ORDER
        check(load().mpm_frame_begin(self._h, what, C.byref(t)))
        return t.value

    def frame_poll(self, ticket: int, wait: bool = True, copy: bool = True):
        """None while the frame's copy runs, else a dict: pos4 / vel4 (n, 4),
        density (n,), diffuse (records), n, top, steps.  copy=False gives
        views of the pinned memory, valid until the second frame_begin after
        this frame's."""
        ready = C.c_int()
        f = capi.Frame()
        check(load().mpm_frame_poll(self._h, ticket, 1 if wait else 0, C.byref(ready),
                                    C.byref(f)))
        if not ready.value:
            return None
        n = f.n

        def view(ptr, count, dtype, shape):
            if not ptr:
                return None
            a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint8)),
                                      (count * np.dtype(dtype).itemsize,))
            a = a.view(dtype).reshape(shape)
            return a.copy() if copy else a

        D = self.cfg.diffuse_capacity
        return {"pos4": view(f.pos4, 4 * n, np.float32, (n, 4)),
                "vel4": view(f.vel4, 4 * n, np.float32, (n, 4)),
                "density": view(f.density, n, np.float32, (n,)),
                "diffuse": view(f.diffuse48, D, DREC_DTYPE, (D,)),
                "n": n, "top": f.top, "steps": f.steps, "ticket": f.ticket}

    def enable_timing(self, on: bool = True):
        check(load().mpm_enable_timing(self._h, 1 if on else 0))

    def stage_times(self) -> np.ndarray:
        ms = np.zeros(8, np.float32)
        check(load().mpm_stage_times(self._h, _p(ms)))
        return ms

    def step_times(self) -> np.ndarray:
        cnt = C.c_uint32()
        check(load().mpm_step_times(self._h, None, 0, C.byref(cnt)))
        ms = np.zeros(max(cnt.value, 1), np.float32)
        check(load().mpm_step_times(self._h, _p(ms), cnt.value, C.byref(cnt)))
        return ms[: cnt.value]

    def tile_stats(self) -> dict:
        o = np.zeros(4, np.uint64)
        check(load().mpm_tile_stats(self._h, _p(o)))
        return {"items": int(o[0]), "bricks": int(o[1]), "particles": int(o[2]),
                "max_item": int(o[3])}

    def stream(self) -> int:
        p = C.c_void_p()
        check(load().mpm_get_stream(self._h, C.byref(p)))
        return p.value or 0


def default_config(gx, gy, gz, capacity, **kw) -> Config:
    return capi.default_config(gx, gy, gz, capacity, **kw)
