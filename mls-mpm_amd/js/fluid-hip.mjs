// fluid-hip.mjs -- FluidHIP: the FluidGPU host surface (fluid-gpu.js) over the
// HIP library, through the Node-API addon (addon/mlsmpm.node).
//
// Same construction, tunables and methods as FluidGPU, so main.js-style code
// keeps working: new FluidHIP(aspectX, aspectY, aspectZ, particleRadius,
// particleNum), the public DT ... LOOKAHEAD_K / HARD_* fields, fillBlock(),
// setWall() / rotateWall() / removeWall(), initGPU(), simFrame().  Differences
// a caller sees:
//   - initGPU(device) takes a HIP device ordinal (default 0), not a
//     GPUDevice; simFrame() takes no command encoder and runs SUBSTEPS steps.
//   - The renderer's buffers are read back explicitly: readPositions(),
//     readVelocities(), readAffine(), readDensity() (stride-16 vec4 layouts of
//     particlePosBuffer / particleVelBuffer).
//   - Like the reference (errors go to WebGPU error scopes, methods return
//     bool), simFrame() returns false on a device error and reports it to
//     this.onerror (if set) and this.lastError; initGPU() throws.
//   - deleteNear() and raycastFluid() accept the reference's leading command
//     encoder argument and ignore it; they complete before returning, so
//     active_particle_num and lastRayHit are current at once and
//     pollDelete() / pollRaycast() have nothing left to do.
//   - deleteNear() keeps the survivors in their order and carries their
//     affine momentum; spawned particles start with zero affine momentum.
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
const addon = require('./addon/mlsmpm.node');

export const MAX_WALLS = 4;
// Parts of a frame (mpm_frame_begin): bit flags.
export const FRAME = { POS: 1, VEL: 2, DENSITY: 4, DIFFUSE: 8, ALL: 15, RESET_ORDER: 16 };

export class FluidHIP {
    constructor(aspectX, aspectY, aspectZ, particleRadius, particleNum) {
        // grid = ceil(aspect / (4 r)) per axis, reference defaults
        const c = addon.configFromAspect(aspectX, aspectY, aspectZ, particleRadius, particleNum);
        this._cfg = c;
        this.grid_X_num = c.grid_x;
        this.grid_Y_num = c.grid_y;
        this.grid_Z_num = c.grid_z;
        this.grid_num = c.grid_x * c.grid_y * c.grid_z;
        this.particle_num = particleNum;
        this.active_particle_num = 0;

        this.particle_pos = new Float32Array(particleNum * 3);
        this.particle_vel = new Float32Array(particleNum * 3);
        this.particle_affine = new Float32Array(particleNum * 9);
        this.particle_mass = new Float32Array(particleNum);

        // JavaScript numbers, as in the reference; rounded to f32 on upload
        this.DT = 0.3;
        this.REST_DENSITY = 8.0;
        this.STIFFNESS = 100.0;
        this.EOS_POWER = 1;
        this.VISCOSITY = 0.01;
        this.GRAVITY = -0.98;
        this.SUBSTEPS = 1;
        this.WALL_MIN = 3.0;
        this.WALL_STIFFNESS = 1.0;
        this.LOOKAHEAD_K = 2.0;
        this.HARD_MIN = 2;
        this.HARD_MAX_X = this.grid_X_num - 2;
        this.HARD_MAX_Y = this.grid_Y_num - 2;
        this.HARD_MAX_Z = this.grid_Z_num - 2;

        this.diffuse = null;
        this.walls = [];
        this._wallPrev = [];
        for (let i = 0; i < MAX_WALLS; i++) {
            this.walls.push({ active: false, min: new Float32Array(3), max: new Float32Array(3),
                              theta: 0, phi: 0 });
            this._wallPrev.push({ min: new Float32Array(3), theta: 0, phi: 0 });
        }
        this._wallF32 = new Float32Array(4 + MAX_WALLS * 20);
        this._wallSent = null;

        this.random = Math.random;  // jitter source of fillBlock / spawnParticles (seedable)
        this.handActive = false;
        this._hand = null;          // last updateHand arguments, re-sent on a rebuild
        this.lastRayHit = null;
        this._deletePending = false;
        this._frameTicket = null;
        this.device = null;
        this.onerror = null;
        this.lastError = null;
        this._scene = null;
        this._sent = null;          // tunables last pushed to the scene
    }

    // Lattice every 0.5 cells from min + 0.25, jitter (rand - 0.5) * 0.05, mass
    // 1; bounds normalised 0..1 per axis; stacks from active_particle_num.
    fillBlock(x0, y0, z0, x1, y1, z1) {
        const lo = [x0 * this.grid_X_num, y0 * this.grid_Y_num, z0 * this.grid_Z_num];
        const hi = [x1 * this.grid_X_num, y1 * this.grid_Y_num, z1 * this.grid_Z_num];
        let p = this.active_particle_num;
        const rnd = this.random;
        fill: for (let z = lo[2] + 0.25; z < hi[2]; z += 0.5) {
            for (let y = lo[1] + 0.25; y < hi[1]; y += 0.5) {
                for (let x = lo[0] + 0.25; x < hi[0]; x += 0.5) {
                    if (p >= this.particle_num) break fill;
                    const o = 3 * p;
                    this.particle_pos[o] = x + (rnd() - 0.5) * 0.05;
                    this.particle_pos[o + 1] = y + (rnd() - 0.5) * 0.05;
                    this.particle_pos[o + 2] = z + (rnd() - 0.5) * 0.05;
                    this.particle_mass[p] = 1.0;
                    p++;
                }
            }
        }
        this.active_particle_num = p;
        return p;
    }

    // ---- moving walls (the wall velocity and angular velocity are derived
    // from the change since the previous simFrame) ----
    setWall(i, min, max) {
        const w = this.walls[i];
        if (!w.active) {  // appearing is not moving: no velocity this frame
            const pv = this._wallPrev[i];
            pv.min.set([min[0], min[1], min[2]]);
            pv.theta = w.theta;
            pv.phi = w.phi;
        }
        w.active = true;
        w.min.set([min[0], min[1], min[2]]);
        w.max.set([max[0], max[1], max[2]]);
    }

    rotateWall(i, dTheta, dPhi = 0) {
        const w = this.walls[i];
        if (!w.active) return;
        w.theta += dTheta;
        w.phi += dPhi;
    }

    removeWall(i) {
        this.walls[i].active = false;
    }

    resetWallTheta(i) {
        this.walls[i].theta = 0;
        this.walls[i].phi = 0;
    }

    // WallParams: active walls packed first, count in the header; each wall
    // centre, cos/sin yaw, local box, cos/sin pitch, linear and angular
    // velocity.  Faces lying in the boundary band are pushed out of the grid
    // when the wall is axis-aligned (so a wall standing on the floor has no
    // bottom face inside the fluid).
    _packWalls() {
        const f = this._wallF32;
        const inv = 1.0 / this.DT;
        const band = this.HARD_MIN + 1.0;
        let n = 0;
        for (let i = 0; i < MAX_WALLS; i++) {
            const w = this.walls[i], pv = this._wallPrev[i];
            let v = [0, 0, 0], dT = 0, dP = 0;
            if (w.active) {
                v = [0, 1, 2].map((a) => (w.min[a] - pv.min[a]) * inv);
                dT = (w.theta - pv.theta) * inv;
                dP = (w.phi - pv.phi) * inv;
            }
            pv.min.set(w.min);
            pv.theta = w.theta;
            pv.phi = w.phi;
            if (!w.active) continue;
            const c = [0, 1, 2].map((a) => (w.min[a] + w.max[a]) * 0.5);
            const cy = Math.cos(w.theta), sy = Math.sin(w.theta);
            const cp = Math.cos(w.phi), sp = Math.sin(w.phi);
            const lmin = [0, 1, 2].map((a) => w.min[a] - c[a]);
            const lmax = [0, 1, 2].map((a) => w.max[a] - c[a]);
            const hardMax = [this.HARD_MAX_X, this.HARD_MAX_Y, this.HARD_MAX_Z];
            const gridN = [this.grid_X_num, this.grid_Y_num, this.grid_Z_num];
            const extend = (a) => {
                if (w.min[a] <= band) lmin[a] = -4.0 - c[a];
                if (w.max[a] >= hardMax[a] - 1.0) lmax[a] = gridN[a] + 4.0 - c[a];
            };
            if (Math.abs(sp) < 0.05 && cp > 0.95) {
                extend(1);
                if (Math.abs(sy) < 0.05 && cy > 0.95) {
                    extend(0);
                    extend(2);
                }
            }
            const b = 4 + n * 20;
            f.set([c[0], c[1], c[2], cy, lmin[0], lmin[1], lmin[2], sy,
                   lmax[0], lmax[1], lmax[2], cp, v[0], v[1], v[2], sp,
                   dP * cy, dT, -dP * sy, 0.0], b);
            n++;
        }
        f[0] = n;
        if (this._wallSent && this._wallSent.every((x, k) => x === f[k])) return;
        this._wallSent = Float32Array.from(f);
        addon.setWalls(this._scene, n ? f : null);
    }

    _config() {
        const c = Object.assign({}, this._cfg);
        c.capacity = this.particle_num;
        // the substep's dt, as simFrame packs it (fluid-gpu.js:649-651)
        c.dt = Math.fround(this.DT / this.SUBSTEPS);
        c.gravity = this.GRAVITY;
        c.rest_density = this.REST_DENSITY;
        c.stiffness = this.STIFFNESS;
        c.eos_power = this.EOS_POWER;
        c.viscosity = this.VISCOSITY;
        c.hard_min = this.HARD_MIN;
        c.hard_max_x = this.HARD_MAX_X;
        c.hard_max_y = this.HARD_MAX_Y;
        c.hard_max_z = this.HARD_MAX_Z;
        c.wall_min = this.WALL_MIN;
        c.wall_stiffness = this.WALL_STIFFNESS;
        c.lookahead_k = this.LOOKAHEAD_K;
        c.device = this.device === null ? 0 : this.device;
        const d = this.diffuse;
        c.diffuse_capacity = d ? d.maxCount : 0;
        if (d) {
            c.ke_threshold = d.KE_THRESHOLD;
            c.pressure_threshold = d.PRESSURE_THRESHOLD;
            c.crest_dot_threshold = d.CREST_DOT_THRESHOLD;
            c.spawn_rate_k = d.SPAWN_RATE_K;
            c.lifetime_foam = d.LIFETIME_FOAM;
            c.buoyancy = d.BUOYANCY;
            c.drag = d.DRAG;
            c.density_spray_max = d.DENSITY_SPRAY_MAX;
            c.density_bubble_min = d.DENSITY_BUBBLE_MIN;
            c.seed = d._seed;
        }
        return c;
    }

    // Creates the device scene and uploads the active particles (throws on
    // failure, with .code = the MPM_ERR_* value).
    initGPU(device = 0) {
        this.device = device;
        this._create(null);
    }

    _create(state) {
        if (this._scene) addon.destroy(this._scene);
        this._scene = null;
        const cfg = this._config();
        this._scene = addon.create(cfg);
        this._sent = cfg;
        this._wallSent = null;
        this._frameTicket = null;   // frames belonged to the previous scene
        const n = this.active_particle_num;
        const hand = this._hand;
        if (state) {
            addon.reset(this._scene, state.pos, state.vel, state.C, this.particle_mass.subarray(0, n));
        } else {
            addon.reset(this._scene, this._inGrid(this.particle_pos.subarray(0, 3 * n)),
                        this.particle_vel.subarray(0, 3 * n),
                        this.particle_affine.subarray(0, 9 * n), this.particle_mass.subarray(0, n));
        }
        if (hand) this.updateHand(...hand);
    }

    // The scene takes particles whose 3x3x3 stencil lies inside the grid:
    // coordinates in [1, grid - 1).  fillBlock's last lattice column can lie
    // beyond grid - 1 (the mobile preset: 0.97 * 29 cells); the reference's
    // first P2G then writes that column's outer stencil cells into the next
    // row of its linear grid, and its G2P clamps the particles to the hard
    // bounds.  Here such coordinates move to the largest float below
    // grid - 1 before the upload.
    _inGrid(pos) {
        const lim = [this.grid_X_num - 1, this.grid_Y_num - 1, this.grid_Z_num - 1];
        const f = new Float32Array(1), u = new Uint32Array(f.buffer);
        const hi = lim.map((l) => { f[0] = l; u[0] -= 1; return f[0]; });
        let out = pos;
        for (let i = 0; i < pos.length; i++) {
            const a = i % 3;
            if (pos[i] >= 1 && pos[i] <= hi[a]) continue;
            if (out === pos) out = Float32Array.from(pos);
            out[i] = Math.min(hi[a], Math.max(1, pos[i]));
        }
        return out;
    }

    // A diffuse pool attached after initGPU: the scene is rebuilt with it,
    // keeping the particles' current state.
    _attachDiffuse(d) {
        this.diffuse = d;
        if (!this._scene) return;
        const n = this.active_particle_num;
        const p4 = new Float32Array(4 * n), v4 = new Float32Array(4 * n), C = new Float32Array(9 * n);
        addon.readParticles(this._scene, p4, v4, C, null);
        const pos = new Float32Array(3 * n), vel = new Float32Array(3 * n);
        for (let i = 0; i < n; i++) {
            for (let a = 0; a < 3; a++) {
                pos[3 * i + a] = p4[4 * i + a];
                vel[3 * i + a] = v4[4 * i + a];
            }
        }
        this._create({ pos, vel, C });
    }

    _pushParams() {
        const c = this._config();
        let same = true;
        for (const k of Object.keys(c)) if (c[k] !== this._sent[k]) { same = false; break; }
        if (same) return;
        addon.setParams(this._scene, c);
        this._sent = c;
    }

    // One frame: SUBSTEPS whole steps of DT / SUBSTEPS each (P2G, grid
    // update, hand, G2P, diffuse).
    simFrame() {
        try {
            if (!this._scene) this.initGPU(this.device === null ? 0 : this.device);
            this._pushParams();
            this._packWalls();
            addon.step(this._scene, this.SUBSTEPS);
            return true;
        } catch (e) {
            this.lastError = e;
            if (this.onerror) this.onerror(e);
            return false;
        }
    }

    step(n = 1) {
        if (!this._scene) this.initGPU(this.device === null ? 0 : this.device);
        this._pushParams();
        this._packWalls();
        addon.step(this._scene, n);
    }

    sync() {
        if (this._scene) addon.sync(this._scene);
    }

    readPositions(out = null) {
        const o = out || new Float32Array(4 * this.active_particle_num);
        addon.readParticles(this._scene, o, null, null, null);
        return o;
    }

    readVelocities(out = null) {
        const o = out || new Float32Array(4 * this.active_particle_num);
        addon.readParticles(this._scene, null, o, null, null);
        return o;
    }

    readAffine(out = null) {
        const o = out || new Float32Array(9 * this.active_particle_num);
        addon.readParticles(this._scene, null, null, o, null);
        return o;
    }

    readDensity(out = null) {
        const o = out || new Float32Array(this.active_particle_num);
        addon.readParticles(this._scene, null, null, null, o);
        return o;
    }

    // -- interaction (fluid-gpu.js:145-180, 479-638) ------------------------

    // Spray `count` particles (at most the free capacity) uniformly in a ball
    // of `radius` around (cx, cy, cz), clamped to the hard bounds, with
    // velocity vel + sprayVel along the offset.  Returns how many were added.
    spawnParticles(cx, cy, cz, count, radius, vel = [0, 0, 0], sprayVel = 3.0) {
        if (this._deletePending) return 0;
        const start = this.active_particle_num;
        const n = Math.min(count, this.particle_num - start);
        if (n <= 0) return 0;
        const pos = new Float32Array(3 * n), v = new Float32Array(3 * n);
        const lo = this.HARD_MIN, hi = [this.HARD_MAX_X, this.HARD_MAX_Y, this.HARD_MAX_Z];
        const c = [cx, cy, cz];
        for (let i = 0; i < n; i++) {
            const r = radius * Math.cbrt(this.random());
            const theta = this.random() * Math.PI * 2;
            const cphi = this.random() * 2 - 1;
            const sphi = Math.sqrt(1 - cphi * cphi);
            const d = [r * sphi * Math.cos(theta), r * sphi * Math.sin(theta), r * cphi];
            const rl = Math.hypot(d[0], d[1], d[2]) || 1;
            for (let a = 0; a < 3; a++) {
                pos[3 * i + a] = Math.min(hi[a], Math.max(lo, c[a] + d[a]));
                v[3 * i + a] = vel[a] + d[a] / rl * sprayVel;
            }
        }
        if (this._scene) {
            addon.append(this._scene, pos, v);
        } else {
            this.particle_pos.set(pos, 3 * start);
            this.particle_vel.set(v, 3 * start);
            this.particle_affine.fill(0, 9 * start, 9 * (start + n));
            this.particle_mass.fill(1, start, start + n);
        }
        this.active_particle_num = start + n;
        return n;
    }

    // The push of cells within `radius` of the line from `eye` through `pos`:
    // vel * (1 - d / radius)^2 * strength, added after the grid update.
    updateHand(pos, vel, radius, strength, active, eye) {
        this.handActive = !!active;
        this._hand = [pos, vel, radius, strength, active, eye];
        if (!this._scene) return;
        if (active) addon.setHand(this._scene, pos, vel, radius, strength, true, eye);
        else addon.setHand(this._scene, null);
    }

    // Removes the particles within `radius` of the line through `origin`
    // along `dir`.  Returns true (the removal is complete on return).
    deleteNear(...args) {
        const [origin, dir, radius] = args.length > 3 ? args.slice(1) : args;
        if (!this._scene) this.initGPU(this.device === null ? 0 : this.device);
        this.active_particle_num = addon.deleteNear(this._scene, origin, dir, radius);
        return true;
    }

    pollDelete() {}

    // Where the ray first meets the fluid within `radius` of its axis:
    // this.lastRayHit = [x, y, z] or null.  Returns true.
    raycastFluid(...args) {
        const [origin, dir, radius] = args.length > 3 ? args.slice(1) : args;
        if (!this._scene) this.initGPU(this.device === null ? 0 : this.device);
        const t = addon.raycast(this._scene, origin, dir, radius);
        this.lastRayHit = t === null ? null
            : [origin[0] + t * dir[0], origin[1] + t * dir[1], origin[2] + t * dir[2]];
        return true;
    }

    pollRaycast() {}

    // -- renderer handoff (fluid-renderer.js:232, 256, 315-316, 850-856) ---

    // Start copying the renderer's buffers (FRAME bits) into pinned memory;
    // the copy overlaps the simFrame() calls that follow.
    beginFrame(what = FRAME.ALL) {
        if (!this.diffuse) what &= ~FRAME.DIFFUSE;
        this._frameTicket = addon.frameBegin(this._scene, what);
        return this._frameTicket;
    }

    // The frame begun last, copied into `out` ({pos4, vel4, density,
    // diffuse}: arrays sized for the frame, any may be missing), or null
    // while its copy still runs (wait = true blocks instead).
    pollFrame(out = {}, wait = false) {
        if (!this._frameTicket) return null;
        const r = addon.framePoll(this._scene, this._frameTicket, wait, out.pos4 || null,
                                  out.vel4 || null, out.density || null, out.diffuse || null);
        if (r) Object.assign(r, out);
        return r;
    }

    stepsTaken() {
        return this._scene ? addon.counts(this._scene).steps : 0;
    }

    destroy() {
        if (this._scene) addon.destroy(this._scene);
        this._scene = null;
    }
}

export { addon };
