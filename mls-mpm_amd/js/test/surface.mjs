// Host-surface check of FluidHIP / DiffuseHIP that needs no GPU: prints one
// JSON object (grid, defaults, fillBlock counts, the error initGPU raises).
import { FluidHIP } from '../fluid-hip.mjs';
import { DiffuseHIP } from '../diffuse-hip.mjs';

const out = {};
// main.js:21-23 desktop preset
const f = new FluidHIP(2, 2, 3, 0.0125, 400000);
out.grid = [f.grid_X_num, f.grid_Y_num, f.grid_Z_num];
out.defaults = { DT: f.DT, GRAVITY: f.GRAVITY, REST_DENSITY: f.REST_DENSITY,
                 STIFFNESS: f.STIFFNESS, EOS_POWER: f.EOS_POWER, VISCOSITY: f.VISCOSITY,
                 HARD_MIN: f.HARD_MIN, HARD_MAX: [f.HARD_MAX_X, f.HARD_MAX_Y, f.HARD_MAX_Z],
                 WALL_MIN: f.WALL_MIN, WALL_STIFFNESS: f.WALL_STIFFNESS,
                 LOOKAHEAD_K: f.LOOKAHEAD_K, SUBSTEPS: f.SUBSTEPS };
const spacing = 0.05;
out.filled = f.fillBlock(spacing, spacing, spacing, 1 - spacing, 1 - spacing, 0.4 - spacing);
out.mass1 = f.particle_mass.subarray(0, out.filled).every((m) => m === 1);
const d = new DiffuseHIP(1 << 16);
out.diffuse = { maxCount: d.maxCount, KE_THRESHOLD: d.KE_THRESHOLD, LIFETIME_FOAM: d.LIFETIME_FOAM,
                BUOYANCY: d.BUOYANCY, DRAG: d.DRAG, seed: d._seed };
try {
    f.initGPU(0);
    out.init = 'ok';
} catch (e) {
    out.init = { code: e.code, message: e.message };
}
out.simFrame = f.simFrame();
// mobile preset (main.js:19-21): the upload moves fillBlock's last column
// inside [1, grid - 1); spawnParticles before initGPU fills the host arrays
const m = new FluidHIP(2, 2, 3, 0.0175, 80000);
let seed = 9;
m.random = () => { seed = (Math.imul(seed, 1664525) + 1013904223) >>> 0; return seed / 4294967296; };
const nm = m.fillBlock(0.03, 0.03, 0.03, 0.97, 0.97, 0.37);
const up = m._inGrid(m.particle_pos.subarray(0, 3 * nm));
let moved = 0;
const mx = [0, 0, 0], mn = [1e9, 1e9, 1e9];
for (let i = 0; i < up.length; i++) {
    if (up[i] !== m.particle_pos[i]) moved++;
    mx[i % 3] = Math.max(mx[i % 3], up[i]);
    mn[i % 3] = Math.min(mn[i % 3], up[i]);
}
const m2 = new FluidHIP(2, 2, 3, 0.0175, 1000);
m2.random = m.random;
const sp = m2.spawnParticles(10, 10, 10, 400, 2.0, [0, -1, 0], 3.0);
let inside = true;
for (let i = 0; i < 3 * sp; i++) {
    const v = m2.particle_pos[i], a = i % 3;
    const hiA = [m2.HARD_MAX_X, m2.HARD_MAX_Y, m2.HARD_MAX_Z][a];
    if (!(v >= m2.HARD_MIN && v <= hiA)) inside = false;
}
out.mobile = { grid: [m.grid_X_num, m.grid_Y_num, m.grid_Z_num], filled: nm, moved, max: mx, min: mn,
               spawned: sp, spawnedInside: inside, active: m2.active_particle_num,
               massOne: m2.particle_mass.subarray(0, sp).every((x) => x === 1) };
out.lastError = f.lastError ? String(f.lastError.code) : null;
console.log(JSON.stringify(out));
