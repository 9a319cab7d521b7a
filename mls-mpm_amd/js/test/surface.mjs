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
out.lastError = f.lastError ? String(f.lastError.code) : null;
console.log(JSON.stringify(out));
