// diffuse-hip.mjs -- DiffuseHIP: the DiffuseGPU host surface (diffuse-gpu.js)
// over the HIP library.  The pool lives inside the fluid's scene (the step
// runs generate + advect right after G2P, as simFrame does with
// fluid.diffuse), so initGPU(device, fluid) attaches it to the fluid and
// step() has nothing to record.  particleBuffer is read back explicitly with
// readParticles() in the 48-byte DiffuseParticle layout the renderer binds.
import { addon } from './fluid-hip.mjs';

export const PARTICLE_BYTES = 48;

export class DiffuseHIP {
    constructor(maxCount) {
        this.maxCount = maxCount;
        this.KE_THRESHOLD = 0.5;
        this.PRESSURE_THRESHOLD = 1.0;
        this.CREST_DOT_THRESHOLD = 0.3;
        this.SPAWN_RATE_K = 3.0;
        this.LIFETIME_FOAM = 3.0;
        this.BUOYANCY = 1.2;
        this.DRAG = 4.0;
        this.DENSITY_SPRAY_MAX = 1.5;
        this.DENSITY_BUBBLE_MIN = 6.0;
        this.aliveCount = 0;
        this._seed = 1;        // advanced on the device before every generate
        this._debugPending = false;
        this.fluid = null;
    }

    initGPU(device, fluid) {
        this.device = device;
        this.fluid = fluid;
        fluid._attachDiffuse(this);
    }

    // Generate + advect run inside fluid.simFrame().
    step() {}

    // Copy the pool (maxCount * 48 bytes) into out (ArrayBuffer / TypedArray).
    readParticles(out = null) {
        const o = out || new ArrayBuffer(this.maxCount * PARTICLE_BYTES);
        const r = addon.readDiffuse(this.fluid._scene, o);
        this.aliveCount = r.alive;
        return o;
    }

    requestDebugReadback() {
        this._debugPending = true;
    }

    pollDebug() {
        if (!this._debugPending || !this.fluid || !this.fluid._scene) return;
        this.aliveCount = addon.readDiffuse(this.fluid._scene, null).alive;
        this._debugPending = false;
    }
}
