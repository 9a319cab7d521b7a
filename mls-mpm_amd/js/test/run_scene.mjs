// Runs a small scene through FluidHIP + DiffuseHIP on the GPU and dumps its
// inputs, the wall uniform of every frame and the state after every frame
// (raw little-endian float32 / bytes) into the directory given as argv[2],
// for comparison with the CPU restatement.
import fs from 'fs';
import path from 'path';
import { FluidHIP, FRAME } from '../fluid-hip.mjs';
import { DiffuseHIP } from '../diffuse-hip.mjs';

const dir = process.argv[2];
const frames = parseInt(process.argv[3] || '6', 10);
const substeps = parseInt(process.argv[4] || '1', 10);
let s = 12345;
const rnd = () => {  // seeded jitter source (LCG, 2^32)
    s = (Math.imul(s, 1664525) + 1013904223) >>> 0;
    return s / 4294967296;
};
const f = new FluidHIP(2, 2, 3, 0.025, 30000);
f.random = rnd;
const n = f.fillBlock(0.1, 0.1, 0.1, 0.6, 0.8, 0.5);
const put = (name, a) => fs.writeFileSync(path.join(dir, name), Buffer.from(a.buffer, a.byteOffset, a.byteLength));
put('pos3.f32', f.particle_pos.subarray(0, 3 * n));
f.SUBSTEPS = substeps;
f.initGPU(0);
const d = new DiffuseHIP(8192);
d.initGPU(0, f);
f.diffuse = d;
f.setWall(0, [13.0, 2.0, 4.0], [15.0, 11.0, 20.0]);
const walls = [];
for (let k = 0; k < frames; k++) {
    if (k >= 1) f.setWall(0, [13.0 - 0.3 * k, 2.0, 4.0], [15.0 - 0.3 * k, 11.0, 20.0]);
    if (k >= 2) f.rotateWall(0, 0.05, 0.0);
    if (!f.simFrame()) throw f.lastError;
    walls.push(Float32Array.from(f._wallF32));
    put(`pos4_${k}.f32`, f.readPositions());
    put(`vel4_${k}.f32`, f.readVelocities());
}
const wall = new Float32Array(walls.length * walls[0].length);
walls.forEach((w, i) => wall.set(w, i * w.length));
put('walls.f32', wall);
const pool = new Uint8Array(d.readParticles());
put('pool.bin', pool);
// the same state through the pinned frame handoff
f.beginFrame(FRAME.ALL | FRAME.RESET_ORDER);
const out = { pos4: new Float32Array(4 * n), vel4: new Float32Array(4 * n),
              density: new Float32Array(n), diffuse: new ArrayBuffer(8192 * 48) };
const fr = f.pollFrame(out, true);
put('frame_pos4.f32', out.pos4);
put('frame_vel4.f32', out.vel4);
put('frame_pool.bin', new Uint8Array(out.diffuse));
f.sync();
console.log(JSON.stringify({ n, frames, alive: d.aliveCount, grid: [f.grid_X_num, f.grid_Y_num, f.grid_Z_num],
                             substeps, frame: { n: fr.n, top: fr.top, steps: fr.steps } }));
f.destroy();
