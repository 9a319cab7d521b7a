// Drives the interaction methods of FluidHIP on the GPU -- updateHand,
// deleteNear, raycastFluid, spawnParticles -- between frames of a small
// seeded scene and dumps every input and the state after every frame (raw
// little-endian float32) into the directory given as argv[2], for comparison
// with the CPU restatement replaying the same calls.
import fs from 'fs';
import path from 'path';
import { FluidHIP } from '../fluid-hip.mjs';

const dir = process.argv[2];
let s = 777;
const rnd = () => {  // seeded jitter and spray source (LCG, 2^32)
    s = (Math.imul(s, 1664525) + 1013904223) >>> 0;
    return s / 4294967296;
};
const f = new FluidHIP(2, 2, 3, 0.025, 40000);
f.random = rnd;
const n0 = f.fillBlock(0.1, 0.1, 0.1, 0.6, 0.8, 0.5);
const put = (name, a) => fs.writeFileSync(path.join(dir, name), Buffer.from(a.buffer, a.byteOffset, a.byteLength));
put('pos3.f32', f.particle_pos.subarray(0, 3 * n0));
f.initGPU(0);
const eye = [10.0, 30.0, 60.0];
const log = [];
const frame = (k) => {
    if (!f.simFrame()) throw f.lastError;
    put(`pos4_${k}.f32`, f.readPositions());
    put(`vel4_${k}.f32`, f.readVelocities());
    log.push({ frame: k, n: f.active_particle_num });
};
f.updateHand([6.0, 6.0, 6.0], [1.5, 0.0, -0.5], 2.5, 0.8, true, eye);
frame(0);
frame(1);
const dn = [0.2, 0.1, 1.0], dl = Math.hypot(...dn);
const dirD = dn.map((x) => x / dl);
f.deleteNear(null, [7.0, 8.0, -5.0], dirD, 2.0);
f.pollDelete();
log.push({ deleted: f.active_particle_num });
f.updateHand([8.0, 10.0, 7.0], [0.0, 1.0, 1.0], 3.0, 0.5, true, eye);
frame(2);
f.raycastFluid(null, [6.0, 30.0, 6.0], [0.0, -1.0, 0.0], 1.0);
f.pollRaycast();
const hit = f.lastRayHit;
const before = f.active_particle_num;
const added = f.spawnParticles(hit ? hit[0] : 6.0, (hit ? hit[1] : 10.0) - 2.0, hit ? hit[2] : 6.0,
                               500, 1.5, [0.0, -1.0, 0.0], 2.0);
// the spawned particles, as appended (slots before .. before + added)
const p4 = f.readPositions(), v4 = f.readVelocities();
put('spawn_pos4.f32', p4.subarray(4 * before, 4 * (before + added)));
put('spawn_vel4.f32', v4.subarray(4 * before, 4 * (before + added)));
f.updateHand([0, 0, 0], [0, 0, 0], 1.0, 1.0, false, eye);
frame(3);
frame(4);
f.sync();
console.log(JSON.stringify({ n0, added, hit, eye, dirD, log,
                             grid: [f.grid_X_num, f.grid_Y_num, f.grid_Z_num] }));
f.destroy();
