// mpm_ops.h -- per-particle and per-cell bodies of the MLS-MPM stages,
// shared by the direct kernels (mpm_kernels.hip), the tiled kernels
// (mpm_tiled.hip) and the batched kernels (mpm_batch.hip).
#pragma once
#include "mpm_device.h"

namespace mpm {

// Shared body of P2G momentum: density, EOS pressure, stress and the APIC
// momentum contribution per stencil cell (WGSL_P2G_MOM, shaders.js:286-370).
// `mass_at(k,j,i)` returns the decoded cell mass of stencil cell (i,j,k);
// `emit(k,j,i,ex,ey,ez)` receives the three encoded contributions.
// Particle inputs of P2G momentum, loaded ahead of their use so a loop can
// prefetch the next particle while computing the current one.
struct MomIn {
  float px, py, pz, vx, vy, vz, m;
  float a[9];
};
template <bool UNIT>
__device__ __forceinline__ void load_mom_in(const PState &S, uint32_t p, MomIn &in) {
  const float4 q = S.rec(p);
  in.px = q.x;
  in.py = q.y;
  in.pz = q.z;
  in.vx = S.at(VX, p);
  in.vy = S.at(VY, p);
  in.vz = S.at(VZ, p);
  in.m = UNIT ? 1.0f : S.at(PM, p);
#pragma unroll
  for (int k = 0; k < 9; k++) in.a[k] = S.at(C00 + k, p);
}

// Returns the particle density; contributions go to emit() unless the
// density is below the reference's 1e-8 cut-off.
template <bool UNIT, class MassAt, class Emit>
__device__ __forceinline__ float p2g_mom_body(const Params &P, const MomIn &in, const PW &pw,
                                              MassAt mass_at, Emit emit) {
  const float vx = in.vx, vy = in.vy, vz = in.vz, m = in.m;
  const float a00 = in.a[0], a01 = in.a[1], a02 = in.a[2];
  const float a10 = in.a[3], a11 = in.a[4], a12 = in.a[5];
  const float a20 = in.a[6], a21 = in.a[7], a22 = in.a[8];

  float density = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++)
        density += ((pw.w[0][i] * pw.w[1][j]) * pw.w[2][k]) * mass_at(k, j, i);
  if (density <= 1e-8f) return density;

  float rest = P.rest_density;
  float vol = m / fmaxf(density, rest * 0.15f);
  float pressure = eos_pressure(P, density);
  float visc = P.viscosity;
  float s00 = -pressure + visc * (a00 + a00);
  float s11 = -pressure + visc * (a11 + a11);
  float s22 = -pressure + visc * (a22 + a22);
  float s01 = visc * (a01 + a10);
  float s02 = visc * (a02 + a20);
  float s12 = visc * (a12 + a21);
  float coef = ((-vol) * 4.0f) * P.dt;

  float d[3][3];  // d[axis][i] = node - particle offset, (-f-1, -f, 1-f)
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float f = pw.f[a];
    d[a][0] = (-f) - 1.0f;
    d[a][1] = -f;
    d[a][2] = 1.0f - f;
  }
  // Products that depend on one axis only, and the (i,j) partial sums:
  // Q_x = (a00*dx_i + a01*dy_j) + a02*dz_k keeps the reference association.
  float qx[3][3], qy[3][3], qz[3][3], sx[3][3], sy[3][3], sz[3][3];
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      float dxi = d[0][i], dyj = d[1][j];
      qx[j][i] = a00 * dxi + a01 * dyj;
      qy[j][i] = a10 * dxi + a11 * dyj;
      qz[j][i] = a20 * dxi + a21 * dyj;
      sx[j][i] = s00 * dxi + s01 * dyj;
      sy[j][i] = s01 * dxi + s11 * dyj;
      sz[j][i] = s02 * dxi + s12 * dyj;
    }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float wzk = pw.w[2][k], dzk = d[2][k];
    float qxz = a02 * dzk, qyz = a12 * dzk, qzz = a22 * dzk;
    float sxz = s02 * dzk, syz = s12 * dzk, szz = s22 * dzk;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      float wyzk = pw.w[1][j] * wzk;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        float w = pw.w[0][i] * wyzk;
        float Qx = qx[j][i] + qxz;
        float Qy = qy[j][i] + qyz;
        float Qz = qz[j][i] + qzz;
        float fxf = coef * (sx[j][i] + sxz);
        float fyf = coef * (sy[j][i] + syz);
        float fzf = coef * (sz[j][i] + szz);
        float ux = UNIT ? (vx + Qx) : m * (vx + Qx);
        float uy = UNIT ? (vy + Qy) : m * (vy + Qy);
        float uz = UNIT ? (vz + Qz) : m * (vz + Qz);
        emit(k, j, i, encode(w * (ux + fxf)), encode(w * (uy + fyf)), encode(w * (uz + fzf)));
      }
    }
  }
  return density;
}

// P2G momentum of particle p, reading its inputs and writing its density.
template <bool UNIT, class MassAt, class Emit>
__device__ __forceinline__ void p2g_mom_particle(const Params &P, const PState &S, uint32_t p,
                                                 const PW &pw, MassAt mass_at, Emit emit) {
  MomIn in;
  load_mom_in<UNIT>(S, p, in);
  S.dens(p) = p2g_mom_body<UNIT>(P, in, pw, mass_at, emit);
}

// Grid update of one cell (WGSL_UPDATE_GRID, shaders.js:443-512).  Returns
// (vx, vy, vz, decoded mass); the mass lane is what DECODE_MASS produced.
__device__ __forceinline__ float4 update_cell(const Params &P, const Walls &W, int32_t im,
                                              int32_t imx, int32_t imy, int32_t imz, int cx,
                                              int cy, int cz) {
  float cm = decode(im);
  if (cm <= 0.0f) return make_float4(0.0f, 0.0f, 0.0f, cm);
  float vx = decode(imx) / cm;
  float vy = decode(imy) / cm;
  float vz = decode(imz) / cm;
  vy += P.gravity * P.dt;
  if (cx < 2 || cx > (int)P.gx - 3) vx = 0.0f;
  if (cy < 2 && vy < 0.0f) vy = 0.0f;
  if (cy > (int)P.gy - 3 && vy > 0.0f) vy = 0.0f;
  if (cz < 2 || cz > (int)P.gz - 3) vz = 0.0f;
  if (cy < 2) {
    vx *= 0.85f;
    vz *= 0.9f;
  }
  if (P.nwalls) {
    F3 cellPos = f3((float)cx + 0.5f, (float)cy + 0.5f, (float)cz + 0.5f);
    F3 wv = f3(vx, vy, vz);
    for (uint32_t wi = 0; wi < P.nwalls; wi++) {
      const Wall &w = W.w[wi];
      F3 lp = wall_to_local(cellPos, w);
      F3 mn = wall_min3(w), mx = wall_max3(w);
      if (lp.x < mn.x - 1.0f || lp.y < mn.y - 1.0f || lp.z < mn.z - 1.0f ||
          lp.x > mx.x + 1.0f || lp.y > mx.y + 1.0f || lp.z > mx.z + 1.0f)
        continue;
      F3 wvel = wall_vel_at(cellPos, w);
      F3 nl;
      float sdf = wall_sdf_normal(lp, mn, mx, nl);
      if (sdf < 0.0f) {
        wv = wvel;
      } else if (sdf < 1.0f) {
        F3 nn = wall_to_world_dir(nl, w);
        F3 vrel = f3(wv.x - wvel.x, wv.y - wvel.y, wv.z - wvel.z);
        float vn = dot3(vrel, nn);
        if (vn < 0.0f) wv = f3(wv.x - vn * nn.x, wv.y - vn * nn.y, wv.z - vn * nn.z);
      }
    }
    vx = wv.x;
    vy = wv.y;
    vz = wv.z;
  }
  return make_float4(vx, vy, vz, cm);
}

// G2P of one particle (WGSL_G2P, shaders.js:589-704) given the 27 gathered
// cell velocities through `vel_at(k,j,i)`.  Writes pos, vel, C.
struct G2POut {
  float C[9];
  float3 pos, vel;
};

template <class VelAt>
__device__ __forceinline__ void g2p_particle(const Params &P, const Walls &W, float px, float py,
                                             float pz, const PW &pw, VelAt vel_at, G2POut &o) {
  float gvx = 0.0f, gvy = 0.0f, gvz = 0.0f;
  float B00 = 0.0f, B01 = 0.0f, B02 = 0.0f;
  float B10 = 0.0f, B11 = 0.0f, B12 = 0.0f;
  float B20 = 0.0f, B21 = 0.0f, B22 = 0.0f;
  float dxs[3], dys[3], dzs[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    dxs[i] = ((float)(pw.b[0] + i) - px) + 0.5f;
    dys[i] = ((float)(pw.b[1] + i) - py) + 0.5f;
    dzs[i] = ((float)(pw.b[2] + i) - pz) + 0.5f;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float wzk = pw.w[2][k], dzk = dzs[k];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      float wyzk = pw.w[1][j] * wzk, dyj = dys[j];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        float w = pw.w[0][i] * wyzk;
        float dxi = dxs[i];
        float3 cv = vel_at(k, j, i);
        float wvx = w * cv.x, wvy = w * cv.y, wvz = w * cv.z;
        gvx += wvx; gvy += wvy; gvz += wvz;
        B00 += wvx * dxi; B01 += wvx * dyj; B02 += wvx * dzk;
        B10 += wvy * dxi; B11 += wvy * dyj; B12 += wvy * dzk;
        B20 += wvz * dxi; B21 += wvz * dyj; B22 += wvz * dzk;
      }
    }
  }
  o.C[0] = B00 * 4.0f; o.C[1] = B01 * 4.0f; o.C[2] = B02 * 4.0f;
  o.C[3] = B10 * 4.0f; o.C[4] = B11 * 4.0f; o.C[5] = B12 * 4.0f;
  o.C[6] = B20 * 4.0f; o.C[7] = B21 * 4.0f; o.C[8] = B22 * 4.0f;

  const float dt = P.dt, lk = P.lookahead_k;
  F3 np = f3(px + dt * gvx, py + dt * gvy, pz + dt * gvz);
  F3 nv = f3(gvx, gvy, gvz);

  // Floor spring, FLOOR_SPRING_BAND = 3 cells (shaders.js:5, 648-654).
  float floorZoneY = P.hard_min + 3.0f;
  float xny = np.y + (nv.y * dt) * lk;
  if (xny < floorZoneY) {
    float distBelowFloor = floorZoneY - xny;
    float blend = fminf(fmaxf(distBelowFloor / 3.0f, 0.0f), 1.0f);
    nv.y += (P.wall_stiffness * blend) * distBelowFloor;
  }

  for (uint32_t wi = 0; wi < P.nwalls; wi++) {  // moving walls (shaders.js:657-693)
    const Wall &w = W.w[wi];
    F3 wmin = wall_min3(w), wmax = wall_max3(w);
    F3 xnw = f3(np.x + (nv.x * dt) * lk, np.y + (nv.y * dt) * lk, np.z + (nv.z * dt) * lk);
    F3 lxn = wall_to_local(xnw, w);
    if (inside_open(lxn, wmin, wmax)) {
      F3 nl;
      float sxn = wall_sdf_normal(lxn, wmin, wmax, nl);
      float dist = -sxn;
      float blend = fminf(fmaxf(dist / P.wall_min, 0.0f), 1.0f);
      F3 nn = wall_to_world_dir(nl, w);
      float s = (P.wall_stiffness * blend) * dist;
      nv = f3(nv.x + nn.x * s, nv.y + nn.y * s, nv.z + nn.z * s);
    }
    F3 lnp = wall_to_local(np, w);
    if (inside_open(lnp, wmin, wmax)) {
      F3 nl;
      float snp = wall_sdf_normal(lnp, wmin, wmax, nl);
      F3 nn = wall_to_world_dir(nl, w);
      float push = fmaxf(snp, -0.5f);
      np = f3(np.x - nn.x * push, np.y - nn.y * push, np.z - nn.z * push);
      F3 wvel = wall_vel_at(np, w);
      float vn = dot3(f3(nv.x - wvel.x, nv.y - wvel.y, nv.z - wvel.z), nn);
      if (vn < 0.0f) nv = f3(nv.x - vn * nn.x, nv.y - vn * nn.y, nv.z - vn * nn.z);
    }
  }

  // hard_clamp_axis (shaders.js:103-108, 696-700).
  if (np.x < P.hard_min) { np.x = P.hard_min; nv.x = 0.0f; }
  if (np.x > P.hard_max_x) { np.x = P.hard_max_x; nv.x = 0.0f; }
  if (np.y < P.hard_min) { np.y = P.hard_min; nv.y = 0.0f; }
  if (np.y > P.hard_max_y) { np.y = P.hard_max_y; nv.y = 0.0f; }
  if (np.z < P.hard_min) { np.z = P.hard_min; nv.z = 0.0f; }
  if (np.z > P.hard_max_z) { np.z = P.hard_max_z; nv.z = 0.0f; }
  o.pos = make_float3(np.x, np.y, np.z);
  o.vel = make_float3(nv.x, nv.y, nv.z);
}

// Spawn decision of WGSL_DIFFUSE_GENERATE (shaders.js:805-829) for one fluid
// particle: kinetic energy, trapped air from the EOS pressure, the wave
// crest from the 6-neighbour decoded-mass gradient (estimate_normal,
// shaders.js:766-787, read from the velocity grid's mass lane), the score
// and the seeded spawn chance.  The reference pops free-list slots with
// unordered atomics (try_spawn, shaders.js:791-798); the callers here mark
// the spawners in a bitmask by reset-order id instead, which a scan ranks.
struct GenEval {
  float score, speed, density;
  uint32_t rand_seed;
};

__device__ __forceinline__ bool gen_decide(const Params &P, const GState &g, float density,
                                           float vx, float vy, float vz, float px, float py,
                                           float pz, float m, uint32_t id, uint32_t seed,
                                           GenEval &e) {
  if (density <= 0.0f || density > P.density_bubble_min) return false;
  float speed2 = (vx * vx + vy * vy) + vz * vz;
  float speed = sqrtf(speed2);
  float ke = (0.5f * m) * speed2;
  float pressure = eos_pressure(P, density);
  float trappedAir = fmaxf(0.0f, pressure - P.pressure_threshold);
  int gxi = (int)floorf(px), gyi = (int)floorf(py), gzi = (int)floorf(pz);
  int X = (int)P.gx, Y = (int)P.gy, Z = (int)P.gz;
  float waveCrest = 0.0f;
  if (!(gxi < 1 || gxi >= X - 1 || gyi < 1 || gyi >= Y - 1 || gzi < 1 || gzi >= Z - 1)) {
    uint32_t c = (uint32_t)((gzi * Y + gyi) * X + gxi);
    float mL = g.v[c - 1].w, mR = g.v[c + 1].w;
    float mD = g.v[c - X].w, mU = g.v[c + X].w;
    float mB = g.v[c - X * Y].w, mF = g.v[c + X * Y].w;
    F3 n = f3(mL - mR, mD - mU, mB - mF);
    float len = sqrtf(dot3(n, n));
    if (len >= 1e-6f && speed > 1e-4f) {
      F3 nn = f3(n.x / len, n.y / len, n.z / len);
      float dotVal = ((vx / speed) * nn.x + (vy / speed) * nn.y) + (vz / speed) * nn.z;
      if (dotVal > P.crest_dot_threshold) waveCrest = dotVal * speed;
    }
  }
  float score = (ke + trappedAir * 2.0f) + waveCrest * 1.5f;
  if (score < P.ke_threshold) return false;
  uint32_t randSeed = id * 9781u + seed;
  float spawnChance = fminf(1.0f, ((score - P.ke_threshold) * P.spawn_rate_k) * P.dt);
  if (diffuse_rand(randSeed) > spawnChance) return false;
  e.score = score;
  e.speed = speed;
  e.density = density;
  e.rand_seed = randSeed;
  return true;
}

// Where a G2P pass that also decides the diffuse spawners marks them: the
// spawner bitmask by id and each spawner's (slot bits, score, speed,
// density); bits == null: no decision.
struct GenOut {
  const uint32_t *seed;  // the last-used seed; this step's generate uses seed + 1
  uint32_t *bits;
  float4 *val;
};

__device__ __forceinline__ void gen_mark(const GenOut &go, uint32_t id, uint32_t slot,
                                         const GenEval &e) {
  go.val[id] = make_float4(__uint_as_float(slot), e.score, e.speed, e.density);
  atomicOr(&go.bits[id >> 5], 1u << (id & 31));
}

// Store a G2P result, with the particle's id, into slot d of a particle
// state (the position and id as one 16-byte record).
__device__ __forceinline__ void store_g2p(const PState &S, uint32_t d, const G2POut &o,
                                          uint32_t id) {
#pragma unroll
  for (int k = 0; k < 9; k++) S.at(C00 + k, d) = o.C[k];
  S.rec(d) = make_float4(o.pos.x, o.pos.y, o.pos.z, __uint_as_float(id));
  S.at(VX, d) = o.vel.x;
  S.at(VY, d) = o.vel.y;
  S.at(VZ, d) = o.vel.z;
}

}  // namespace mpm
