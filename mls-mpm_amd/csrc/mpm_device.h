// mpm_device.h -- device-side arithmetic shared by every MLS-MPM kernel.
//
// Each helper restates one WGSL helper of the reference (shaders.js) with the
// same operation order.  The library is compiled with -ffp-contract=off and
// correctly rounded f32 division and sqrt, so every expression rounds once per
// operator, exactly like the CPU restatement the tests compare against.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

constexpr float kFixed = 100000.0f;  // FIXED (shaders.js:15)
constexpr float kEncMax = 20000.0f;  // encode clamp (shaders.js:21-24)

// Particle state fields.  PX, PY, PZ and the reset-order id of a slot form
// one 16-byte record (x, y, z, id bits), so that the re-sort's gather of a
// particle from its old slot is one load and its store one store; the other
// fields are planes of `cap` floats (structure of arrays).
enum Plane : int {
  PX = 0, PY, PZ, VX, VY, VZ,
  C00, C01, C02, C10, C11, C12, C20, C21, C22,
  PM,
  NPLANES
};
// Floats per slot of the particle state: the 4-float record, then VX .. PM.
constexpr int kStateFloats = NPLANES + 1;

// Step constants: the Params uniform (shaders.js:27-49) plus the diffuse
// GenParams / AdvectParams (shaders.js:734-743, 872-877).
struct Params {
  uint32_t gx, gy, gz, n;
  float dt, gravity, rest_density, stiffness, eos_power, viscosity;
  float hard_min, hard_max_x, hard_max_y, hard_max_z;
  float wall_min, wall_stiffness, lookahead_k;
  float ke_threshold, spawn_rate_k, density_spray_max, density_bubble_min;
  float lifetime_foam, pressure_threshold, crest_dot_threshold, buoyancy, drag;
  uint32_t dmax;
  uint32_t unit_mass;  // 1 when every particle mass is exactly 1.0f
  uint32_t nwalls;
  uint32_t lifetime_mode;  // 0: reference foam lifetime, 1: uniform [0.5, 3] s
};

// One moving wall, 20 floats (shaders.js:123-129).
struct Wall { float v[20]; };
struct Walls { Wall w[4]; };

// Particle state of one scene.
struct PState {
  float *pl;         // kStateFloats * cap floats: cap (x, y, z, id) records, then VX .. PM
  float *density;    // particle_density, by slot
  uint32_t cap;
  // A slot's byte offset fits in 32 bits (capacity <= 2^26): a uniform
  // plane base plus a 32-bit lane offset, which the compiler can address as
  // a scalar base and one vector offset instead of a 64-bit address per plane.
  __device__ __forceinline__ float &at(int k, uint32_t p) const {
    if (k < 3)
      return *reinterpret_cast<float *>(reinterpret_cast<char *>(pl) + (p << 4) + (k << 2));
    return *reinterpret_cast<float *>(reinterpret_cast<char *>(pl + (size_t)(k + 1) * cap) +
                                      (p << 2));
  }
  // The slot's record: position and, in .w, the id's bits.
  __device__ __forceinline__ float4 &rec(uint32_t p) const {
    return *reinterpret_cast<float4 *>(reinterpret_cast<char *>(pl) + (p << 4));
  }
  // Reset-order index of the particle held in slot p.
  __device__ __forceinline__ uint32_t &idr(uint32_t p) const {
    return *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(pl) + (p << 4) + 12);
  }
  __device__ __forceinline__ float &dens(uint32_t p) const {
    return *reinterpret_cast<float *>(reinterpret_cast<char *>(density) + (p << 2));
  }
};

// Grid of one scene: fixed-point accumulators (structure of arrays) and the
// f32 result of the grid update, (vx, vy, vz, decoded mass) per cell.
struct GState {
  int32_t *m, *mx, *my, *mz;
  float4 *v;
};

// DiffuseParticle, 48 bytes (shaders.js:725-732).
struct DRec {
  float4 pos;
  float4 vel;
  uint32_t ptype;
  uint32_t alive;
  float lifetime;
  float pad;
};
static_assert(sizeof(DRec) == 48, "DiffuseParticle is 48 bytes");

// encode / decode (shaders.js:21-25).  The clamp keeps |v * FIXED| <= 2e9,
// inside the i32 range, so the conversion truncates without saturating.
__device__ __forceinline__ int32_t encode(float v) {
  float c = fminf(fmaxf(v, -kEncMax), kEncMax);
  return (int32_t)(c * kFixed);
}
__device__ __forceinline__ float decode(int32_t v) { return (float)v / kFixed; }

// quadratic_weights (shaders.js:53-59) and the base cell / fraction of
// WGSL_COMPUTE_WEIGHTS (shaders.js:226-230).
struct PW {
  int b[3];
  float f[3];
  float w[3][3];
};
__device__ __forceinline__ void quad_w(float f, float *w) {
  w[0] = (0.5f * (0.5f - f)) * (0.5f - f);
  w[1] = 0.75f - f * f;
  w[2] = (0.5f * (0.5f + f)) * (0.5f + f);
}
__device__ __forceinline__ void axis_weights(float x, int &b, float &f, float *w) {
  int ci = (int)floorf(x);
  b = ci - 1;
  f = (x - (float)ci) - 0.5f;
  quad_w(f, w);
}
__device__ __forceinline__ void particle_weights(float x, float y, float z, PW &o) {
  axis_weights(x, o.b[0], o.f[0], o.w[0]);
  axis_weights(y, o.b[1], o.f[1], o.w[1]);
  axis_weights(z, o.b[2], o.f[2], o.w[2]);
}

// pow(x, eos_power) with eos_power == 1 (the reference default): x exactly.
// Other exponents are refused when the scene is configured.
__device__ __forceinline__ float eos_pressure(const Params &P, float density) {
  float rest = P.rest_density;
  float pressure = P.stiffness * ((density / rest) - 1.0f);
  float pfloor = (density > rest * 0.7f) ? -0.1f : 0.0f;
  return fmaxf(pressure, pfloor);
}

// cell_idx_checked (shaders.js:93-98).
__device__ __forceinline__ int cell_idx_checked(int cx, int cy, int cz, const Params &P) {
  if (cx < 0 || cx >= (int)P.gx || cy < 0 || cy >= (int)P.gy || cz < 0 || cz >= (int)P.gz)
    return -1;
  return (cz * (int)P.gy + cy) * (int)P.gx + cx;
}

// xorshift32 and diffuse_rand (shaders.js:756-762).
__device__ __forceinline__ uint32_t xorshift32(uint32_t x) {
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  return x;
}
__device__ __forceinline__ float diffuse_rand(uint32_t seed) {
  return (float)xorshift32(seed) / 4294967295.0f;
}

// ---- moving walls (shaders.js:137-193) ------------------------------------
struct F3 { float x, y, z; };
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

__device__ __forceinline__ F3 wall_to_local(F3 p, const Wall &w) {
  F3 d = f3(p.x - w.v[0], p.y - w.v[1], p.z - w.v[2]);
  float cy = w.v[3], sy = w.v[7], cp = w.v[11], sp = w.v[15];
  float x1 = cy * d.x - sy * d.z;
  float z1 = sy * d.x + cy * d.z;
  return f3(x1, cp * d.y + sp * z1, (-sp) * d.y + cp * z1);
}
__device__ __forceinline__ F3 wall_to_world_dir(F3 v, const Wall &w) {
  float cy = w.v[3], sy = w.v[7], cp = w.v[11], sp = w.v[15];
  float y1 = cp * v.y - sp * v.z;
  float z1 = sp * v.y + cp * v.z;
  return f3(cy * v.x + sy * z1, y1, (-sy) * v.x + cy * z1);
}
__device__ __forceinline__ F3 wall_vel_at(F3 p, const Wall &w) {
  F3 r = f3(p.x - w.v[0], p.y - w.v[1], p.z - w.v[2]);
  float ox = w.v[16], oy = w.v[17], oz = w.v[18];
  F3 c = f3(oy * r.z - oz * r.y, oz * r.x - ox * r.z, ox * r.y - oy * r.x);
  return f3(w.v[12] + c.x, w.v[13] + c.y, w.v[14] + c.z);
}
// Returns the signed distance; the (wall-local) normal goes to n.
__device__ __forceinline__ float wall_sdf_normal(F3 lp, F3 mn, F3 mx, F3 &n) {
  F3 dmin = f3(mn.x - lp.x, mn.y - lp.y, mn.z - lp.z);
  F3 dmax = f3(lp.x - mx.x, lp.y - mx.y, lp.z - mx.z);
  F3 q = f3(fmaxf(dmin.x, dmax.x), fmaxf(dmin.y, dmax.y), fmaxf(dmin.z, dmax.z));
  float qmax = fmaxf(q.x, fmaxf(q.y, q.z));
  if (qmax > 0.0f) {
    F3 nearest = f3(fminf(fmaxf(lp.x, mn.x), mx.x), fminf(fmaxf(lp.y, mn.y), mx.y),
                    fminf(fmaxf(lp.z, mn.z), mx.z));
    F3 diff = f3(lp.x - nearest.x, lp.y - nearest.y, lp.z - nearest.z);
    float len = sqrtf(dot3(diff, diff));
    if (len > 1e-6f) n = f3(diff.x / len, diff.y / len, diff.z / len);
    else n = f3(0.0f, 1.0f, 0.0f);
    return len;
  }
  if (qmax == q.x) n = f3(dmax.x > dmin.x ? 1.0f : -1.0f, 0.0f, 0.0f);
  else if (qmax == q.y) n = f3(0.0f, dmax.y > dmin.y ? 1.0f : -1.0f, 0.0f);
  else n = f3(0.0f, 0.0f, dmax.z > dmin.z ? 1.0f : -1.0f);
  return qmax;
}
__device__ __forceinline__ F3 wall_min3(const Wall &w) { return f3(w.v[4], w.v[5], w.v[6]); }
__device__ __forceinline__ F3 wall_max3(const Wall &w) { return f3(w.v[8], w.v[9], w.v[10]); }
__device__ __forceinline__ bool inside_open(F3 l, F3 mn, F3 mx) {
  return l.x > mn.x && l.y > mn.y && l.z > mn.z && l.x < mx.x && l.y < mx.y && l.z < mx.z;
}

}  // namespace mpm
