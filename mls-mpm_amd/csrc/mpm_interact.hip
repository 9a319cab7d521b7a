// mpm_interact.hip -- the interaction operations of FluidGPU: the hand push
// applied inside the step, deleteNear's compaction, raycastFluid and the
// upload of spawnParticles (fluid-gpu.js:145-180, 479-638; shaders.js:80-84,
// 394-433, 521-577).
#include "mpm_kernels.h"

namespace mpm {

// ray_perp_dist (shaders.js:80-84).
__device__ __forceinline__ float ray_perp_dist(F3 p, const Ray &r) {
  const F3 to = f3(p.x - r.ox, p.y - r.oy, p.z - r.oz);
  const F3 d = f3(r.dx, r.dy, r.dz);
  const float t = dot3(to, d);
  const F3 perp = f3(to.x - t * d.x, to.y - t * d.y, to.z - t * d.z);
  return sqrtf(dot3(perp, perp));
}

// WGSL_APPLY_HAND (shaders.js:410-432), grid-stride over the box the host
// computed.  The parameters are read from device memory so that a captured
// step replays the latest push.  With act set (brick-sparse grid update), the
// bricks of the cells it changes are marked active, so that the next step's
// resolve rewrites them as the reference's dense update would.
__global__ __launch_bounds__(256) void k_apply_hand(Params P, GState g, uint8_t *act,
                                                    uint32_t nbx, uint32_t nby,
                                                    const HandDev *hp) {
  const HandDev h = *hp;
  if (!h.on) return;
  const int ex = h.bmax[0] - h.bmin[0] + 1, ey = h.bmax[1] - h.bmin[1] + 1,
            ez = h.bmax[2] - h.bmin[2] + 1;
  if (ex <= 0 || ey <= 0 || ez <= 0) return;
  const uint64_t total = (uint64_t)ex * (uint64_t)ey * (uint64_t)ez;
  const F3 e0 = f3(h.px - h.ex, h.py - h.ey, h.pz - h.ez);
  const float len = sqrtf(dot3(e0, e0));
  const F3 rd = f3(e0.x / len, e0.y / len, e0.z / len);
  const int X = (int)P.gx, Y = (int)P.gy, Z = (int)P.gz;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const int cx = h.bmin[0] + (int)(i % (uint64_t)ex);
    const int cy = h.bmin[1] + (int)((i / (uint64_t)ex) % (uint64_t)ey);
    const int cz = h.bmin[2] + (int)(i / ((uint64_t)ex * (uint64_t)ey));
    if (cx < 0 || cy < 0 || cz < 0 || cx >= X || cy >= Y || cz >= Z) continue;
    const F3 to = f3(((float)cx + 0.5f) - h.px, ((float)cy + 0.5f) - h.py,
                     ((float)cz + 0.5f) - h.pz);
    const float t0 = dot3(to, rd);
    const F3 perp = f3(to.x - t0 * rd.x, to.y - t0 * rd.y, to.z - t0 * rd.z);
    const float d = sqrtf(dot3(perp, perp));
    if (d >= h.radius) continue;
    const float t = 1.0f - d / h.radius;
    const float w = (t * t) * h.strength;
    const size_t idx = ((size_t)cz * Y + cy) * X + cx;
    float4 v = g.v[idx];
    v.x += h.vx * w;
    v.y += h.vy * w;
    v.z += h.vz * w;
    g.v[idx] = v;
    if (act) act[((uint32_t)(cz >> 3) * nby + (uint32_t)(cy >> 3)) * nbx + (uint32_t)(cx >> 3)] |= 1u;
  }
}

// WGSL_COMPACT's test (shaders.js:538-540): keep bit by reset-order id.
__global__ __launch_bounds__(256) void k_del_mark(PState S, uint32_t n, Ray r, uint32_t *bits) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  if (ray_perp_dist(f3(S.at(PX, s), S.at(PY, s), S.at(PZ, s)), r) < r.radius) return;
  const uint32_t id = S.id[s];
  atomicOr(&bits[id >> 5], 1u << (id & 31));
}

// Survivors to slot (and id) = their rank among the survivors in id order.
__global__ __launch_bounds__(256) void k_del_move(PState S, PState D, uint32_t n,
                                                  const uint32_t *bits, const uint32_t *prefix) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t id = S.id[s], w = bits[id >> 5], b = id & 31;
  if (!((w >> b) & 1u)) return;
  const uint32_t d = prefix[id >> 5] + (uint32_t)__popc(w & ((1u << b) - 1u));
#pragma unroll
  for (int k = 0; k < NPLANES; k++) D.at(k, d) = S.at(k, s);
  D.density[d] = S.density[s];
  D.id[d] = d;
}

// WGSL_RAYCAST (shaders.js:562-576): atomic minimum of t's bits (positive
// floats order like their bit patterns).
__global__ __launch_bounds__(256) void k_raycast(PState S, uint32_t n, Ray r, uint32_t *hit) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const F3 p = f3(S.at(PX, s), S.at(PY, s), S.at(PZ, s));
  const F3 to = f3(p.x - r.ox, p.y - r.oy, p.z - r.oz);
  const float t = dot3(to, f3(r.dx, r.dy, r.dz));
  if (t <= 0.0f) return;
  if (ray_perp_dist(p, r) >= r.radius) return;
  atomicMin(hit, __float_as_uint(t));
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

static inline uint32_t nblocks(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

void launch_apply_hand(const Params &P, const GState &g, uint8_t *act, uint32_t nbx,
                       uint32_t nby, const HandDev *h, hipStream_t st) {
  k_apply_hand<<<1024, 256, 0, st>>>(P, g, act, nbx, nby, h);
}

void launch_delete_mark_move(const PState &S, const PState &D, uint32_t n, const Ray &r,
                             uint32_t *bits, uint32_t *prefix, uint32_t *scratch,
                             uint32_t *total, uint32_t nwords, hipStream_t st) {
  k_fill_u32<<<nblocks(nwords, 256), 256, 0, st>>>(bits, nwords, 0u);
  if (n) k_del_mark<<<nblocks(n, 256), 256, 0, st>>>(S, n, r, bits);
  launch_scan(bits, nwords, true, prefix, scratch, total, st);
  if (n) k_del_move<<<nblocks(n, 256), 256, 0, st>>>(S, D, n, bits, prefix);
}

void launch_raycast(const PState &S, uint32_t n, const Ray &r, uint32_t *hit, hipStream_t st) {
  k_fill_u32<<<1, 1, 0, st>>>(hit, 1, 0xffffffffu);
  if (n) k_raycast<<<nblocks(n, 256), 256, 0, st>>>(S, n, r, hit);
}

}  // namespace mpm
