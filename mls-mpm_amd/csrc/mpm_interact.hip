// mpm_interact.hip -- the interaction operations of FluidGPU: the hand push
// applied inside the step, deleteNear's compaction, raycastFluid and the
// upload of spawnParticles (fluid-gpu.js:145-180, 479-638; shaders.js:80-84,
// 394-433, 521-577).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "mpm_kernels.h"
#include "mpm_scene.h"

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
  const uint32_t id = S.idr(s);
  atomicOr(&bits[id >> 5], 1u << (id & 31));
}

// Survivors to slot (and id) = their rank among the survivors in id order.
__global__ __launch_bounds__(256) void k_del_move(PState S, PState D, uint32_t n,
                                                  const uint32_t *bits, const uint32_t *prefix) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t id = S.idr(s), w = bits[id >> 5], b = id & 31;
  if (!((w >> b) & 1u)) return;
  const uint32_t d = prefix[id >> 5] + (uint32_t)__popc(w & ((1u << b) - 1u));
  const float4 q = S.rec(s);
  D.rec(d) = make_float4(q.x, q.y, q.z, __uint_as_float(d));
#pragma unroll
  for (int k = VX; k < NPLANES; k++) D.at(k, d) = S.at(k, s);
  D.density[d] = S.density[s];
}

// Ray count of the allocation-free raycast: its scratch holds this many
// rays and hit words.
constexpr uint32_t kRaycastMaxRays = MPM_RAYCAST_NOALLOC_MAX;

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

// Many rays in one pass over the particles, for the allocation-free entry
// point: every thread tests its particle against each ray with the same
// arithmetic as k_raycast; a wavefront folds its lanes' hits of a ray with
// shuffles before one atomic minimum.  rays and hit live in the library's
// static scratch.
__global__ __launch_bounds__(256) void k_raycast_many(PState S, uint32_t n, const Ray *rays,
                                                      uint32_t nrays, uint32_t *hit) {
  __shared__ Ray lr[kRaycastMaxRays];
  for (uint32_t i = threadIdx.x; i < nrays; i += blockDim.x) lr[i] = rays[i];
  __syncthreads();
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = s < n;
  F3 p = f3(0.0f, 0.0f, 0.0f);
  if (live) p = f3(S.at(PX, s), S.at(PY, s), S.at(PZ, s));
  for (uint32_t i = 0; i < nrays; i++) {
    const Ray r = lr[i];
    uint32_t cand = 0xffffffffu;
    if (live) {
      const F3 to = f3(p.x - r.ox, p.y - r.oy, p.z - r.oz);
      const float t = dot3(to, f3(r.dx, r.dy, r.dz));
      if (t > 0.0f && ray_perp_dist(p, r) < r.radius) cand = __float_as_uint(t);
    }
    if (!__ballot(cand != 0xffffffffu)) continue;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const uint32_t o = __shfl_xor(cand, off, 64);
      cand = o < cand ? o : cand;
    }
    if (__lane_id() == 0) atomicMin(&hit[i], cand);
  }
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

// ---------------------------------------------------------------------------
// Allocation-free raycast.  Its scratch is this file's static device array:
// the rays (32 bytes each) and their hit words, one copy per GPU, held until
// the process ends.  Host code reaches a device's copy only through the
// address hipGetSymbolAddress returns on that device, looked up once per
// device at the first call there.  Every call on one GPU shares that copy:
// calls that overlap on one GPU race.
constexpr size_t kRaycastScratchBytes = (size_t)kRaycastMaxRays * (sizeof(Ray) + 4);
static __device__ __align__(256) unsigned char scratch[kRaycastScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];

static int scratch_of_device(int dev, void **out) {
  if (dev < 0 || dev >= 64) return set_error(MPM_ERR_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(raycast scratch)");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return MPM_OK;
}

int scene_raycast_noalloc(mpm_scene *s, uint32_t nrays, const float *origins3, const float *dirs3,
                          const float *radii, float *t_out, int *hit_out) {
  if (nrays > kRaycastMaxRays) {
    char msg[160];
    snprintf(msg, sizeof msg, "mpm_raycast_noalloc: %u rays, at most %u per call", nrays,
             kRaycastMaxRays);
    return set_error(MPM_ERR_INVALID, msg);
  }
  if (!nrays) return MPM_OK;
  if (!origins3 || !dirs3 || !radii || !t_out || !hit_out)
    return set_error(MPM_ERR_INVALID, "mpm_raycast_noalloc: NULL argument");
  Ray rays[kRaycastMaxRays];
  for (uint32_t i = 0; i < nrays; i++) {
    for (int a = 0; a < 3; a++)
      if (!std::isfinite(origins3[3 * i + a]) || !std::isfinite(dirs3[3 * i + a]))
        return set_error(MPM_ERR_INVALID, "mpm_raycast_noalloc: origin and dir must be finite");
    if (!(radii[i] >= 0.0f) || !std::isfinite(radii[i]))
      return set_error(MPM_ERR_INVALID, "mpm_raycast_noalloc: radius must be finite and >= 0");
    Ray &r = rays[i];
    r.ox = origins3[3 * i];
    r.oy = origins3[3 * i + 1];
    r.oz = origins3[3 * i + 2];
    r.dx = dirs3[3 * i];
    r.dy = dirs3[3 * i + 1];
    r.dz = dirs3[3 * i + 2];
    r.radius = radii[i];
    r.pad = 0.0f;
  }
  hipError_t e = hipSetDevice(s->cfg.device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  void *base = nullptr;
  int rc = scratch_of_device(s->cfg.device, &base);
  if (rc) return rc;
  Ray *drays = (Ray *)base;
  uint32_t *dhit = (uint32_t *)((char *)base + (size_t)kRaycastMaxRays * sizeof(Ray));
  PState S;
  S.pl = s->planes[s->cur];
  S.density = s->dens[s->cur];
  S.cap = s->cfg.capacity ? s->cfg.capacity : 1;
  const uint32_t n = s->n;
  uint32_t bits[kRaycastMaxRays];
  e = hipMemcpyAsync(drays, rays, (size_t)nrays * sizeof(Ray), hipMemcpyHostToDevice, s->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(rays)");
  // the previous call's hits stay in the scratch: clear every word used
  k_fill_u32<<<nblocks(nrays, 256), 256, 0, s->stream>>>(dhit, nrays, 0xffffffffu);
  if (n) k_raycast_many<<<nblocks(n, 256), 256, 0, s->stream>>>(S, n, drays, nrays, dhit);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "k_raycast_many");
  e = hipMemcpyAsync(bits, dhit, (size_t)nrays * 4, hipMemcpyDeviceToHost, s->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(hits)");
  e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  for (uint32_t i = 0; i < nrays; i++) {
    const int h = bits[i] != 0xffffffffu;
    hit_out[i] = h;
    float f = 0.0f;
    if (h) std::memcpy(&f, &bits[i], 4);
    t_out[i] = f;
  }
  return MPM_OK;
}

}  // namespace mpm
