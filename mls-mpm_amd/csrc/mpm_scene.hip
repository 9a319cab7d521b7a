// mpm_scene.hip -- scene lifetime, state upload/download and the per-step
// launch sequence of FluidGPU.simFrame + DiffuseGPU.step
// (fluid-gpu.js:648-684, diffuse-gpu.js:148-166).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <utility>
#include <vector>

#include "mpm_scene.h"

namespace mpm {

#define HIPCHK(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

// Formats into a buffer of the calling thread's stack: the server loop and
// its client threads report errors concurrently.
static int failf(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int failf(int code, const char *fmt, ...) {
  char msg[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof msg, fmt, ap);
  va_end(ap);
  return set_error(code, msg);
}

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

int validate_config(const mpm_config &c) {
  if (c.grid_x < 5 || c.grid_y < 5 || c.grid_z < 5)
    return failf(MPM_ERR_INVALID, "grid must be at least 5 cells per axis (got %ux%ux%u)",
                 c.grid_x, c.grid_y, c.grid_z);
  uint64_t G = (uint64_t)c.grid_x * c.grid_y * c.grid_z;
  if (G > (1ull << 30)) return failf(MPM_ERR_INVALID, "grid has %llu cells, limit 2^30",
                                     (unsigned long long)G);
  if (c.capacity > (1u << 26))
    return failf(MPM_ERR_INVALID, "capacity %u exceeds 2^26", c.capacity);
  if (c.diffuse_capacity > (1u << 30))
    return failf(MPM_ERR_INVALID, "diffuse_capacity %u exceeds 2^30", c.diffuse_capacity);
  if (c.tile_ppt && !tiled_ppt_supported(c.tile_ppt))
    return failf(MPM_ERR_INVALID, "tile_ppt %u: supported 4, 8, 16", c.tile_ppt);
  if (c.tile_margin && !tiled_margin_supported(c.tile_margin))
    return failf(MPM_ERR_INVALID, "tile_margin %u: supported 1, 2", c.tile_margin);
  if ((c.tile_threads && !tiled_threads_supported(c.tile_threads)) ||
      (c.g2p_threads && !tiled_threads_supported(c.g2p_threads)))
    return failf(MPM_ERR_INVALID, "tile_threads / g2p_threads: supported 256");
  if (c.lifetime_mode > 1)
    return failf(MPM_ERR_INVALID, "lifetime_mode %u: 0 (reference) or 1 (uniform)",
                 c.lifetime_mode);
  if (c.eos_power != 1.0f)
    return failf(MPM_ERR_UNSUPPORTED, "eos_power %g: only 1 (the reference default) is built",
                 (double)c.eos_power);
  const float vals[] = {c.dt, c.gravity, c.rest_density, c.stiffness, c.viscosity, c.hard_min,
                        c.hard_max_x, c.hard_max_y, c.hard_max_z, c.wall_min, c.wall_stiffness,
                        c.lookahead_k, c.ke_threshold, c.pressure_threshold,
                        c.crest_dot_threshold, c.spawn_rate_k, c.lifetime_foam, c.buoyancy,
                        c.drag, c.density_spray_max, c.density_bubble_min};
  for (float v : vals)
    if (!std::isfinite(v)) return failf(MPM_ERR_INVALID, "physics constants must be finite");
  if (!(c.rest_density > 0.0f)) return failf(MPM_ERR_INVALID, "rest_density must be > 0");
  // Positions are clamped to [hard_min, hard_max] every step; the stencil of
  // a particle at x touches floor(x)-1 .. floor(x)+1, so the box must lie in
  // [1, grid-1).
  if (!(c.hard_min >= 1.0f) || !(c.hard_max_x < (float)c.grid_x - 1.0f) ||
      !(c.hard_max_y < (float)c.grid_y - 1.0f) || !(c.hard_max_z < (float)c.grid_z - 1.0f) ||
      !(c.hard_min <= c.hard_max_x && c.hard_min <= c.hard_max_y && c.hard_min <= c.hard_max_z))
    return failf(MPM_ERR_INVALID,
                 "hard bounds must satisfy 1 <= hard_min <= hard_max_* < grid_* - 1");
  return MPM_OK;
}

// Device bytes a scene of this configuration holds (used for admission).
size_t scene_footprint(const mpm_config &c) {
  const size_t cap = c.capacity ? c.capacity : 1, D = c.diffuse_capacity;
  const size_t G = (size_t)c.grid_x * c.grid_y * c.grid_z;
  const size_t nb = (size_t)cdiv(c.grid_x, 8) * cdiv(c.grid_y, 8) * cdiv(c.grid_z, 8);
  const size_t nit = nb + cap / 1024 + 1;
  size_t b = 0;
  b += 2 * cap * (kStateFloats * 4 + 4);  // particle records and planes, density (two orders)
  b += G * (8 * 4 + 16);                 // i32 grids, overflow grids, float4 velocity/mass
  b += cap * 12 + nb * 24;               // keys, ranks, src_of; counts, offsets, first, count, spans
  b += (nb / 64 + 1) * 18 * 4 + 2304;    // re-sort scan: group bases, running totals
  b += nit * (16 + 4 * 4000);            // work items and ghost cores
  b += nb * 8 + 65536;                   // flags, lists, scan scratch
  b += cap * 72;                         // read-back staging
  if (D) b += D * (48 + 4 + 4) + cap * 16 + (cap / 32 + 2 * (D / 32) + 3) * 8 + 4096;
  return b;
}

static int dalloc(mpm_scene *s, void **p, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (s->nowned >= (int)(sizeof s->owned / sizeof s->owned[0]))
    return failf(MPM_ERR_INVALID, "internal: too many allocations");
  hipError_t e = s->arena ? (hipError_t)s->arena->alloc(bytes, p) : hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return failf(MPM_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  s->owned[s->nowned++] = *p;
  s->bytes += bytes;
  return MPM_OK;
}

#define DALLOC(ptr, bytes)                                          \
  do {                                                              \
    int rc_ = dalloc(s, (void **)&(ptr), (size_t)(bytes));          \
    if (rc_) return rc_;                                            \
  } while (0)

Params scene_params(const mpm_scene *s) {
  const mpm_config &c = s->cfg;
  Params P{};
  P.gx = c.grid_x;
  P.gy = c.grid_y;
  P.gz = c.grid_z;
  P.n = s->n;
  P.dt = c.dt;
  P.gravity = c.gravity;
  P.rest_density = c.rest_density;
  P.stiffness = c.stiffness;
  P.eos_power = c.eos_power;
  P.viscosity = c.viscosity;
  P.hard_min = c.hard_min;
  P.hard_max_x = c.hard_max_x;
  P.hard_max_y = c.hard_max_y;
  P.hard_max_z = c.hard_max_z;
  P.wall_min = c.wall_min;
  P.wall_stiffness = c.wall_stiffness;
  P.lookahead_k = c.lookahead_k;
  P.ke_threshold = c.ke_threshold;
  P.spawn_rate_k = c.spawn_rate_k;
  P.density_spray_max = c.density_spray_max;
  P.density_bubble_min = c.density_bubble_min;
  P.lifetime_foam = c.lifetime_foam;
  P.pressure_threshold = c.pressure_threshold;
  P.crest_dot_threshold = c.crest_dot_threshold;
  P.buoyancy = c.buoyancy;
  P.drag = c.drag;
  P.dmax = c.diffuse_capacity;
  P.unit_mass = (uint32_t)s->unit_mass;
  P.nwalls = s->nwalls;
  P.lifetime_mode = c.lifetime_mode;
  return P;
}

static PState pstate(const mpm_scene *s, int which) {
  PState S;
  S.pl = s->planes[which];
  S.density = s->dens[which];
  S.cap = s->cfg.capacity ? s->cfg.capacity : 1;
  return S;
}

// Brick work with the current work-item set c and the next set c ^ 1.
BrickWork bw_of(const mpm_scene *s, int c) {
  BrickWork b = s->bw;
  b.items = s->items_buf[c];
  b.items_w = s->items_buf[c ^ 1];
  b.nitems = s->nitems_buf[c];
  b.nitems_w = s->nitems_buf[c ^ 1];
  b.first = s->first_buf[c];
  b.first_w = s->first_buf[c ^ 1];
  b.count = s->count_buf[c];
  b.count_w = s->count_buf[c ^ 1];
  b.occ = s->occ_buf[c];
  b.occ_w = s->occ_buf[c ^ 1];
  return b;
}

// Descriptor of the scene for the batched kernels; phase 1 = the buffer
// parity after one more re-sorting step.
void scene_dev(const mpm_scene *s, int phase, SceneDev *d) {
  memset((void *)d, 0, sizeof *d);
  d->P = scene_params(s);
  const int c = s->cur ^ phase;
  d->S = pstate(s, c);
  d->D = pstate(s, c ^ 1);
  d->g = s->g;
  d->bw = bw_of(s, c);
  d->ko = key_out(s->sort, d->bw);
  d->ko.cols = s->tc.ppt;
  d->offsets = s->sort.offsets;
  d->src_of = s->sort.src_of;
  d->spans = s->sort.spans;
  d->so = src_out(s->sort, s->tc, d->D.density);
  d->ch = 256u * s->tc.ppt;
  d->keep = (s->cfg.flags & MPM_FLAG_KEEP_GRID) ? 1 : 0;
  d->W = s->walls;
}

static int reset_diffuse(mpm_scene *s) {
  uint32_t D = s->cfg.diffuse_capacity;
  if (!D) return MPM_OK;
  std::vector<uint32_t> fl(D);
  for (uint32_t i = 0; i < D; i++) fl[i] = i;
  HIPCHK(hipMemsetAsync(s->d.pool, 0, (size_t)D * sizeof(DRec), s->stream));
  HIPCHK(hipMemsetAsync(s->d.alivebits, 0, (size_t)((D + 31) / 32) * 4, s->stream));
  HIPCHK(hipMemcpyAsync(s->d.free_list, fl.data(), (size_t)D * 4, hipMemcpyHostToDevice,
                        s->stream));
  int32_t top = (int32_t)D;
  HIPCHK(hipMemcpyAsync(s->d.top, &top, 4, hipMemcpyHostToDevice, s->stream));
  uint32_t seed = s->cfg.seed;
  HIPCHK(hipMemcpyAsync(s->d.seed, &seed, 4, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

// Rank of every 8x8x8 brick when bricks are visited in Morton (Z) order, so
// that consecutive sort keys stay spatially close across brick boundaries.
static uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}
static std::vector<uint32_t> morton_tile_rank(uint32_t tx, uint32_t ty, uint32_t tz) {
  std::vector<std::pair<uint64_t, uint32_t>> v;
  v.reserve((size_t)tx * ty * tz);
  for (uint32_t z = 0; z < tz; z++)
    for (uint32_t y = 0; y < ty; y++)
      for (uint32_t x = 0; x < tx; x++)
        v.push_back({spread3(x) | (spread3(y) << 1) | (spread3(z) << 2), (z * ty + y) * tx + x});
  std::sort(v.begin(), v.end());
  std::vector<uint32_t> rank(v.size());
  for (size_t i = 0; i < v.size(); i++) rank[v[i].second] = (uint32_t)i;
  return rank;
}

// Zero every grid-sized accumulator and the per-brick flags (create, reset).
static int zero_grids(mpm_scene *s) {
  const size_t G = s->G;
  int32_t *i32[] = {s->g.m, s->g.mx, s->g.my, s->g.mz, s->bw.om, s->bw.ox, s->bw.oy, s->bw.oz};
  for (int32_t *p : i32)
    if (p) HIPCHK(hipMemsetAsync(p, 0, G * 4, s->stream));
  HIPCHK(hipMemsetAsync(s->g.v, 0, G * 16, s->stream));
  uint8_t *u8[] = {s->bw.ofl, s->bw.act, s->occ_buf[0], s->occ_buf[1]};
  for (uint8_t *p : u8)
    if (p) HIPCHK(hipMemsetAsync(p, 0, s->bw.nbricks, s->stream));
  return MPM_OK;
}

int scene_init(mpm_scene *s, const mpm_config &c, Arena *arena) {
  s->cfg = c;
  s->arena = arena;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return failf(MPM_ERR_HIP, "no HIP device available (%s)",
                 e == hipSuccess ? "0 devices" : hipGetErrorString(e));
  if (c.device < 0 || c.device >= ndev)
    return failf(MPM_ERR_INVALID, "device %d out of range (%d devices)", c.device, ndev);
  HIPCHK(hipSetDevice(c.device));
  HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  const size_t cap = c.capacity ? c.capacity : 1;
  s->G = c.grid_x * c.grid_y * c.grid_z;
  const size_t G = s->G;
  for (int b = 0; b < 2; b++) {
    DALLOC(s->planes[b], cap * kStateFloats * 4);
  }
  DALLOC(s->dens[0], cap * 4);
  DALLOC(s->dens[1], cap * 4);
  // brick-sorted layout, LDS-staged P2G
  s->tc.ppt = c.tile_ppt ? c.tile_ppt : 8;
  s->tc.margin = c.tile_margin ? c.tile_margin : 2;
  s->tc.threads = c.tile_threads ? c.tile_threads : 256;
  s->tc.g2p_threads = c.g2p_threads ? c.g2p_threads : 256;
  {
    BrickWork &bw = s->bw;
    bw.nbx = cdiv(c.grid_x, 8);
    bw.nby = cdiv(c.grid_y, 8);
    bw.nbz = cdiv(c.grid_z, 8);
    bw.nbricks = bw.nbx * bw.nby * bw.nbz;
    s->sort.nkeys = bw.nbricks;
    const uint32_t nit = max_items((uint32_t)cap, bw.nbricks, 4);
    for (int k = 0; k < 2; k++) {
      DALLOC(s->items_buf[k], (size_t)nit * 16);
      DALLOC(s->nitems_buf[k], 4);
      DALLOC(s->first_buf[k], (size_t)bw.nbricks * 4);
      DALLOC(s->count_buf[k], (size_t)bw.nbricks * 4);
      DALLOC(s->occ_buf[k], bw.nbricks);
      HIPCHK(hipMemsetAsync(s->nitems_buf[k], 0, 4, s->stream));
      HIPCHK(hipMemsetAsync(s->count_buf[k], 0, (size_t)bw.nbricks * 4, s->stream));
      HIPCHK(hipMemsetAsync(s->occ_buf[k], 0, bw.nbricks, s->stream));
    }
    DALLOC(s->brick_of_rank, (size_t)bw.nbricks * 4);
    DALLOC(s->sort.brick_rank, (size_t)bw.nbricks * 4);
    bw.brick_of_rank = s->brick_of_rank;
    bw.brick_rank = s->sort.brick_rank;
    const size_t core = 1000;  // 10^3 ghost cells per item
    DALLOC(bw.gm, (size_t)nit * core * 4);
    DALLOC(bw.gx, (size_t)nit * core * 4);
    DALLOC(bw.gy, (size_t)nit * core * 4);
    DALLOC(bw.gz, (size_t)nit * core * 4);
    DALLOC(bw.om, G * 4);
    DALLOC(bw.ox, G * 4);
    DALLOC(bw.oy, G * 4);
    DALLOC(bw.oz, G * 4);
    DALLOC(bw.ofl, bw.nbricks);
    DALLOC(bw.act, bw.nbricks);
    DALLOC(bw.list, (size_t)bw.nbricks * 4);
    DALLOC(bw.nlist, 4);
    HIPCHK(hipMemsetAsync(bw.nlist, 0, 4, s->stream));
    DALLOC(s->sort.keys, cap * 4);
    DALLOC(s->sort.rank, cap * 4);
    DALLOC(s->sort.counts, (size_t)s->sort.nkeys * 4);
    DALLOC(s->sort.offsets, (size_t)s->sort.nkeys * 4);
    DALLOC(s->sort.scratch, (size_t)scan_scratch_words(s->sort.nkeys) * 4);
    DALLOC(s->sort.src_of, cap * 4);
    const size_t ngroups = (s->sort.nkeys + 63) / 64;
    DALLOC(s->sort.gsum, ngroups * kScanStride * 4);
    DALLOC(s->sort.gacc, kScanStride * kAccStride * 4);
    DALLOC(s->sort.spans, (size_t)s->sort.nkeys * 8);
    HIPCHK(hipMemsetAsync(s->sort.counts, 0, (size_t)s->sort.nkeys * 4, s->stream));
    std::vector<uint32_t> rank = morton_tile_rank(bw.nbx, bw.nby, bw.nbz);
    std::vector<uint32_t> inv(rank.size());
    for (size_t i = 0; i < rank.size(); i++) inv[rank[i]] = (uint32_t)i;
    HIPCHK(hipMemcpy(s->sort.brick_rank, rank.data(), rank.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->brick_of_rank, inv.data(), inv.size() * 4, hipMemcpyHostToDevice));
  }
  DALLOC(s->g.m, G * 4);
  DALLOC(s->g.mx, G * 4);
  DALLOC(s->g.my, G * 4);
  DALLOC(s->g.mz, G * 4);
  DALLOC(s->g.v, G * 16);
  {
    int rz = zero_grids(s);
    if (rz) return rz;
  }
  const uint32_t D = c.diffuse_capacity;
  if (D) {
    uint32_t nwg = cdiv(cap, 32), nwk = cdiv(D, 32);
    uint32_t nw = nwg > nwk ? nwg : nwk;
    DALLOC(s->d.pool, (size_t)D * sizeof(DRec));
    DALLOC(s->d.free_list, (size_t)D * 4);
    DALLOC(s->d.top, 4);
    DALLOC(s->d.seed, 4);
    DALLOC(s->d.genbits, (size_t)nwg * 4);
    DALLOC(s->d.genval, cap * 16);
    DALLOC(s->d.killbits, (size_t)nwk * 4);
    DALLOC(s->d.alivebits, (size_t)nwk * 4);
    DALLOC(s->d.list, (size_t)D * 4);
    DALLOC(s->d.scratch, (size_t)scan_scratch_words(nw) * 4);
    DALLOC(s->d.total, 8);  // spawners; the stack top after the spawns
    HIPCHK(hipMemsetAsync(s->d.genbits, 0, (size_t)nwg * 4, s->stream));
    HIPCHK(hipMemsetAsync(s->d.killbits, 0, (size_t)nwk * 4, s->stream));
  }
  int rc = reset_diffuse(s);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

void scene_free(mpm_scene *s) {
  if (!s) return;
  scene_drop_graph(s);
  if (s->stream) {
    (void)hipSetDevice(s->cfg.device);
    (void)hipStreamSynchronize(s->stream);
  }
  if (s->copy_stream) (void)hipStreamSynchronize(s->copy_stream);
  for (mpm_scene::Frame &f : s->frames) {
    if (f.host) (void)hipHostFree(f.host);
    if (f.ready) (void)hipEventDestroy(f.ready);
    if (f.done) (void)hipEventDestroy(f.done);
    f = mpm_scene::Frame{};
  }
  if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
  s->copy_stream = nullptr;
  for (int i = 0; i < s->nowned; i++) {
    if (s->arena) s->arena->release(s->owned[i]);
    else (void)hipFree(s->owned[i]);
  }
  s->nowned = 0;
  if (s->stage) {
    if (s->arena) s->arena->release(s->stage);
    else (void)hipFree(s->stage);
  }
  s->stage = nullptr;
  for (int i = 0; i < mpm_scene::kEvPerStep * mpm_scene::kEvSteps; i++)
    if (s->ev[i]) (void)hipEventDestroy(s->ev[i]);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  s->stream = nullptr;
}

static int ensure_stage(mpm_scene *s, size_t bytes) {
  if (s->stage_bytes >= bytes) return MPM_OK;
  if (s->stage) {
    if (s->arena) s->arena->release(s->stage);
    else HIPCHK(hipFree(s->stage));
  }
  s->stage = nullptr;
  s->stage_bytes = 0;
  hipError_t e =
      s->arena ? (hipError_t)s->arena->alloc(bytes, &s->stage) : hipMalloc(&s->stage, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return failf(MPM_ERR_NOMEM, "staging hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  s->stage_bytes = bytes;
  return MPM_OK;
}

int scene_reset(mpm_scene *s, uint32_t n, const float *pos3, const float *vel3, const float *C9,
                const float *mass, bool device_ptrs) {
  const mpm_config &c = s->cfg;
  if (n > c.capacity)
    return failf(MPM_ERR_CAPACITY, "reset of %u particles exceeds capacity %u", n, c.capacity);
  if (n && !pos3) return failf(MPM_ERR_INVALID, "reset: pos3 is NULL");
  int unit = 1;
  if (!device_ptrs) {
    const float lim[3] = {(float)c.grid_x - 1.0f, (float)c.grid_y - 1.0f, (float)c.grid_z - 1.0f};
    for (uint64_t i = 0; i < (uint64_t)n * 3; i++) {
      float x = pos3[i];
      if (!(x >= 1.0f && x < lim[i % 3]))
        return failf(MPM_ERR_RANGE,
                     "particle %llu: coordinate %u = %g outside [1, %g) or not finite",
                     (unsigned long long)(i / 3), (unsigned)(i % 3), (double)x, (double)lim[i % 3]);
    }
    if (vel3)
      for (uint64_t i = 0; i < (uint64_t)n * 3; i++)
        if (!std::isfinite(vel3[i]))
          return failf(MPM_ERR_RANGE, "particle %llu: velocity not finite",
                       (unsigned long long)(i / 3));
    if (C9)
      for (uint64_t i = 0; i < (uint64_t)n * 9; i++)
        if (!std::isfinite(C9[i]))
          return failf(MPM_ERR_RANGE, "particle %llu: C not finite", (unsigned long long)(i / 9));
    if (mass)
      for (uint64_t i = 0; i < n; i++) {
        if (!std::isfinite(mass[i]))
          return failf(MPM_ERR_RANGE, "particle %llu: mass not finite", (unsigned long long)i);
        if (mass[i] != 1.0f) unit = 0;
      }
  }
  HIPCHK(hipSetDevice(c.device));
  HIPCHK(hipStreamSynchronize(s->stream));
  // A captured step depends on the host state only through the launch
  // parameters (particle count, unit mass; walls and constants drop the
  // graphs where they change) and its start state (cur, since_sort), so a
  // reset to the same count and mass kind keeps the graphs.
  if (n != s->n || unit != s->unit_mass) scene_drop_graph(s);
  s->version++;
  s->n = n;
  s->cur = 0;
  s->steps = 0;
  s->since_sort = 0;
  s->sorted_once = false;
  s->unit_mass = unit;
  PState S = pstate(s, 0);
  if (n) {
    const float *dp = pos3, *dv = vel3, *dc = C9, *dm = mass;
    if (!device_ptrs) {
      size_t need = (size_t)n * 4 * (3 + (vel3 ? 3 : 0) + (C9 ? 9 : 0) + (mass ? 1 : 0));
      int rc = ensure_stage(s, need);
      if (rc) return rc;
      char *st = (char *)s->stage;
      size_t off = 0;
      auto put = [&](const float *h, size_t cnt, const float *&d) -> int {
        if (!h) {
          d = nullptr;
          return MPM_OK;
        }
        HIPCHK(hipMemcpyAsync(st + off, h, cnt * 4, hipMemcpyHostToDevice, s->stream));
        d = (const float *)(st + off);
        off += cnt * 4;
        return MPM_OK;
      };
      int r2 = put(pos3, (size_t)n * 3, dp);
      if (!r2) r2 = put(vel3, (size_t)n * 3, dv);
      if (!r2) r2 = put(C9, (size_t)n * 9, dc);
      if (!r2) r2 = put(mass, (size_t)n, dm);
      if (r2) return r2;
    }
    launch_import_particles(S, n, dp, dv, dc, dm, s->stream);
    HIPCHK(hipGetLastError());
  }
  int rc = zero_grids(s);
  if (rc) return rc;
  // On the brick path the caller's order is sorted into bricks here, with
  // the upload (after the brick flags were cleared: the sort writes the
  // occupancy of its work-item set), so the first step is an ordinary one;
  // the steps' kernels and results are those of sorting in the first step.
  if (n && !(c.flags & MPM_FLAG_NO_LDS) && c.sort_period) {
    const Params P = scene_params(s);
    launch_sort(P, S, pstate(s, 1), s->sort, bw_of(s, 0), s->tc, s->stream);
    HIPCHK(hipGetLastError());
    s->cur = 1;
    s->sorted_once = true;
  }
  rc = reset_diffuse(s);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

// Add one step's 8 events (start, sort, P2G mass, P2G momentum, update,
// relayout (an empty interval: the re-sort's preparation runs inside the P2G
// stages), G2P, diffuse) to the stage sums.
static int fold_step_events(mpm_scene *s, hipEvent_t *e) {
  const int E = mpm_scene::kEvPerStep;
  float ms[E - 1];
  for (int i = 0; i < E - 1; i++) HIPCHK(hipEventElapsedTime(&ms[i], e[i], e[i + 1]));
  float tot;
  HIPCHK(hipEventElapsedTime(&tot, e[0], e[E - 1]));
  s->stage_sum[5] += ms[0];
  s->stage_sum[0] += ms[1];
  s->stage_sum[1] += ms[2];
  s->stage_sum[2] += ms[3];
  s->stage_sum[7] += ms[4];
  s->stage_sum[3] += ms[5];
  s->stage_sum[4] += ms[6];
  s->stage_sum[6] += tot;
  s->step_ms.push_back(tot);
  s->timed_steps++;
  return MPM_OK;
}

static int fold_graph_events(mpm_scene *s);

// Fold the steps recorded by eager enqueues.
static int fold_eager(mpm_scene *s) {
  const int E = mpm_scene::kEvPerStep;
  for (int k = 0; k < s->ev_used; k++) {
    hipEvent_t *e = &s->ev[k * E];
    HIPCHK(hipEventSynchronize(e[E - 1]));
    int rc = fold_step_events(s, e);
    if (rc) return rc;
  }
  s->ev_used = 0;
  return MPM_OK;
}

// Fold every recorded step (eager pool and pending graph replay).
static int fold_timing(mpm_scene *s) {
  int rc = fold_eager(s);
  if (rc) return rc;
  return fold_graph_events(s);
}

// Enqueue one step.  e (optional) receives 8 events: start, after the sort,
// P2G mass, P2G momentum, grid update, (empty) relayout, G2P, diffuse.
// When captured (graph_events), the events become event-record nodes.
static int enqueue_step(mpm_scene *s, hipEvent_t *e, bool graph_events = false) {
  const bool keep = (s->cfg.flags & MPM_FLAG_KEEP_GRID) != 0;
  Params P = scene_params(s);
  PState S = pstate(s, s->cur);
  hipStream_t st = s->stream;
  auto mark = [&](int i) -> int {
    if (e) {
      if (graph_events) HIPCHK(hipEventRecordWithFlags(e[i], st, hipEventRecordExternal));
      else HIPCHK(hipEventRecord(e[i], st));
    }
    return MPM_OK;
  };
  int rc = mark(0);
  if (rc) return rc;
  const uint32_t period = s->cfg.sort_period;
  // The LDS-staged scatter works on brick-sorted particles; without a sort
  // period the direct per-particle form runs, and so does a step without
  // particles (its dense update leaves the whole grid as an empty step does,
  // whatever the brick flags say).
  const bool lds = !(s->cfg.flags & MPM_FLAG_NO_LDS) && period != 0;
  const bool tiled = lds && s->n;
  // Layout policy: particles not sorted since they were loaded or changed
  // (the direct path after a reset; appended or deleted particles) take the
  // full sort first; afterwards every `period` steps the P2G mass pass
  // records new bricks and G2P writes the new state in sorted order.  The
  // direct path has no key recording and sorts with the full sort.
  bool relayout = false;
  if (period && s->n) {
    bool due = !s->sorted_once || s->since_sort >= period;
    if (due && (!tiled || !s->sorted_once)) {
      PState dst = pstate(s, s->cur ^ 1);
      launch_sort(P, S, dst, s->sort, bw_of(s, s->cur), s->tc, st);
      s->cur ^= 1;
      S = dst;
      s->since_sort = 0;
      s->sorted_once = true;
    } else if (due) {
      relayout = true;
    }
  }
  const BrickWork bw = bw_of(s, s->cur);
  if ((rc = mark(1))) return rc;
  // the dense update expects zero integer grids; the brick form leaves them
  // zero only with KEEP_GRID off and outside the bricks it resolved
  if ((keep || lds) && !tiled) launch_clear_grid(s->g, s->G, st);
  KeyOut ko{};
  if (relayout) {
    ko = key_out(s->sort, bw);
    ko.acc = s->sort.gacc;
    ko.cols = s->tc.ppt;
  }
  if (tiled) launch_p2g_mass_tiled(P, S, s->g, s->tc, bw, ko, st);
  else launch_p2g_mass_global(P, S, s->g, st);
  // A re-sort's scan (offsets, next work items) runs inside the brick-list
  // and mass-resolve kernels and its inverse permutation inside P2G
  // momentum: one stream, no extra launches.
  if (tiled) launch_resolve_mass(P, s->g, bw, st, relayout ? scan_args(s->sort, s->tc) : ScanArgs{});
  if ((rc = mark(2))) return rc;
  if (tiled)
    launch_p2g_mom_tiled(P, S, s->g, s->tc, bw, st,
                         relayout ? src_out(s->sort, s->tc, pstate(s, s->cur ^ 1).density)
                                  : SrcOut{});
  else launch_p2g_mom_global(P, S, s->g, st);
  if ((rc = mark(3))) return rc;
  if (tiled) launch_resolve_update(P, s->g, s->tc, bw, s->walls, keep, st);
  else launch_update_dense(P, s->g, s->walls, !keep, st);
  if (s->hand_on)
    launch_apply_hand(P, s->g, lds ? s->bw.act : nullptr, s->bw.nbx, s->bw.nby, s->hand_dev, st);
  if ((rc = mark(4))) return rc;
  if ((rc = mark(5))) return rc;
  // with a diffuse pool, the tiled G2P also decides the spawners
  const GenOut go = (tiled && P.dmax) ? GenOut{s->d.seed, s->d.genbits, s->d.genval} : GenOut{};
  if (relayout) {
    PState dst = pstate(s, s->cur ^ 1);
    launch_g2p_tiled(P, S, dst, s->g, s->walls, s->tc, bw, s->sort.src_of, st, go);
    s->cur ^= 1;
    S = dst;
    s->since_sort = 0;
  } else if (tiled) {
    launch_g2p_tiled(P, S, S, s->g, s->walls, s->tc, bw, nullptr, st, go);
  } else {
    launch_g2p_global(P, S, s->g, s->walls, st);
  }
  if ((rc = mark(6))) return rc;
  // the resolve list starts every step empty (the re-sorting G2P empties it)
  if (tiled && !relayout) launch_clear_list(bw, st);
  launch_diffuse(P, S, s->g, s->walls, s->d, st, go.bits != nullptr);
  if ((rc = mark(7))) return rc;
  HIPCHK(hipGetLastError());
  s->steps++;
  s->since_sort++;
  return MPM_OK;
}

static void drop_set(mpm_scene::GraphSet &gs) {
  for (hipGraphExec_t &g : gs.exec) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

// Releases every captured graph after the stream has drained (a replay may
// still run) and the pending timed replay's events are folded.
void scene_drop_graph(mpm_scene *s) {
  bool any = false;
  for (const mpm_scene::GraphSet &gs : s->gset) any = any || gs.exec[0];
  if (any && s->stream) {
    (void)hipSetDevice(s->cfg.device);
    (void)hipStreamSynchronize(s->stream);
    (void)fold_graph_events(s);
  }
  for (mpm_scene::GraphSet &gs : s->gset) drop_set(gs);
  s->graph_len = 0;
  s->graph_next = 0;
  s->graph_pending = -1;
}

// Steps per captured graph: a whole number of re-sort cycles (the buffer
// parity flips at every re-sort, so 2 * period steps return to the start).
static uint32_t graph_steps(uint32_t period) {
  if (period == 0) return 0;  // direct form: never captured
  uint32_t cyc = 2 * period;
  uint32_t k = (16 + cyc - 1) / cyc;
  return cyc * k;
}

// Capture graph_steps() steps from the current layout state into gs.exec[half];
// timed graphs record their events into ev[half * L * kEvPerStep ...].
static int capture_one(mpm_scene *s, mpm_scene::GraphSet &gs, int half, bool timed) {
  const int cur0 = s->cur;
  const uint32_t since0 = s->since_sort;
  const uint64_t steps0 = s->steps;
  const uint32_t L = graph_steps(s->cfg.sort_period);
  HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int rc = MPM_OK;
  for (uint32_t i = 0; i < L && !rc; i++) {
    hipEvent_t *e = timed ? &s->ev[((size_t)half * L + i) * mpm_scene::kEvPerStep] : nullptr;
    rc = enqueue_step(s, e, true);
  }
  hipGraph_t g = nullptr;
  hipError_t err = hipStreamEndCapture(s->stream, &g);
  gs.post_cur = s->cur;
  gs.post_since = s->since_sort;
  // capture only recorded the work: put the host state back
  s->cur = cur0;
  s->since_sort = since0;
  s->steps = steps0;
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (err != hipSuccess) return hip_fail(err, "hipStreamEndCapture");
  err = hipGraphInstantiate(&gs.exec[half], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (err != hipSuccess) {
    gs.exec[half] = nullptr;
    return hip_fail(err, "hipGraphInstantiate");
  }
  return MPM_OK;
}

// The graph set for the current start state, captured if absent.
static int find_graphs(mpm_scene *s, mpm_scene::GraphSet **out) {
  mpm_scene::GraphSet *lru = &s->gset[0];
  for (mpm_scene::GraphSet &gs : s->gset) {
    if (gs.exec[0] && gs.timed == s->timing && gs.cur == s->cur && gs.since == s->since_sort) {
      gs.used = ++s->graph_clock;
      *out = &gs;
      return MPM_OK;
    }
    if (!gs.exec[0] || gs.used < lru->used) lru = &gs;
  }
  if (lru->exec[0]) {
    // the replaced graphs may still run
    HIPCHK(hipStreamSynchronize(s->stream));
    drop_set(*lru);
  }
  for (int half = 0; half < (s->timing ? 2 : 1); half++) {
    int rc = capture_one(s, *lru, half, s->timing);
    if (rc) {
      drop_set(*lru);
      return rc;
    }
  }
  lru->timed = s->timing;
  lru->cur = s->cur;
  lru->since = s->since_sort;
  lru->used = ++s->graph_clock;
  s->graph_len = graph_steps(s->cfg.sort_period);
  *out = lru;
  return MPM_OK;
}

// Fold the events of a finished timed replay (waits for it).
static int fold_graph_events(mpm_scene *s) {
  if (s->graph_pending < 0) return MPM_OK;
  const int E = mpm_scene::kEvPerStep;
  const size_t base = (size_t)s->graph_pending * s->graph_len * E;
  for (uint32_t k = 0; k < s->graph_len; k++) {
    hipEvent_t *e = &s->ev[base + (size_t)k * E];
    HIPCHK(hipEventSynchronize(e[E - 1]));
    int rc = fold_step_events(s, e);
    if (rc) return rc;
  }
  s->graph_pending = -1;
  return MPM_OK;
}

int scene_step(mpm_scene *s, uint32_t nsteps) {
  HIPCHK(hipSetDevice(s->cfg.device));
  s->version++;
  const uint32_t period = s->cfg.sort_period;
  const bool tiled = !(s->cfg.flags & MPM_FLAG_NO_LDS) && period != 0;
  while (nsteps > 0) {
    const uint32_t L = graph_steps(period);
    const bool graphable = L > 0 && !(s->cfg.flags & MPM_FLAG_NO_GRAPH) && tiled && s->sorted_once &&
                           s->n > 0 && (!s->timing || 2 * L <= (uint32_t)mpm_scene::kEvSteps);
    if (graphable && nsteps >= L) {
      // eager-step events share the pool with the graphs' events
      int rc = fold_eager(s);
      if (rc) return rc;
      mpm_scene::GraphSet *gs = nullptr;
      if ((rc = find_graphs(s, &gs))) return rc;
      if (s->timing) {
        const int half = s->graph_next;
        HIPCHK(hipGraphLaunch(gs->exec[half], s->stream));
        // the replay that used the other half was enqueued before this one
        if ((rc = fold_graph_events(s))) return rc;
        s->graph_pending = half;
        s->graph_next = half ^ 1;
      } else {
        HIPCHK(hipGraphLaunch(gs->exec[0], s->stream));
      }
      s->cur = gs->post_cur;
      s->since_sort = gs->post_since;
      s->steps += s->graph_len;
      nsteps -= s->graph_len;
      continue;
    }
    hipEvent_t *e = nullptr;
    if (s->timing) {
      int rc = fold_graph_events(s);
      if (rc) return rc;
      if (s->ev_used == mpm_scene::kEvSteps) {
        if ((rc = fold_eager(s))) return rc;
      }
      e = &s->ev[s->ev_used * mpm_scene::kEvPerStep];
      s->ev_used++;
    }
    int rc = enqueue_step(s, e);
    if (rc) return rc;
    nsteps--;
  }
  return MPM_OK;
}

// Capture, without running anything, the graphs that the next multi-step
// call will replay: those of the current start state, or, when the next step
// must first sort the particles (eagerly: after an append or a delete),
// those of the state that step leaves (buffers swapped, one step since the
// sort).
int scene_prepare(mpm_scene *s) {
  const uint32_t period = s->cfg.sort_period;
  const bool tiled = !(s->cfg.flags & MPM_FLAG_NO_LDS) && period != 0;
  const uint32_t L = graph_steps(period);
  if (s->batch || !L || (s->cfg.flags & MPM_FLAG_NO_GRAPH) || !tiled || !s->n ||
      (s->timing && 2 * L > (uint32_t)mpm_scene::kEvSteps))
    return MPM_OK;
  HIPCHK(hipSetDevice(s->cfg.device));
  const int cur0 = s->cur;
  const uint32_t since0 = s->since_sort;
  const bool sorted0 = s->sorted_once;
  if (!s->sorted_once) {
    s->cur ^= 1;
    s->since_sort = 1;
    s->sorted_once = true;
  }
  int rc = MPM_OK;
  if (s->timing) rc = fold_timing(s);
  mpm_scene::GraphSet *gs = nullptr;
  if (!rc) rc = find_graphs(s, &gs);
  s->cur = cur0;
  s->since_sort = since0;
  s->sorted_once = sorted0;
  return rc;
}

// The particles in reset order as one block of 18 floats per particle (pos4,
// vel4, C9, density planes), without waiting: the export kernel on stream
// st, then one copy of the block into dst (host memory, pinned for the copy
// to run beside other work) on stream copy, after event `exported`; event
// `done` marks the copy's end.  The scene's buffers must outlive the copy.
int scene_export_async(mpm_scene *s, hipStream_t st, hipStream_t copy, void *dst,
                       hipEvent_t exported, hipEvent_t done) {
  HIPCHK(hipSetDevice(s->cfg.device));
  const size_t n = s->n;
  const size_t need = n * 4 * (4 + 4 + 9 + 1);
  int rc = ensure_stage(s, need ? need : 256);
  if (rc) return rc;
  char *stg = (char *)s->stage;
  launch_export_particles(pstate(s, s->cur), (uint32_t)n, (float4 *)stg, (float4 *)(stg + n * 16),
                          (float *)(stg + n * 32), (float *)(stg + n * 68), st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(exported, st));
  HIPCHK(hipStreamWaitEvent(copy, exported, 0));
  if (need) HIPCHK(hipMemcpyAsync(dst, stg, need, hipMemcpyDeviceToHost, copy));
  HIPCHK(hipEventRecord(done, copy));
  return MPM_OK;
}

int scene_read_particles(mpm_scene *s, float *pos4, float *vel4, float *C9, float *density) {
  HIPCHK(hipSetDevice(s->cfg.device));
  const uint32_t n = s->n;
  if (!n) return MPM_OK;
  size_t need = (size_t)n * 4 * (4 + 4 + 9 + 1);
  int rc = ensure_stage(s, need);
  if (rc) return rc;
  char *st = (char *)s->stage;
  float4 *dpos = pos4 ? (float4 *)st : nullptr;
  float4 *dvel = vel4 ? (float4 *)(st + (size_t)n * 16) : nullptr;
  float *dC = C9 ? (float *)(st + (size_t)n * 32) : nullptr;
  float *dd = density ? (float *)(st + (size_t)n * 68) : nullptr;
  launch_export_particles(pstate(s, s->cur), n, dpos, dvel, dC, dd, s->stream);
  HIPCHK(hipGetLastError());
  if (pos4) HIPCHK(hipMemcpyAsync(pos4, dpos, (size_t)n * 16, hipMemcpyDeviceToHost, s->stream));
  if (vel4) HIPCHK(hipMemcpyAsync(vel4, dvel, (size_t)n * 16, hipMemcpyDeviceToHost, s->stream));
  if (C9) HIPCHK(hipMemcpyAsync(C9, dC, (size_t)n * 36, hipMemcpyDeviceToHost, s->stream));
  if (density)
    HIPCHK(hipMemcpyAsync(density, dd, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

int scene_read_grid(mpm_scene *s, int32_t *m, int32_t *mv3, float *mf, float *v3) {
  if ((m || mv3) && !(s->cfg.flags & MPM_FLAG_KEEP_GRID))
    return failf(MPM_ERR_STATE,
                 "i32 grids are cleared inside the step; create the scene with "
                 "MPM_FLAG_KEEP_GRID to read them");
  HIPCHK(hipSetDevice(s->cfg.device));
  const size_t G = s->G;
  int rc = ensure_stage(s, G * 4 * 8);
  if (rc) return rc;
  char *st = (char *)s->stage;
  int32_t *dm = m ? (int32_t *)st : nullptr;
  int32_t *dmv = mv3 ? (int32_t *)(st + G * 4) : nullptr;
  float *dmf = mf ? (float *)(st + G * 16) : nullptr;
  float *dv = v3 ? (float *)(st + G * 20) : nullptr;
  launch_export_grid(s->g, (uint32_t)G, dm, dmv, dmf, dv, s->stream);
  HIPCHK(hipGetLastError());
  if (m) HIPCHK(hipMemcpyAsync(m, dm, G * 4, hipMemcpyDeviceToHost, s->stream));
  if (mv3) HIPCHK(hipMemcpyAsync(mv3, dmv, G * 12, hipMemcpyDeviceToHost, s->stream));
  if (mf) HIPCHK(hipMemcpyAsync(mf, dmf, G * 4, hipMemcpyDeviceToHost, s->stream));
  if (v3) HIPCHK(hipMemcpyAsync(v3, dv, G * 12, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

int scene_read_diffuse(mpm_scene *s, void *rec48, uint32_t *free_list, int32_t *top) {
  const uint32_t D = s->cfg.diffuse_capacity;
  if (!D) return failf(MPM_ERR_STATE, "scene has no diffuse pool (diffuse_capacity = 0)");
  HIPCHK(hipSetDevice(s->cfg.device));
  if (rec48)
    HIPCHK(hipMemcpyAsync(rec48, s->d.pool, (size_t)D * 48, hipMemcpyDeviceToHost, s->stream));
  if (free_list)
    HIPCHK(hipMemcpyAsync(free_list, s->d.free_list, (size_t)D * 4, hipMemcpyDeviceToHost,
                          s->stream));
  if (top) HIPCHK(hipMemcpyAsync(top, s->d.top, 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

int scene_write_diffuse(mpm_scene *s, const void *rec48, const uint32_t *free_list, int32_t top) {
  const uint32_t D = s->cfg.diffuse_capacity;
  if (!D) return failf(MPM_ERR_STATE, "scene has no diffuse pool (diffuse_capacity = 0)");
  if (top < 0 || (uint32_t)top > D)
    return failf(MPM_ERR_INVALID, "free-list top %d outside [0, %u]", top, D);
  if (free_list)
    for (uint32_t i = 0; i < D; i++)
      if (free_list[i] >= D)
        return failf(MPM_ERR_INVALID, "free_list[%u] = %u is not a slot index", i, free_list[i]);
  HIPCHK(hipSetDevice(s->cfg.device));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (rec48) {
    HIPCHK(hipMemcpyAsync(s->d.pool, rec48, (size_t)D * 48, hipMemcpyHostToDevice, s->stream));
    launch_alive_bits(s->d, D, s->stream);
    HIPCHK(hipGetLastError());
  }
  if (free_list)
    HIPCHK(hipMemcpyAsync(s->d.free_list, free_list, (size_t)D * 4, hipMemcpyHostToDevice,
                          s->stream));
  HIPCHK(hipMemcpyAsync(s->d.top, &top, 4, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

int scene_set_seed(mpm_scene *s, uint32_t seed) {
  s->cfg.seed = seed;
  if (!s->cfg.diffuse_capacity) return MPM_OK;
  HIPCHK(hipSetDevice(s->cfg.device));
  HIPCHK(hipMemcpyAsync(s->d.seed, &seed, 4, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

// ---------------------------------------------------------------------------
// Interaction (fluid-gpu.js:145-180, 479-638).

// One allocation for the hand parameters and the delete / raycast scratch.
static int ensure_interact(mpm_scene *s) {
  if (s->hand_dev) return MPM_OK;
  const size_t cap = s->cfg.capacity ? s->cfg.capacity : 1;
  const size_t nw = (cap + 31) / 32, sw = scan_scratch_words((uint32_t)nw);
  const size_t bytes = 256 + 256 + (2 * nw + sw) * 4;
  char *base = nullptr;
  DALLOC(base, bytes);
  s->hand_dev = (HandDev *)base;
  s->itotal = (uint32_t *)(base + 256);
  s->ihit = s->itotal + 1;
  s->ibits = (uint32_t *)(base + 512);
  s->iprefix = s->ibits + nw;
  s->iscratch = s->iprefix + nw;
  HandDev off{};
  off.bmax[0] = off.bmax[1] = off.bmax[2] = -1;
  HIPCHK(hipMemcpy(s->hand_dev, &off, sizeof off, hipMemcpyHostToDevice));
  return MPM_OK;
}

// The box of grid cells the push can reach: the line through the hand along
// the eye ray clipped to the grid grown by the radius (lineBoxIntersect,
// math.js:149-164), its ends grown by the radius again
// (FluidGPU.updateHand, fluid-gpu.js:494-523).  In double, like the JS.
static void hand_box(const mpm_config &c, const mpm_hand &h, HandDev &d) {
  for (int a = 0; a < 3; a++) {
    d.bmin[a] = 0;
    d.bmax[a] = -1;
  }
  const double pos[3] = {h.pos[0], h.pos[1], h.pos[2]};
  const double g[3] = {(double)c.grid_x, (double)c.grid_y, (double)c.grid_z};
  double dir[3] = {pos[0] - h.eye[0], pos[1] - h.eye[1], pos[2] - h.eye[2]};
  double dl = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  if (dl == 0.0) dl = 1.0;
  for (double &v : dir) v /= dl;
  const double R = h.radius;
  double tmin = -INFINITY, tmax = INFINITY;
  for (int a = 0; a < 3; a++) {
    const double lo = -R, hi = g[a] + R;
    if (std::fabs(dir[a]) < 1e-9) {
      if (pos[a] < lo || pos[a] > hi) return;
      continue;
    }
    double t1 = (lo - pos[a]) / dir[a], t2 = (hi - pos[a]) / dir[a];
    if (t1 > t2) std::swap(t1, t2);
    tmin = std::max(tmin, t1);
    tmax = std::min(tmax, t2);
    if (tmin > tmax) return;
  }
  for (int a = 0; a < 3; a++) {
    const double p0 = pos[a] + tmin * dir[a], p1 = pos[a] + tmax * dir[a];
    d.bmin[a] = (int32_t)std::max(0.0, std::floor(std::min(p0, p1) - R));
    d.bmax[a] = (int32_t)std::min(g[a] - 1.0, std::ceil(std::max(p0, p1) + R));
  }
}

int scene_set_hand(mpm_scene *s, const mpm_hand *hand) {
  const bool active = hand && hand->active;
  if (active) {
    const float v[] = {hand->pos[0], hand->pos[1], hand->pos[2], hand->radius, hand->vel[0],
                       hand->vel[1], hand->vel[2], hand->strength, hand->eye[0], hand->eye[1],
                       hand->eye[2]};
    for (float x : v)
      if (!std::isfinite(x)) return failf(MPM_ERR_INVALID, "mpm_set_hand: values must be finite");
    if (s->batch)
      return failf(MPM_ERR_STATE, "mpm_set_hand: the batched step applies no hand push");
  }
  if (!active && !s->hand_on) return MPM_OK;
  HIPCHK(hipSetDevice(s->cfg.device));
  int rc = ensure_interact(s);
  if (rc) return rc;
  HandDev d{};
  if (active) {
    d.px = hand->pos[0];
    d.py = hand->pos[1];
    d.pz = hand->pos[2];
    d.radius = hand->radius;
    d.vx = hand->vel[0];
    d.vy = hand->vel[1];
    d.vz = hand->vel[2];
    d.strength = hand->strength;
    d.ex = hand->eye[0];
    d.ey = hand->eye[1];
    d.ez = hand->eye[2];
    d.on = 1u;
    hand_box(s->cfg, *hand, d);
  }
  // ordered after the steps already enqueued, before the next ones
  HIPCHK(hipMemcpyAsync(s->hand_dev, &d, sizeof d, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (active && !s->hand_on) {
    scene_drop_graph(s);  // the step gains the hand kernel
    s->hand_on = true;
    s->version++;
  } else if (!active) {
    // steps captured with the kernel stay correct: it reads `on` = 0
    s->hand_on = false;
  }
  return MPM_OK;
}

static Ray make_ray(const float *o, const float *d, float radius) {
  Ray r{};
  r.ox = o[0];
  r.oy = o[1];
  r.oz = o[2];
  r.dx = d[0];
  r.dy = d[1];
  r.dz = d[2];
  r.radius = radius;
  return r;
}

static int check_ray(const float *o, const float *d, float radius, const char *what) {
  if (!o || !d) return failf(MPM_ERR_INVALID, "%s: origin or dir is NULL", what);
  for (int a = 0; a < 3; a++)
    if (!std::isfinite(o[a]) || !std::isfinite(d[a]))
      return failf(MPM_ERR_INVALID, "%s: origin and dir must be finite", what);
  if (!(radius >= 0.0f) || !std::isfinite(radius))
    return failf(MPM_ERR_INVALID, "%s: radius must be finite and >= 0", what);
  return MPM_OK;
}

int scene_delete_near(mpm_scene *s, const float *origin, const float *dir, float radius,
                      uint32_t *remaining) {
  int rc = check_ray(origin, dir, radius, "mpm_delete_near");
  if (rc) return rc;
  HIPCHK(hipSetDevice(s->cfg.device));
  HIPCHK(hipStreamSynchronize(s->stream));
  if ((rc = ensure_interact(s))) return rc;
  const uint32_t n = s->n, nw = (n + 31) / 32;
  uint32_t kept = 0;
  if (n) {
    launch_delete_mark_move(pstate(s, s->cur), pstate(s, s->cur ^ 1), n,
                            make_ray(origin, dir, radius), s->ibits, s->iprefix, s->iscratch,
                            s->itotal, nw, s->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&kept, s->itotal, 4, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    // the survivors are in reset order in the other buffer: re-sort next step
    scene_drop_graph(s);
    s->cur ^= 1;
    s->n = kept;
    s->since_sort = 0;
    s->sorted_once = false;
    s->version++;
    if (!kept && s->batch) {
      // batched steps skip scenes without particles: leave the grid as the
      // next (empty) step would
      if ((rc = zero_grids(s))) return rc;
      HIPCHK(hipStreamSynchronize(s->stream));
    }
  }
  if (remaining) *remaining = kept;
  return MPM_OK;
}

int scene_raycast(mpm_scene *s, const float *origin, const float *dir, float radius, float *t,
                  int *hit) {
  int rc = check_ray(origin, dir, radius, "mpm_raycast");
  if (rc) return rc;
  HIPCHK(hipSetDevice(s->cfg.device));
  if ((rc = ensure_interact(s))) return rc;
  launch_raycast(pstate(s, s->cur), s->n, make_ray(origin, dir, radius), s->ihit, s->stream);
  HIPCHK(hipGetLastError());
  uint32_t bits = 0xffffffffu;
  HIPCHK(hipMemcpyAsync(&bits, s->ihit, 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  const int h = bits != 0xffffffffu;
  if (hit) *hit = h;
  if (t) {
    float f = 0.0f;
    if (h) std::memcpy(&f, &bits, 4);
    *t = f;
  }
  return MPM_OK;
}

int scene_append(mpm_scene *s, uint32_t k, const float *pos3, const float *vel3) {
  const mpm_config &c = s->cfg;
  if (!k) return MPM_OK;
  if (!pos3) return failf(MPM_ERR_INVALID, "mpm_append_particles: pos3 is NULL");
  if ((uint64_t)s->n + k > c.capacity)
    return failf(MPM_ERR_CAPACITY, "append of %u particles to %u exceeds capacity %u", k, s->n,
                 c.capacity);
  const float lim[3] = {(float)c.grid_x - 1.0f, (float)c.grid_y - 1.0f, (float)c.grid_z - 1.0f};
  for (uint64_t i = 0; i < (uint64_t)k * 3; i++) {
    const float x = pos3[i];
    if (!(x >= 1.0f && x < lim[i % 3]))
      return failf(MPM_ERR_RANGE, "appended particle %llu: coordinate %u = %g outside [1, %g)",
                   (unsigned long long)(i / 3), (unsigned)(i % 3), (double)x,
                   (double)lim[i % 3]);
  }
  if (vel3)
    for (uint64_t i = 0; i < (uint64_t)k * 3; i++)
      if (!std::isfinite(vel3[i]))
        return failf(MPM_ERR_RANGE, "appended particle %llu: velocity not finite",
                     (unsigned long long)(i / 3));
  HIPCHK(hipSetDevice(c.device));
  HIPCHK(hipStreamSynchronize(s->stream));
  int rc = ensure_stage(s, (size_t)k * 4 * (vel3 ? 6 : 3));
  if (rc) return rc;
  float *dp = (float *)s->stage, *dv = vel3 ? dp + (size_t)k * 3 : nullptr;
  HIPCHK(hipMemcpyAsync(dp, pos3, (size_t)k * 12, hipMemcpyHostToDevice, s->stream));
  if (vel3) HIPCHK(hipMemcpyAsync(dv, vel3, (size_t)k * 12, hipMemcpyHostToDevice, s->stream));
  launch_import_particles(pstate(s, s->cur), k, dp, dv, nullptr, nullptr, s->stream, s->n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  // the new particles sit after the sorted ones: re-sort next step
  scene_drop_graph(s);
  s->n += k;
  s->since_sort = 0;
  s->sorted_once = false;
  s->version++;
  return MPM_OK;
}

// ---------------------------------------------------------------------------
// Renderer handoff (fluid-renderer.js:232, 256, 315-316, 850-856).

int scene_export_device(mpm_scene *s, const mpm_export *dst, uint32_t flags) {
  if (!dst) return failf(MPM_ERR_INVALID, "mpm_export_device: dst is NULL");
  if (flags & ~MPM_EXPORT_RESET_ORDER)
    return failf(MPM_ERR_INVALID, "mpm_export_device: flags 0x%x (only MPM_EXPORT_RESET_ORDER)",
                 flags);
  if (dst->diffuse48 && !s->cfg.diffuse_capacity)
    return failf(MPM_ERR_STATE, "scene has no diffuse pool (diffuse_capacity = 0)");
  HIPCHK(hipSetDevice(s->cfg.device));
  launch_export_particles(pstate(s, s->cur), s->n, (float4 *)dst->pos4, (float4 *)dst->vel4,
                          nullptr, dst->density, s->stream,
                          (flags & MPM_EXPORT_RESET_ORDER) != 0);
  if (dst->diffuse48)
    launch_copy16(s->d.pool, dst->diffuse48, (size_t)s->cfg.diffuse_capacity * 48, s->stream);
  HIPCHK(hipGetLastError());
  return MPM_OK;
}

// The two frames, at first use: staging and pinned memory for every part at
// capacity, so a frame never reallocates.
static int ensure_frames(mpm_scene *s) {
  if (s->copy_stream) return MPM_OK;
  const size_t cap = s->cfg.capacity, D = s->cfg.diffuse_capacity;
  s->frame_bytes = cap * 36 + D * 48 + 4 * 256;
  // (a failure part way leaves what was made, for the next call to finish)
  for (mpm_scene::Frame &f : s->frames) {
    if (!f.dev) DALLOC(f.dev, s->frame_bytes);
    if (!f.host) {
      hipError_t e = hipHostMalloc((void **)&f.host, s->frame_bytes, hipHostMallocDefault);
      if (e != hipSuccess) {
        f.host = nullptr;
        (void)hipGetLastError();
        return failf(MPM_ERR_NOMEM, "pinned hipHostMalloc(%zu) failed: %s", s->frame_bytes,
                     hipGetErrorString(e));
      }
    }
    if (!f.ready) HIPCHK(hipEventCreateWithFlags(&f.ready, hipEventDisableTiming));
    if (!f.done) HIPCHK(hipEventCreateWithFlags(&f.done, hipEventDisableTiming));
  }
  HIPCHK(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
  return MPM_OK;
}

int scene_frame_begin(mpm_scene *s, uint32_t what, uint64_t *ticket) {
  if (!ticket) return failf(MPM_ERR_INVALID, "mpm_frame_begin: ticket is NULL");
  if (what & ~(MPM_EXPORT_ALL | MPM_EXPORT_RESET_ORDER) || !(what & MPM_EXPORT_ALL))
    return failf(MPM_ERR_INVALID, "mpm_frame_begin: what 0x%x", what);
  const size_t D = s->cfg.diffuse_capacity;
  if ((what & MPM_EXPORT_DIFFUSE) && !D)
    return failf(MPM_ERR_STATE, "scene has no diffuse pool (diffuse_capacity = 0)");
  HIPCHK(hipSetDevice(s->cfg.device));
  int rc = ensure_frames(s);
  if (rc) return rc;
  mpm_scene::Frame &f = s->frames[s->frame_next & 1];
  // this frame's staging and host memory are free once its last copy is
  if (f.ticket) HIPCHK(hipStreamWaitEvent(s->stream, f.done, 0));
  const size_t n = s->n;
  // the parts asked for, packed, so that one copy moves the frame
  size_t o = 0;
  auto part = [&](int k, uint32_t bit, size_t bytes) {
    f.off[k] = o;
    if (what & bit) o = (o + bytes + 255) & ~(size_t)255;
  };
  part(0, MPM_EXPORT_POS, n * 16);
  part(1, MPM_EXPORT_VEL, n * 16);
  part(2, MPM_EXPORT_DENSITY, n * 4);
  part(3, MPM_EXPORT_DIFFUSE, D * 48);
  part(4, MPM_EXPORT_DIFFUSE, 4);
  launch_export_particles(pstate(s, s->cur), (uint32_t)n,
                          (what & MPM_EXPORT_POS) ? (float4 *)(f.dev + f.off[0]) : nullptr,
                          (what & MPM_EXPORT_VEL) ? (float4 *)(f.dev + f.off[1]) : nullptr,
                          nullptr,
                          (what & MPM_EXPORT_DENSITY) ? (float *)(f.dev + f.off[2]) : nullptr,
                          s->stream, (what & MPM_EXPORT_RESET_ORDER) != 0);
  if (what & MPM_EXPORT_DIFFUSE) {
    launch_copy16(s->d.pool, f.dev + f.off[3], D * 48, s->stream);
    HIPCHK(hipMemcpyAsync(f.dev + f.off[4], s->d.top, 4, hipMemcpyDeviceToDevice, s->stream));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(f.ready, s->stream));
  HIPCHK(hipStreamWaitEvent(s->copy_stream, f.ready, 0));
  if (o) HIPCHK(hipMemcpyAsync(f.host, f.dev, o, hipMemcpyDeviceToHost, s->copy_stream));
  HIPCHK(hipEventRecord(f.done, s->copy_stream));
  f.ticket = s->frame_next++;
  f.steps = s->steps;
  f.n = (uint32_t)n;
  f.what = what;
  *ticket = f.ticket;
  return MPM_OK;
}

int scene_frame_poll(mpm_scene *s, uint64_t ticket, int wait, int *ready, mpm_frame *out) {
  if (!ready || !out) return failf(MPM_ERR_INVALID, "mpm_frame_poll: NULL argument");
  *ready = 0;
  mpm_scene::Frame *f = nullptr;
  for (mpm_scene::Frame &x : s->frames)
    if (ticket && x.ticket == ticket) f = &x;
  if (!f)
    return failf(MPM_ERR_NOT_FOUND, "frame %llu is unknown or its memory was reused",
                 (unsigned long long)ticket);
  HIPCHK(hipSetDevice(s->cfg.device));
  if (wait) {
    HIPCHK(hipEventSynchronize(f->done));
  } else {
    hipError_t e = hipEventQuery(f->done);
    if (e == hipErrorNotReady) {
      (void)hipGetLastError();
      return MPM_OK;
    }
    if (e != hipSuccess) return hip_fail(e, "hipEventQuery(frame)");
  }
  const uint32_t w = f->what;
  out->pos4 = (w & MPM_EXPORT_POS) ? (const float *)(f->host + f->off[0]) : nullptr;
  out->vel4 = (w & MPM_EXPORT_VEL) ? (const float *)(f->host + f->off[1]) : nullptr;
  out->density = (w & MPM_EXPORT_DENSITY) ? (const float *)(f->host + f->off[2]) : nullptr;
  out->diffuse48 = (w & MPM_EXPORT_DIFFUSE) ? (const void *)(f->host + f->off[3]) : nullptr;
  out->top = 0;
  if (w & MPM_EXPORT_DIFFUSE) std::memcpy(&out->top, f->host + f->off[4], 4);
  out->n = f->n;
  out->steps = f->steps;
  out->ticket = f->ticket;
  *ready = 1;
  return MPM_OK;
}

int scene_set_walls(mpm_scene *s, const float *wallp) {
  scene_drop_graph(s);
  s->version++;
  if (!wallp) {
    s->nwalls = 0;
    memset(&s->walls, 0, sizeof s->walls);
    return MPM_OK;
  }
  float c = wallp[0];
  if (!(c >= 0.0f) || c > 4.0f)
    return failf(MPM_ERR_INVALID, "wall count %g outside [0, 4]", (double)c);
  for (int i = 0; i < 4 + 4 * 20; i++)
    if (!std::isfinite(wallp[i])) return failf(MPM_ERR_INVALID, "wall parameters must be finite");
  s->nwalls = (uint32_t)c;
  memcpy(&s->walls, wallp + 4, sizeof s->walls);
  return MPM_OK;
}

int scene_enable_timing(mpm_scene *s, bool on) {
  HIPCHK(hipSetDevice(s->cfg.device));
  HIPCHK(hipStreamSynchronize(s->stream));
  // the event pool is created at the first enable
  for (int i = 0; on && i < mpm_scene::kEvPerStep * mpm_scene::kEvSteps; i++)
    if (!s->ev[i]) HIPCHK(hipEventCreate(&s->ev[i]));
  s->timing = on;
  s->ev_used = 0;
  s->graph_pending = -1;
  for (double &v : s->stage_sum) v = 0.0;
  s->timed_steps = 0;
  s->step_ms.clear();
  return MPM_OK;
}

int scene_stage_times(mpm_scene *s, float *ms8) {
  HIPCHK(hipSetDevice(s->cfg.device));
  int rc = fold_timing(s);
  if (rc) return rc;
  for (int i = 0; i < 8; i++)
    ms8[i] = s->timed_steps ? (float)(s->stage_sum[i] / (double)s->timed_steps) : 0.0f;
  return MPM_OK;
}

int scene_tile_stats(mpm_scene *s, uint64_t out[4]) {
  HIPCHK(hipSetDevice(s->cfg.device));
  HIPCHK(hipStreamSynchronize(s->stream));
  uint32_t nit = 0;
  HIPCHK(hipMemcpy(&nit, s->nitems_buf[s->cur], 4, hipMemcpyDeviceToHost));
  std::vector<uint4> it(nit);
  if (nit)
    HIPCHK(hipMemcpy(it.data(), s->items_buf[s->cur], (size_t)nit * 16, hipMemcpyDeviceToHost));
  uint64_t mx = 0;
  for (const uint4 &q : it)
    if (q.z > mx) mx = q.z;
  out[0] = nit;
  out[1] = s->bw.nbricks;
  out[2] = s->n;
  out[3] = mx;
  return MPM_OK;
}

int scene_step_times(mpm_scene *s, float *ms, uint32_t cap, uint32_t *count) {
  HIPCHK(hipSetDevice(s->cfg.device));
  int rc = fold_timing(s);
  if (rc) return rc;
  uint32_t n = (uint32_t)s->step_ms.size();
  if (count) *count = n;
  if (ms)
    for (uint32_t i = 0; i < n && i < cap; i++) ms[i] = s->step_ms[i];
  return MPM_OK;
}

}  // namespace mpm
