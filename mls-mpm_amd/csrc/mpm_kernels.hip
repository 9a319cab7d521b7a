// mpm_kernels.hip -- gfx950 kernels of the MLS-MPM fluid step.
//
// Stage order per step (fluid-gpu.js:663-681, diffuse-gpu.js:148-166):
//   [clear] -> P2G mass -> P2G momentum -> grid update -> G2P
//   -> diffuse generate -> diffuse advect
// The reference's separate COMPUTE_WEIGHTS and DECODE_MASS dispatches are
// folded into the kernels that consume them (weights are recomputed from the
// position, which is cheaper than a 60-byte-per-particle cache; decoded mass
// is produced where it is read).
//
// This file holds the direct forms: one thread per particle (or cell), i32
// global atomics.  They are the always-correct path and the fallback of the
// tiled path in mpm_tiled.hip.
#include "mpm_kernels.h"
#include "mpm_ops.h"

namespace mpm {

// ---------------------------------------------------------------------------
// Grid clear (WGSL_CLEAR, shaders.js:196-210).
__global__ __launch_bounds__(256) void k_clear_grid(GState g, uint32_t G) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G) return;
  g.m[c] = 0;
  g.mx[c] = 0;
  g.my[c] = 0;
  g.mz[c] = 0;
}

// ---------------------------------------------------------------------------
// P2G mass (WGSL_P2G_MASS, shaders.js:242-274): 27 bounds-checked i32 atomics.
template <bool UNIT>
__global__ __launch_bounds__(256) void k_p2g_mass_global(Params P, PState S, GState g) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  PW pw;
  particle_weights(S.at(PX, p), S.at(PY, p), S.at(PZ, p), pw);
  float pm = UNIT ? 1.0f : S.at(PM, p);
#pragma unroll
  for (int k = 0; k < 3; k++) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
      float wxy[3];
#pragma unroll
      for (int i = 0; i < 3; i++) wxy[i] = pw.w[0][i] * pw.w[1][j];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        int idx = cell_idx_checked(pw.b[0] + i, pw.b[1] + j, pw.b[2] + k, P);
        if (idx < 0) continue;
        float w = wxy[i] * pw.w[2][k];
        atomicAdd(&g.m[idx], encode(UNIT ? w : w * pm));
      }
    }
  }
}

// ---------------------------------------------------------------------------
// P2G momentum with global atomics; decodes the cell mass on the fly.
template <bool UNIT>
__global__ __launch_bounds__(256) void k_p2g_mom_global(Params P, PState S, GState g) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  PW pw;
  particle_weights(S.at(PX, p), S.at(PY, p), S.at(PZ, p), pw);
  const uint32_t gx = P.gx, slice = P.gx * P.gy;
  uint32_t base = (uint32_t)((pw.b[2] * (int)P.gy + pw.b[1]) * (int)gx + pw.b[0]);
  p2g_mom_particle<UNIT>(
      P, S, p, pw,
      [&](int k, int j, int i) { return decode(g.m[base + k * slice + j * gx + i]); },
      [&](int k, int j, int i, int32_t ex, int32_t ey, int32_t ez) {
        uint32_t c = base + k * slice + j * gx + i;
        atomicAdd(&g.mx[c], ex);
        atomicAdd(&g.my[c], ey);
        atomicAdd(&g.mz[c], ez);
      });
}

// ---------------------------------------------------------------------------
// Dense grid update over every cell; optionally clears the accumulators it
// consumed (the next step's WGSL_CLEAR, fused).
__global__ __launch_bounds__(256) void k_update_dense(Params P, GState g, Walls W, int clear) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t G = P.gx * P.gy * P.gz;
  if (c >= G) return;
  int32_t im = g.m[c], imx = g.mx[c], imy = g.my[c], imz = g.mz[c];
  int cx = (int)(c % P.gx);
  int cy = (int)((c / P.gx) % P.gy);
  int cz = (int)(c / (P.gx * P.gy));
  g.v[c] = update_cell(P, W, im, imx, imy, imz, cx, cy, cz);
  if (clear) {
    g.m[c] = 0;
    g.mx[c] = 0;
    g.my[c] = 0;
    g.mz[c] = 0;
  }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_g2p_global(Params P, PState S, GState g, Walls W) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  const float4 q = S.rec(p);
  float px = q.x, py = q.y, pz = q.z;
  PW pw;
  particle_weights(px, py, pz, pw);
  const uint32_t gx = P.gx, slice = P.gx * P.gy;
  uint32_t base = (uint32_t)((pw.b[2] * (int)P.gy + pw.b[1]) * (int)gx + pw.b[0]);
  G2POut o;
  g2p_particle(P, W, px, py, pz, pw, [&](int k, int j, int i) {
    float4 v = g.v[base + k * slice + j * gx + i];
    return make_float3(v.x, v.y, v.z);
  }, o);
  store_g2p(S, p, o, __float_as_uint(q.w));
}

// G2P that also applies the new particle order: destination slot d takes the
// particle at source slot src_of[d], whose cell key and rank the P2G mass
// pass recorded.  Reads are gathered from the old order, writes are
// coalesced in the new one; mass, id and density move along.
__global__ __launch_bounds__(256) void k_g2p_relayout(Params P, PState S, PState D, GState g,
                                                      Walls W, const uint32_t *src_of) {
  uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= P.n) return;
  uint32_t p = src_of[d];
  const float4 q = S.rec(p);
  float px = q.x, py = q.y, pz = q.z;
  PW pw;
  particle_weights(px, py, pz, pw);
  const uint32_t gx = P.gx, slice = P.gx * P.gy;
  uint32_t base = (uint32_t)((pw.b[2] * (int)P.gy + pw.b[1]) * (int)gx + pw.b[0]);
  G2POut o;
  g2p_particle(P, W, px, py, pz, pw, [&](int k, int j, int i) {
    float4 v = g.v[base + k * slice + j * gx + i];
    return make_float3(v.x, v.y, v.z);
  }, o);
  store_g2p(D, d, o, __float_as_uint(q.w));
  D.at(PM, d) = S.at(PM, p);
  D.density[d] = S.density[p];
}

// ---------------------------------------------------------------------------
// Diffuse generate (WGSL_DIFFUSE_GENERATE, shaders.js:800-855).
//
// The reference pops free-list slots with unordered atomics (try_spawn,
// shaders.js:791-798).  Here the pops are serialised in particle-index order,
// one of the orders the reference allows: a decide pass marks spawners in a
// bitmask indexed by reset-order id, a scan ranks them, and an apply pass
// gives the r-th spawner free_list[top-1-r] while r < top.
__device__ __forceinline__ bool gen_eval(const Params &P, const PState &S, const GState &g,
                                         uint32_t s, uint32_t seed, GenEval &e) {
  const float density = S.density[s];
  if (density <= 0.0f || density > P.density_bubble_min) return false;
  return gen_decide(P, g, density, S.at(VX, s), S.at(VY, s), S.at(VZ, s), S.at(PX, s),
                    S.at(PY, s), S.at(PZ, s), P.unit_mass ? 1.0f : S.at(PM, s), S.idr(s), seed, e);
}

// Marks the spawners by id and keeps what the apply pass needs of each: its
// slot and the score, speed and density it was decided on.
__global__ __launch_bounds__(256) void k_gen_decide(Params P, PState S, GState g,
                                                    const uint32_t *seedp, uint32_t *bits,
                                                    float4 *val) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.n) return;
  GenEval e;
  if (!gen_eval(P, S, g, s, *seedp + 1u, e)) return;
  gen_mark(GenOut{seedp, bits, val}, S.idr(s), s, e);
}

// The r-th spawner in id order (list[r], from the scan) takes the slot
// free_list[top - 1 - r] while r < top; grid-stride over the spawners.
__global__ __launch_bounds__(256) void k_gen_apply(Params P, PState S, const uint32_t *seedp,
                                                   const uint32_t *list, const uint32_t *total,
                                                   const float4 *val, const uint32_t *free_list,
                                                   const int32_t *topp, DRec *pool,
                                                   uint32_t *alivebits) {
  const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t top0 = *topp;
  const uint32_t nsp = *total, seed = *seedp + 1u;
  const uint32_t lim = top0 > 0 ? (nsp < (uint32_t)top0 ? nsp : (uint32_t)top0) : 0u;
  for (uint32_t r = t0; r < lim; r += gridDim.x * blockDim.x) {
    const int32_t fi = top0 - 1 - (int32_t)r;
    if (fi < 0 || (uint32_t)fi >= P.dmax) continue;  // only on an inconsistent pool
    const uint32_t slot = free_list[fi];
    if (slot >= P.dmax) continue;
    const uint32_t id = list[r];
    const float4 v = val[id];
    const uint32_t s = __float_as_uint(v.x);
    const float score = v.y, speed = v.z, density = v.w;
    const uint32_t rand_seed = id * 9781u + seed;
    float jitter = 0.3f;
    float jx = ((diffuse_rand(rand_seed + 1u) - 0.5f) * jitter) * speed;
    float jy = ((diffuse_rand(rand_seed + 2u) - 0.5f) * jitter) * speed;
    float jz = ((diffuse_rand(rand_seed + 3u) - 0.5f) * jitter) * speed;
    uint32_t ptype = (density < P.density_spray_max) ? 0u : 1u;
    float lifetime = 999.0f;
    if (ptype == 1u)
      lifetime = P.lifetime_mode ? 0.5f + 2.5f * diffuse_rand(rand_seed + 4u)
                                 : P.lifetime_foam * fminf(1.0f, score / P.ke_threshold);
    DRec rec;
    rec.pos = make_float4(S.at(PX, s), S.at(PY, s), S.at(PZ, s), 0.0f);
    rec.vel = make_float4(S.at(VX, s) + jx, S.at(VY, s) + jy, S.at(VZ, s) + jz, 0.0f);
    rec.ptype = ptype;
    rec.alive = 1u;
    rec.lifetime = lifetime;
    rec.pad = 0.0f;
    pool[slot] = rec;
    atomicOr(&alivebits[slot >> 5], 1u << (slot & 31));
  }
}

// No particles: no spawners.
__global__ void k_zero_total(uint32_t *p) { *p = 0u; }

// top -= min(spawners, top); the step's seed becomes the last-used seed.
// The stack after the spawns (the spawned count capped by the free slots)
// and the next seed.  Run by one thread of the advect kernel that follows,
// which reads neither; the new top is also kept in topsnap, the base the
// kill push indexes from while it moves top on.
__device__ __forceinline__ void gen_finish(int32_t *topp, const uint32_t *total, uint32_t *seedp,
                                           int32_t *topsnap) {
  int32_t top = *topp;
  int32_t t = (int32_t)*total;
  int32_t avail = top > 0 ? top : 0;
  top -= t < avail ? t : avail;
  *topp = top;
  *topsnap = top;
  *seedp = *seedp + 1u;
}

// ---------------------------------------------------------------------------
// Diffuse advect (WGSL_DIFFUSE_ADVECT, shaders.js:929-996).  Kills mark a
// bitmask by slot; a scan then pushes the killed slots in slot order.
// sample_grid_velocity (shaders.js:889-911).  All 27 cells are loaded before
// the sum, at indexes clamped into the grid, so the loads overlap; a cell
// outside the grid then adds an exact zero, as the reference skips it (the sum
// starts at +0 and can never become -0, so adding +0 leaves it unchanged).
__device__ __forceinline__ F3 sample_grid_velocity(const Params &P, const GState &g, F3 pos) {
  PW pw;
  particle_weights(pos.x, pos.y, pos.z, pw);
  const int gx = (int)P.gx, gy = (int)P.gy, gz = (int)P.gz;
  int cx[3], cy[3], cz[3];
  bool ox[3], oy[3], oz[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int x = pw.b[0] + a, y = pw.b[1] + a, z = pw.b[2] + a;
    ox[a] = x >= 0 && x < gx;
    oy[a] = y >= 0 && y < gy;
    oz[a] = z >= 0 && z < gz;
    cx[a] = min(max(x, 0), gx - 1);
    cy[a] = min(max(y, 0), gy - 1);
    cz[a] = min(max(z, 0), gz - 1);
  }
  float4 c[27];
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++)
        c[(k * 3 + j) * 3 + i] = g.v[(cz[k] * gy + cy[j]) * gx + cx[i]];
  F3 v = f3(0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const bool in = ox[i] && oy[j] && oz[k];
        const float w = (pw.w[0][i] * pw.w[1][j]) * pw.w[2][k];
        const float4 q = c[(k * 3 + j) * 3 + i];
        v.x += in ? w * q.x : 0.0f;
        v.y += in ? w * q.y : 0.0f;
        v.z += in ? w * q.z : 0.0f;
      }
  return v;
}

__device__ __forceinline__ void reflect_axis(float &p, float &v, float lo, float hi) {
  float np = p, nv = v;
  if (np < lo) {
    np = lo + (lo - np);
    nv = (-nv) * 0.3f;
  } else if (np > hi) {
    np = hi - (np - hi);
    nv = (-nv) * 0.3f;
  }
  p = fminf(fmaxf(np, lo), hi);
  v = nv;
}

// The alive bitmask mirrors the records' alive flags, so dead slots are
// skipped without touching the pool (32 slots per word read); a kill clears
// both.  Thread 0 also closes the spawns (gen_finish): nothing here reads the
// stack top or the seed.  (Advecting a compacted list of the alive slots
// instead measured slower: the compaction's atomics plus the list's
// indirection cost more than the skipped threads.)
__global__ __launch_bounds__(256) void k_advect(Params P, GState g, DRec *pool, uint32_t *killbits,
                                                uint32_t *alivebits, Walls W, int32_t *topp,
                                                const uint32_t *total, uint32_t *seedp,
                                                int32_t *topsnap) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) gen_finish(topp, total, seedp, topsnap);
  if (i >= P.dmax) return;
  if (!((alivebits[i >> 5] >> (i & 31)) & 1u)) return;
  DRec p = pool[i];
  const float dt = P.dt, gr = P.gravity;
  F3 pos3 = f3(p.pos.x, p.pos.y, p.pos.z);
  bool kill = false;
  // foam and bubbles sample the fluid's velocity (one inlined gather for both)
  const F3 gv = (p.ptype == 1u || p.ptype == 2u) ? sample_grid_velocity(P, g, pos3)
                                                 : f3(0.0f, 0.0f, 0.0f);
  if (p.ptype == 0u) {
    p.vel.y += gr * dt;
    p.pos = make_float4(pos3.x + dt * p.vel.x, pos3.y + dt * p.vel.y, pos3.z + dt * p.vel.z, 0.0f);
  } else if (p.ptype == 1u) {
    p.pos = make_float4(pos3.x + dt * gv.x, pos3.y + dt * gv.y, pos3.z + dt * gv.z, 0.0f);
    p.lifetime -= dt;
    if (p.lifetime <= 0.0f) kill = true;
  } else if (p.ptype == 2u) {
    p.vel.x += (dt * P.drag) * (gv.x - p.vel.x);
    p.vel.y += dt * (P.buoyancy * (-gr) + P.drag * (gv.y - p.vel.y));
    p.vel.z += (dt * P.drag) * (gv.z - p.vel.z);
    p.pos = make_float4(pos3.x + dt * p.vel.x, pos3.y + dt * p.vel.y, pos3.z + dt * p.vel.z, 0.0f);
  }
  if (!kill) {
    for (uint32_t wi = 0; wi < P.nwalls; wi++) {
      const Wall &w = W.w[wi];
      F3 mn = wall_min3(w), mx = wall_max3(w);
      F3 lp = wall_to_local(f3(p.pos.x, p.pos.y, p.pos.z), w);
      if (!inside_open(lp, mn, mx)) continue;
      F3 nl;
      float sdf = wall_sdf_normal(lp, mn, mx, nl);
      F3 nn = wall_to_world_dir(nl, w);
      p.pos = make_float4(p.pos.x - nn.x * sdf, p.pos.y - nn.y * sdf, p.pos.z - nn.z * sdf, 0.0f);
      F3 wvel = wall_vel_at(f3(p.pos.x, p.pos.y, p.pos.z), w);
      F3 vrel = f3(p.vel.x - wvel.x, p.vel.y - wvel.y, p.vel.z - wvel.z);
      float vn = dot3(vrel, nn);
      if (vn < 0.0f) {
        float e = 1.0f + 0.3f;
        p.vel = make_float4(wvel.x + (vrel.x - (e * vn) * nn.x), wvel.y + (vrel.y - (e * vn) * nn.y),
                            wvel.z + (vrel.z - (e * vn) * nn.z), 0.0f);
      }
    }
    if (p.pos.y < P.hard_min) kill = true;
  }
  if (kill) {  // kill_particle (shaders.js:914-918): payload kept, alive = 0
    pool[i].alive = 0u;
    atomicAnd(&alivebits[i >> 5], ~(1u << (i & 31)));
    atomicOr(&killbits[i >> 5], 1u << (i & 31));
    return;
  }
  reflect_axis(p.pos.x, p.vel.x, P.hard_min, P.hard_max_x);
  reflect_axis(p.pos.y, p.vel.y, P.hard_min, P.hard_max_y);
  reflect_axis(p.pos.z, p.vel.z, P.hard_min, P.hard_max_z);
  p.pos.w = 0.0f;
  p.vel.w = 0.0f;
  pool[i] = p;
}

// Alive bitmask from the records (after the host wrote the pool).
__global__ __launch_bounds__(256) void k_alive_bits(const DRec *pool, uint32_t *alivebits,
                                                    uint32_t dmax) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= (dmax + 31u) / 32u) return;
  uint32_t bits = 0;
  for (uint32_t b = 0; b < 32u && w * 32u + b < dmax; b++)
    if (pool[w * 32u + b].alive != 0u) bits |= 1u << b;
  alivebits[w] = bits;
}

// ---------------------------------------------------------------------------
// Exclusive scan of u32 counts (or of the popcounts of bitmask words):
// tile reduce, then the tile scan with the carry-in of the tiles before it.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;

template <bool POPC>
__device__ __forceinline__ uint32_t scan_load(const uint32_t *in, uint32_t n, uint32_t i) {
  if (i >= n) return 0u;
  uint32_t v = in[i];
  return POPC ? (uint32_t)__popc(v) : v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t &excl) {
  __shared__ uint32_t wsum[kScanThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  uint32_t before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; w++) {
    uint32_t s = wsum[w];
    if (w < wid) before += s;
    total += s;
  }
  __syncthreads();
  excl = before + x - v;
  return total;
}

template <bool POPC>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t *in, uint32_t n,
                                                             uint32_t *partials) {
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) s += scan_load<POPC>(in, n, base + i);
  uint32_t excl;
  uint32_t total = block_exclusive_scan(s, excl);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// `clear` (the input array itself, or null) is zeroed behind the scan, which
// leaves a histogram ready for its next use without a separate fill.  Each
// block adds up the tile sums before its own (one per thread for up to 256
// tiles), and block 0 stores the total: no pass over the tile sums alone.
template <bool POPC>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t *in, uint32_t n,
                                                           const uint32_t *partials, uint32_t nb,
                                                           uint32_t *out, uint32_t *clear,
                                                           uint32_t *total) {
  uint32_t before = 0, all = 0;
  for (uint32_t i = threadIdx.x; i < nb; i += kScanThreads) {
    const uint32_t p = partials[i];
    all += p;
    if (i < blockIdx.x) before += p;
  }
  uint32_t unused;
  before = block_exclusive_scan(before, unused);
  all = block_exclusive_scan(all, unused);
  if (blockIdx.x == 0 && threadIdx.x == 0 && total) *total = all;
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = scan_load<POPC>(in, n, base + i);
    s += v[i];
  }
  uint32_t excl;
  block_exclusive_scan(s, excl);
  uint32_t run = before + excl;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    if (base + i < n) {
      out[base + i] = run;
      if (clear) clear[base + i] = 0u;
    }
    run += v[i];
  }
}

// Down pass over a bitmask that writes the index of every set bit at its
// rank: out[*base + rank] for ranks whose position lies in [0, cap) (base
// null: 0), and clears the bitmask behind it.  Block 0 stores the number of
// set bits in *total (if set) and *base + that number in *top (if set).
// Spawners are listed for the apply pass this way (ranks by id), and the
// killed slots are pushed onto the free list above the top (in slot order).
__global__ __launch_bounds__(kScanThreads) void k_scan_emit(uint32_t *bits, uint32_t n,
                                                           const uint32_t *partials, uint32_t nb,
                                                           uint32_t *out, const int32_t *base,
                                                           uint32_t cap, uint32_t *total,
                                                           int32_t *top) {
  uint32_t before = 0, all = 0;
  for (uint32_t i = threadIdx.x; i < nb; i += kScanThreads) {
    const uint32_t p = partials[i];
    all += p;
    if (i < blockIdx.x) before += p;
  }
  uint32_t unused;
  before = block_exclusive_scan(before, unused);
  all = block_exclusive_scan(all, unused);
  const int32_t b0 = base ? *base : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (total) *total = all;
    if (top) *top = b0 + (int32_t)all;
  }
  const uint32_t wbase = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = wbase + i < n ? bits[wbase + i] : 0u;
    sum += (uint32_t)__popc(v[i]);
  }
  uint32_t excl;
  block_exclusive_scan(sum, excl);
  int64_t pos = (int64_t)b0 + before + excl;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    uint32_t word = v[i];
    if (!word) continue;
    bits[wbase + i] = 0u;
    while (word) {
      const int bit = __ffs(word) - 1;
      word &= word - 1;
      if (pos >= 0 && pos < (int64_t)cap) out[pos] = (wbase + i) * 32u + (uint32_t)bit;
      pos++;
    }
  }
}

// ---------------------------------------------------------------------------
// Export to the reference layouts (reset order), for reads.
// The renderer layouts of the particles: into reset-order rows (RESET) or
// into the storage order (contiguous reads and writes).
template <bool RESET>
__global__ __launch_bounds__(256) void k_export_particles(PState S, uint32_t n, float4 *pos4,
                                                          float4 *vel4, float *C9, float *dens) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t id = RESET ? S.idr(s) : s;
  if (pos4) pos4[id] = make_float4(S.at(PX, s), S.at(PY, s), S.at(PZ, s), 0.0f);
  if (vel4) vel4[id] = make_float4(S.at(VX, s), S.at(VY, s), S.at(VZ, s), 0.0f);
  if (C9) {
#pragma unroll
    for (int k = 0; k < 9; k++) C9[(size_t)id * 9 + k] = S.at(C00 + k, s);
  }
  if (dens) dens[id] = S.density[s];
}

__global__ __launch_bounds__(256) void k_export_grid(GState g, uint32_t G, int32_t *m,
                                                     int32_t *mv3, float *mf, float *v3) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G) return;
  if (m) m[c] = g.m[c];
  if (mv3) {
    mv3[3 * (size_t)c] = g.mx[c];
    mv3[3 * (size_t)c + 1] = g.my[c];
    mv3[3 * (size_t)c + 2] = g.mz[c];
  }
  float4 v = g.v[c];
  if (mf) mf[c] = v.w;
  if (v3) {
    v3[3 * (size_t)c] = v.x;
    v3[3 * (size_t)c + 1] = v.y;
    v3[3 * (size_t)c + 2] = v.z;
  }
}

// Scatter host-order particles into slots (identity order) at reset.
__global__ __launch_bounds__(256) void k_import_particles(PState S, uint32_t n, const float *pos3,
                                                          const float *vel3, const float *C9,
                                                          const float *mass, uint32_t first) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t q = first + p;
  S.rec(q) = make_float4(pos3[3 * (size_t)p], pos3[3 * (size_t)p + 1], pos3[3 * (size_t)p + 2],
                         __uint_as_float(q));
  S.at(VX, q) = vel3 ? vel3[3 * (size_t)p] : 0.0f;
  S.at(VY, q) = vel3 ? vel3[3 * (size_t)p + 1] : 0.0f;
  S.at(VZ, q) = vel3 ? vel3[3 * (size_t)p + 2] : 0.0f;
#pragma unroll
  for (int k = 0; k < 9; k++) S.at(C00 + k, q) = C9 ? C9[(size_t)p * 9 + k] : 0.0f;
  S.at(PM, q) = mass ? mass[p] : 1.0f;
  S.density[q] = 0.0f;

}

// ---------------------------------------------------------------------------
// Host launchers.
static inline uint32_t blocks(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

void launch_clear_grid(const GState &g, uint32_t G, hipStream_t st) {
  if (G) k_clear_grid<<<blocks(G, 256), 256, 0, st>>>(g, G);
}
void launch_p2g_mass_global(const Params &P, const PState &S, const GState &g, hipStream_t st) {
  if (!P.n) return;
  if (P.unit_mass) k_p2g_mass_global<true><<<blocks(P.n, 256), 256, 0, st>>>(P, S, g);
  else k_p2g_mass_global<false><<<blocks(P.n, 256), 256, 0, st>>>(P, S, g);
}
void launch_p2g_mom_global(const Params &P, const PState &S, const GState &g, hipStream_t st) {
  if (!P.n) return;
  if (P.unit_mass) k_p2g_mom_global<true><<<blocks(P.n, 256), 256, 0, st>>>(P, S, g);
  else k_p2g_mom_global<false><<<blocks(P.n, 256), 256, 0, st>>>(P, S, g);
}
void launch_update_dense(const Params &P, const GState &g, const Walls &W, bool clear,
                         hipStream_t st) {
  uint32_t G = P.gx * P.gy * P.gz;
  k_update_dense<<<blocks(G, 256), 256, 0, st>>>(P, g, W, clear ? 1 : 0);
}
void launch_g2p_global(const Params &P, const PState &S, const GState &g, const Walls &W,
                       hipStream_t st) {
  if (!P.n) return;
  k_g2p_global<<<blocks(P.n, 256), 256, 0, st>>>(P, S, g, W);
}
void launch_g2p_relayout(const Params &P, const PState &S, const PState &D, const GState &g,
                         const Walls &W, const uint32_t *src_of, hipStream_t st) {
  if (!P.n) return;
  k_g2p_relayout<<<blocks(P.n, 256), 256, 0, st>>>(P, S, D, g, W, src_of);
}

uint32_t scan_scratch_words(uint32_t n) { return blocks(n, kScanTile) + 1; }

void launch_scan(const uint32_t *in, uint32_t n, bool popc, uint32_t *out, uint32_t *scratch,
                 uint32_t *total, hipStream_t st, uint32_t *clear) {
  uint32_t nb = blocks(n, kScanTile);
  if (nb == 0) nb = 1;
  if (popc) k_scan_reduce<true><<<nb, kScanThreads, 0, st>>>(in, n, scratch);
  else k_scan_reduce<false><<<nb, kScanThreads, 0, st>>>(in, n, scratch);
  if (popc) k_scan_down<true><<<nb, kScanThreads, 0, st>>>(in, n, scratch, nb, out, clear, total);
  else k_scan_down<false><<<nb, kScanThreads, 0, st>>>(in, n, scratch, nb, out, clear, total);
}

void launch_alive_bits(const DiffuseBufs &D, uint32_t dmax, hipStream_t st) {
  if (dmax) k_alive_bits<<<blocks((dmax + 31u) / 32u, 256), 256, 0, st>>>(D.pool, D.alivebits, dmax);
}

// The diffuse step: decide (spawners by id; or done by G2P) -> scan that
// lists them in id order -> apply (one thread per spawner) -> advect over
// the alive slots (closes the spawns) -> scan that pushes the killed slots.
void launch_diffuse(const Params &P, const PState &S, const GState &g, const Walls &W,
                    const DiffuseBufs &D, hipStream_t st, bool decided) {
  if (!P.dmax) return;
  int32_t *topsnap = (int32_t *)(D.total + 1);
  if (P.n) {
    const uint32_t nwg = blocks(P.n, 32);
    if (!decided)
      k_gen_decide<<<blocks(P.n, 256), 256, 0, st>>>(P, S, g, D.seed, D.genbits, D.genval);
    const uint32_t nb = blocks(nwg, kScanTile);
    k_scan_reduce<true><<<nb, kScanThreads, 0, st>>>(D.genbits, nwg, D.scratch);
    k_scan_emit<<<nb, kScanThreads, 0, st>>>(D.genbits, nwg, D.scratch, nb, D.list, nullptr,
                                             P.dmax, D.total, nullptr);
    const uint32_t ab = blocks(P.dmax, 256);
    k_gen_apply<<<ab < 1024 ? ab : 1024, 256, 0, st>>>(P, S, D.seed, D.list, D.total, D.genval,
                                                       D.free_list, D.top, D.pool, D.alivebits);
  } else {
    k_zero_total<<<1, 1, 0, st>>>(D.total);
  }
  k_advect<<<blocks(P.dmax, 256), 256, 0, st>>>(P, g, D.pool, D.killbits, D.alivebits, W, D.top,
                                               D.total, D.seed, topsnap);
  const uint32_t nwk = blocks(P.dmax, 32), kb = blocks(nwk, kScanTile);
  k_scan_reduce<true><<<kb, kScanThreads, 0, st>>>(D.killbits, nwk, D.scratch);
  k_scan_emit<<<kb, kScanThreads, 0, st>>>(D.killbits, nwk, D.scratch, kb, D.free_list, topsnap,
                                           P.dmax, nullptr, D.top);
}

void launch_export_particles(const PState &S, uint32_t n, float4 *pos4, float4 *vel4, float *C9,
                             float *dens, hipStream_t st, bool reset_order) {
  if (!n) return;
  if (reset_order) k_export_particles<true><<<blocks(n, 256), 256, 0, st>>>(S, n, pos4, vel4, C9, dens);
  else k_export_particles<false><<<blocks(n, 256), 256, 0, st>>>(S, n, pos4, vel4, C9, dens);
}

__global__ __launch_bounds__(256) void k_copy16(const uint4 *src, uint4 *dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

void launch_copy16(const void *src, void *dst, size_t bytes, hipStream_t st) {
  const size_t n = bytes / 16;
  if (!n) return;
  const size_t want = (n + 255) / 256;
  k_copy16<<<(uint32_t)(want < 4096 ? want : 4096), 256, 0, st>>>((const uint4 *)src, (uint4 *)dst, n);
}
void launch_export_grid(const GState &g, uint32_t G, int32_t *m, int32_t *mv3, float *mf,
                        float *v3, hipStream_t st) {
  if (G) k_export_grid<<<blocks(G, 256), 256, 0, st>>>(g, G, m, mv3, mf, v3);
}
void launch_import_particles(const PState &S, uint32_t n, const float *pos3, const float *vel3,
                             const float *C9, const float *mass, hipStream_t st, uint32_t first) {
  if (n) k_import_particles<<<blocks(n, 256), 256, 0, st>>>(S, n, pos3, vel3, C9, mass, first);
}

}  // namespace mpm
