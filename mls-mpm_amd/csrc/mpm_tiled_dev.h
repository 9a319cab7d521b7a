// mpm_tiled_dev.h -- device bodies of the brick-ordered, LDS-staged P2G /
// G2P and of the re-sort plumbing (internal).  Each body takes its block
// (work item, listed brick) or thread (brick, particle) index explicitly, so
// the single-scene kernels (mpm_tiled.hip) and the batched persistent kernels
// (mpm_batch.hip) run the same code.
#pragma once
#include "mpm_kernels.h"
#include "mpm_ops.h"
#include "mpm_tiled.h"

namespace mpm {

constexpr int kTB = 256;    // row width of an item (slots per row)
constexpr int kMaxPpt = 16; // largest particles-per-thread of a work item
constexpr int kCore = 10;   // brick + 1-cell halo per axis
constexpr int kCoreV = kCore * kCore * kCore;

// Ghost core layout, kCoreV cells per item: the cells x = 1..8 of every
// (z, y) row first, 8 contiguous per row, rows z-major (800 cells); then the
// x = 0 plane and the x = 9 plane, (z, y)-major (100 cells each).  Every
// region a neighbouring brick's resolve reads is then made of runs of at
// least 8 contiguous cells; in x-fastest order its x faces would be one cell
// every 10.
constexpr int kGhostRows = kCore * kCore * (kCore - 2);
__device__ __forceinline__ int ghost_off(int cx, int cy, int cz) {
  const int yz = cz * kCore + cy;
  return cx == 0 ? kGhostRows + yz
                 : (cx == kCore - 1 ? kGhostRows + kCore * kCore + yz : yz * (kCore - 2) + cx - 1);
}
// The inverse: ghost index d to its core cell.
__device__ __forceinline__ void ghost_cell(int d, int &cx, int &cy, int &cz) {
  int yz;
  if (d < kGhostRows) {
    yz = d >> 3;
    cx = (d & 7) + 1;
  } else {
    const int e = d - kGhostRows;
    cx = e < kCore * kCore ? 0 : kCore - 1;
    yz = e < kCore * kCore ? e : e - kCore * kCore;
  }
  cz = yz / kCore;
  cy = yz - cz * kCore;
}
// Resolve blocks: cells per thread and threads per block (one 8^3 brick).
constexpr int kResolveCells = 2;
constexpr int kResolveThreads = 512 / kResolveCells;
constexpr uint32_t kFar = 31u;  // rank tag: particle left the 27-brick neighbourhood

// Brick of a cell, clamped into the grid.
__device__ __forceinline__ void cell_brick(const Params &P, float x, float y, float z, int &bx,
                                           int &by, int &bz) {
  bx = min(max((int)floorf(x), 0), (int)P.gx - 1) >> 3;
  by = min(max((int)floorf(y), 0), (int)P.gy - 1) >> 3;
  bz = min(max((int)floorf(z), 0), (int)P.gz - 1) >> 3;
}

__device__ __forceinline__ void brick_coords(const BrickWork &bw, uint32_t b, int &bx, int &by,
                                             int &bz) {
  bx = (int)(b % bw.nbx);
  by = (int)((b / bw.nbx) % bw.nby);
  bz = (int)(b / (bw.nbx * bw.nby));
}

__device__ __forceinline__ uint32_t brick_index(const BrickWork &bw, int bx, int by, int bz) {
  return ((uint32_t)bz * bw.nby + (uint32_t)by) * bw.nbx + (uint32_t)bx;
}

// Global index of window cell c (edge E, origin o), or -1 outside the grid.
template <int E>
__device__ __forceinline__ int window_cell(const Params &P, int ox, int oy, int oz, int c) {
  int wz = c / (E * E), rem = c - wz * (E * E);
  int wy = rem / E, wx = rem - wy * E;
  int x = ox + wx, y = oy + wy, z = oz + wz;
  if (x < 0 || y < 0 || z < 0 || x >= (int)P.gx || y >= (int)P.gy || z >= (int)P.gz) return -1;
  return (z * (int)P.gy + y) * (int)P.gx + x;
}

// window_cell without branches: gi is the cell's global index inside the
// grid, else 0 (a valid address); returns whether the cell is inside.
template <int E>
__device__ __forceinline__ bool window_cell_sel(const Params &P, int ox, int oy, int oz, int c,
                                                int &gi) {
  const int wz = c / (E * E), rem = c - wz * (E * E);
  const int wy = rem / E, wx = rem - wy * E;
  const int x = ox + wx, y = oy + wy, z = oz + wz;
  const bool in = ((unsigned)x < P.gx) & ((unsigned)y < P.gy) & ((unsigned)z < P.gz);
  // a mask, not a select: a select here is compiled into a branch
  gi = ((z * (int)P.gy + y) * (int)P.gx + x) & -(int)in;
  return in;
}

__device__ __forceinline__ void flag_cell_brick(const BrickWork &bw, int x, int y, int z) {
  bw.ofl[brick_index(bw, x >> 3, y >> 3, z >> 3)] = 1;
}

// A particle whose stencil leaves the window core contributes to margin
// cells (or, past the margin, straight to the overflow grid): flag the
// bricks of its stencil corners so the resolve folds the overflow in.
__device__ __forceinline__ void flag_stencil(const Params &P, const BrickWork &bw, const PW &pw) {
#pragma unroll
  for (int c = 0; c < 8; c++) {
    int x = min(max(pw.b[0] + ((c & 1) ? 2 : 0), 0), (int)P.gx - 1);
    int y = min(max(pw.b[1] + ((c & 2) ? 2 : 0), 0), (int)P.gy - 1);
    int z = min(max(pw.b[2] + ((c & 4) ? 2 : 0), 0), (int)P.gz - 1);
    flag_cell_brick(bw, x, y, z);
  }
}

// Window position of a particle's base cell and whether its stencil lies in
// the window / in the window core.
struct WinPos {
  int lb;
  bool inside, in_core;
};
template <int M>
__device__ __forceinline__ WinPos win_pos(const PW &pw, int ox, int oy, int oz) {
  constexpr int E = 10 + 2 * M;
  int lx = pw.b[0] - ox, ly = pw.b[1] - oy, lz = pw.b[2] - oz;
  WinPos w;
  w.inside = lx >= 0 && ly >= 0 && lz >= 0 && lx + 2 < E && ly + 2 < E && lz + 2 < E;
  w.in_core = lx >= M && ly >= M && lz >= M && lx + 2 < M + kCore && ly + 2 < M + kCore &&
              lz + 2 < M + kCore;
  w.lb = (lz * E + ly) * E + lx;
  return w;
}

// floor(r / d) for r < 2^13 and 1 <= d <= 16, or r < 2^7 and 1 <= d <= 64.
// (r + 1/2) / d lies at least 1/(2d) from every integer (1/32, or 1/128),
// and the float product is within 2^-21 * r of it (2^-8, or 2^-14), so the
// truncation is exact (the reciprocal may be off by an ulp).
__device__ __forceinline__ uint32_t div_small(uint32_t r, uint32_t d) {
  return (uint32_t)(((float)r + 0.5f) * __builtin_amdgcn_rcpf((float)d));
}

// Offset in a block of n consecutive ranks (n <= 64) of the block's p-th
// particle, taken column by column: the offsets in order of (offset mod K,
// offset), a bijection of [0, n).  Offsets K apart are the slots of
// neighbouring threads in one row of an item (item_slot), so consecutive
// particles of the block land in neighbouring threads of one row.
__device__ __forceinline__ uint32_t rank_in_block(uint32_t p, uint32_t n, uint32_t K) {
  const uint32_t q = div_small(n, K), rm = n - q * K, big = rm * (q + 1u);
  uint32_t j, m;
  if (p < big) {
    j = div_small(p, q + 1u);
    m = p - j * (q + 1u);
  } else {  // q >= 1 here (p < n)
    const uint32_t p2 = p - big, u = div_small(p2, q);
    j = rm + u;
    m = p2 - u * q;
  }
  return j + K * m;
}

// Slot of the particle of rank r among the c particles of an item: ranks are
// stored column-major over rows of 256 slots.  With K = ceil(c/256), the
// first L = c - 256(K-1) columns hold K ranks, the others K-1.
__device__ __forceinline__ uint32_t item_slot(uint32_t r, uint32_t c) {
  const uint32_t K = (c + 255u) >> 8, L = c - 256u * (K - 1u);
  uint32_t t, k;
  if (r < L * K) {
    t = div_small(r, K);
    k = r - t * K;
  } else {  // K >= 2 here
    const uint32_t r2 = r - L * K, q = div_small(r2, K - 1u);
    t = L + q;
    k = r2 - q * (K - 1u);
  }
  return k * 256u + t;
}

// The thread index as an opaque value.  Unrolled staging loops derive their
// window offsets from it; in the batched kernels, which run the item bodies
// in a loop over work units, those offsets would otherwise be hoisted out of
// the loop and held in registers across every item.
__device__ __forceinline__ int opaque_lane(int t) {
  asm volatile("" : "+v"(t));
  return t;
}

// Work item blk of a launch and whether it exists.  The item buffers hold an
// entry for every block of a launch (max_items), so the item is loaded
// together with the count rather than after it.
__device__ __forceinline__ bool load_item(const uint4 *items, const uint32_t *nitems,
                                          uint32_t blk, uint4 &it) {
  const uint32_t n = *nitems;
  it = items[blk];
  // Both operands evaluated: the compiler keeps the item load ahead of the
  // branch instead of sinking it behind the count's round trip.  A count is
  // never all ones.
  return (blk < n) & (it.z != 0xffffffffu);
}

// ---------------------------------------------------------------------------
// P2G mass over brick work items (WGSL_P2G_MASS, shaders.js:242-274).  With
// ko.counts set, also assigns every particle its brick and rank for the
// re-sort.
template <int M, bool UNIT>
__device__ __forceinline__ void p2g_mass_item(const Params &P, const PState &S,
                                              const BrickWork &bw, const KeyOut &ko, uint32_t blk) {
  constexpr int E = 10 + 2 * M, V = E * E * E;
  __shared__ int32_t lds[V];
  // the 27 neighbour bricks: particle counts, reserved bases, Morton ranks
  __shared__ uint32_t lcnt[27], lbase[27], ldest[27];
  // An item of at most kStageRanks particles keeps its (neighbour, rank)
  // tags here until the bases are known, 16 bits each: neighbour << 11 |
  // rank among the item's particles bound there (< 2048); a particle beyond
  // the 27 bricks takes its final rank in ko.rank at once and the tag kFar.
  // A larger item keeps (neighbour << 26 | rank) tags in ko.rank.  The
  // 16-bit tags keep the block's LDS small enough for 8 blocks per CU.
  constexpr uint32_t kStageRanks = 8u * kTB;
  __shared__ uint16_t lrk[kStageRanks];
  // the re-sort scan's running totals start every step at zero (the brick
  // list kernel after this pass adds to them)
  if (ko.acc && blk == 0 && threadIdx.x < kScanStride) ko.acc[threadIdx.x * kAccStride] = 0u;
  uint4 it;
  if (!load_item(bw.items, bw.nitems, blk, it)) return;
  int ibx, iby, ibz;
  brick_coords(bw, it.w, ibx, iby, ibz);
  const int ox = ibx * 8 - 1 - M, oy = iby * 8 - 1 - M, oz = ibz * 8 - 1 - M;
  const int t = threadIdx.x, nt = blockDim.x;
  const bool stage = it.z <= kStageRanks;
  for (int c = t; c < V; c += nt) lds[c] = 0;
  if (t < 27) {
    lcnt[t] = 0;
    if (ko.counts) {
      const int bx = ibx + t % 3 - 1, by = iby + (t / 3) % 3 - 1, bz = ibz + t / 9 - 1;
      const bool in = bx >= 0 && by >= 0 && bz >= 0 && bx < (int)bw.nbx && by < (int)bw.nby &&
                      bz < (int)bw.nbz;
      ldest[t] = in ? bw.brick_rank[brick_index(bw, bx, by, bz)] : 0u;
    }
  }
  __syncthreads();
  for (uint32_t idx = t; idx < it.z; idx += nt) {
    uint32_t s = it.y + idx;
    const float4 q = S.rec(s);
    float px = q.x, py = q.y, pz = q.z;
    PW pw;
    particle_weights(px, py, pz, pw);
    if (ko.counts) {
      int bx, by, bz;
      cell_brick(P, px, py, pz, bx, by, bz);
      int dx = bx - ibx + 1, dy = by - iby + 1, dz = bz - ibz + 1;
      const bool near = dx >= 0 && dx < 3 && dy >= 0 && dy < 3 && dz >= 0 && dz < 3;
      const uint32_t nb = near ? (uint32_t)((dz * 3 + dy) * 3 + dx) : kFar;
      // a near particle's brick is one of the 27 (inside the grid: cell_brick
      // clamps), whose ranks are in LDS
      const uint32_t dest = near ? ldest[nb] : bw.brick_rank[brick_index(bw, bx, by, bz)];
      ko.keys[s] = dest;
      // Most particles stay in their own brick (nb 13): those of a
      // wavefront take consecutive ranks with one LDS atomic.
      const uint64_t own = __ballot(nb == 13u);
      uint32_t r;
      if (nb == 13u) {
        const uint32_t lane = __lane_id();
        const int leader = __ffsll((long long)own) - 1;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&lcnt[13], (uint32_t)__popcll(own));
        base = __shfl(base, leader, 64);
        const uint32_t p = (uint32_t)__popcll(own & ((1ull << lane) - 1ull));
        r = base + (ko.cols > 1u ? rank_in_block(p, (uint32_t)__popcll(own), ko.cols) : p);
      } else if (near) {
        r = atomicAdd(&lcnt[nb], 1u);
      } else {
        r = atomicAdd(&ko.counts[dest], 1u);
      }
      if (stage) {
        lrk[idx] = (uint16_t)(near ? (nb << 11) | r : (kFar << 11));
        if (!near) ko.rank[s] = r;
      } else {
        ko.rank[s] = (nb << 26) | r;
      }
    }
    float pm = UNIT ? 1.0f : S.at(PM, s);
    WinPos wp = win_pos<M>(pw, ox, oy, oz);
    if (!wp.in_core) flag_stencil(P, bw, pw);
    if (wp.inside) {
#pragma unroll
      for (int kk = 0; kk < 3; kk++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          float wxy[3];
#pragma unroll
          for (int i = 0; i < 3; i++) wxy[i] = pw.w[0][i] * pw.w[1][j];
#pragma unroll
          for (int i = 0; i < 3; i++) {
            float w = wxy[i] * pw.w[2][kk];
            atomicAdd(&lds[wp.lb + (kk * E + j) * E + i], encode(UNIT ? w : w * pm));
          }
        }
    } else {  // beyond the window: overflow atomics
      for (int kk = 0; kk < 3; kk++)
        for (int j = 0; j < 3; j++)
          for (int i = 0; i < 3; i++) {
            int gi = cell_idx_checked(pw.b[0] + i, pw.b[1] + j, pw.b[2] + kk, P);
            if (gi < 0) continue;
            float w = (pw.w[0][i] * pw.w[1][j]) * pw.w[2][kk];
            atomicAdd(&bw.om[gi], encode(UNIT ? w : w * pm));
          }
    }
  }
  __syncthreads();
  if (ko.counts) {
    if (t < 27) {
      uint32_t c = lcnt[t], base = 0;
      if (c) base = atomicAdd(&ko.counts[ldest[t]], c);
      lbase[t] = base;
    }
    __syncthreads();
    for (uint32_t idx = t; idx < it.z; idx += nt) {
      uint32_t s = it.y + idx;
      if (stage) {
        const uint32_t v = lrk[idx], nb = v >> 11;
        if (nb != kFar) ko.rank[s] = lbase[nb] + (v & 0x7ffu);
      } else {
        const uint32_t r = ko.rank[s], nb = r >> 26;
        ko.rank[s] = (nb != kFar ? lbase[nb] : 0u) + (r & 0x3ffffffu);
      }
    }
  }
  // core to the ghost buffer, non-zero margin cells to the overflow grid
  int32_t *gh = bw.gm + (size_t)it.x * kCoreV;  // the item's ghost core
  for (int c = t; c < kCoreV; c += nt) {
    int cx, cy, cz;
    ghost_cell(c, cx, cy, cz);
    gh[c] = lds[((cz + M) * E + (cy + M)) * E + (cx + M)];
  }
  if (M > 0) {
    for (int c = t; c < V; c += nt) {
      int32_t v = lds[c];
      if (v == 0) continue;
      int wz = c / (E * E), wy = (c / E) % E, wx = c % E;
      bool core = wx >= M && wy >= M && wz >= M && wx < M + kCore && wy < M + kCore &&
                  wz < M + kCore;
      if (core) continue;
      int gi = window_cell<E>(P, ox, oy, oz, c);
      if (gi >= 0) atomicAdd(&bw.om[gi], v);
    }
  }
}

// ---------------------------------------------------------------------------
// P2G momentum over brick work items (WGSL_P2G_MOM, shaders.js:286-370): the
// decoded mass window and the momentum window live in LDS.  SRC: the pass
// also writes the re-sort's inverse permutation, so.src_of[new slot] = old
// slot, from each particle's new brick and rank (the scan has run).
template <int M, bool UNIT, bool SRC = false>
__device__ __forceinline__ void p2g_mom_item(const Params &P, const PState &S, const GState &g,
                                             const BrickWork &bw, uint32_t blk,
                                             const SrcOut &so = SrcOut{}) {
  constexpr int E = 10 + 2 * M, V = E * E * E;
  __shared__ float lm[V];
  __shared__ int32_t lv[3 * V];
  uint4 it;
  if (!load_item(bw.items, bw.nitems, blk, it)) return;
  int ibx, iby, ibz;
  brick_coords(bw, it.w, ibx, iby, ibz);
  const int ox = ibx * 8 - 1 - M, oy = iby * 8 - 1 - M, oz = ibz * 8 - 1 - M;
  const int t = threadIdx.x, nt = blockDim.x;
  // The decoded mass window: every load of a thread issued before the first
  // decode (blocks of kTB threads), so the window costs one memory latency.
  // The loads are unconditional (cells outside the grid read cell 0 and are
  // then zeroed): a load under a branch is waited for at the branch's end.
  constexpr int kPer = (V + kTB - 1) / kTB;
  int32_t mw[kPer];
  bool in[kPer];
  const int tw = opaque_lane(t);
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int c = tw + u * kTB;
    int gi;
    in[u] = window_cell_sel<E>(P, ox, oy, oz, c < V ? c : V - 1, gi) && c < V;
    mw[u] = g.m[gi];
  }
  // No decode is scheduled among the loads (that would wait for each one).
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int u = 0; u < kPer; u++) mw[u] = in[u] ? mw[u] : 0;
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int c = tw + u * kTB;
    if (c < V) {
      lm[c] = decode(mw[u]);
      lv[3 * c] = 0;
      lv[3 * c + 1] = 0;
      lv[3 * c + 2] = 0;
    }
  }
  __syncthreads();
  const uint32_t gx = P.gx, slice = P.gx * P.gy;
  // Software pipeline: the next particle's inputs load while this one computes.
  // The prefetch index past the item's end is clamped to its last particle
  // (in bounds, unconditional).
  uint32_t idx = t;
  MomIn cur;
  uint32_t ck = 0, cr = 0;  // SRC: the particle's new brick (Morton rank) and rank in it
  if (idx < it.z) {
    load_mom_in<UNIT>(S, it.y + idx, cur);
    if (SRC) {
      ck = so.keys[it.y + idx];
      cr = so.rank[it.y + idx];
    }
  }
  while (idx < it.z) {
    const uint32_t s = it.y + idx, nidx = idx + nt;
    const uint32_t sn = it.y + (nidx < it.z ? nidx : it.z - 1);
    MomIn nxt;
    load_mom_in<UNIT>(S, sn, nxt);
    uint32_t nk = 0, nr = 0;
    uint2 span = make_uint2(0u, 0u);  // the new brick's [first, end) slots
    if (SRC) {
      nk = so.keys[sn];
      nr = so.rank[sn];
      span = so.spans[ck];
    }
    PW pw;
    particle_weights(cur.px, cur.py, cur.pz, pw);
    WinPos wp = win_pos<M>(pw, ox, oy, oz);
    float density;
    if (wp.inside) {
      const int lb = wp.lb;
      density = p2g_mom_body<UNIT>(
          P, cur, pw, [&](int kk, int j, int i) { return lm[lb + (kk * E + j) * E + i]; },
          [&](int kk, int j, int i, int32_t ex, int32_t ey, int32_t ez) {
            int c = lb + (kk * E + j) * E + i;
            atomicAdd(&lv[3 * c], ex);
            atomicAdd(&lv[3 * c + 1], ey);
            atomicAdd(&lv[3 * c + 2], ez);
          });
    } else {  // beyond the window: global gather, overflow atomics
      uint32_t base = (uint32_t)((pw.b[2] * (int)P.gy + pw.b[1]) * (int)gx + pw.b[0]);
      density = p2g_mom_body<UNIT>(
          P, cur, pw,
          [&](int kk, int j, int i) { return decode(g.m[base + kk * slice + j * gx + i]); },
          [&](int kk, int j, int i, int32_t ex, int32_t ey, int32_t ez) {
            uint32_t c = base + kk * slice + j * gx + i;
            atomicAdd(&bw.ox[c], ex);
            atomicAdd(&bw.oy[c], ey);
            atomicAdd(&bw.oz[c], ez);
          });
    }
    if (SRC) {
      const uint32_t ch = 1u << so.chs, istart = span.x + (cr & ~(ch - 1u));
      const uint32_t d = istart + item_slot(cr & (ch - 1u), min(span.y - istart, ch));
      so.src_of[d] = s;
      if (so.dens_w) so.dens_w[d] = density;
      else S.dens(s) = density;
    } else {
      S.dens(s) = density;
    }
    cur = nxt;
    ck = nk;
    cr = nr;
    idx = nidx;
  }
  __syncthreads();
  const size_t gb = (size_t)it.x * kCoreV;  // the item's ghost core
  for (int c = t; c < kCoreV; c += nt) {
    int cx, cy, cz;
    ghost_cell(c, cx, cy, cz);
    int w = ((cz + M) * E + (cy + M)) * E + (cx + M);
    bw.gx[gb + c] = lv[3 * w];
    bw.gy[gb + c] = lv[3 * w + 1];
    bw.gz[gb + c] = lv[3 * w + 2];
  }
  if (M > 0) {
    for (int c = t; c < V; c += nt) {
      int32_t ex = lv[3 * c], ey = lv[3 * c + 1], ez = lv[3 * c + 2];
      if ((ex | ey | ez) == 0) continue;
      int wz = c / (E * E), wy = (c / E) % E, wx = c % E;
      bool core = wx >= M && wy >= M && wz >= M && wx < M + kCore && wy < M + kCore &&
                  wz < M + kCore;
      if (core) continue;
      int gi = window_cell<E>(P, ox, oy, oz, c);
      if (gi < 0) continue;
      if (ex) atomicAdd(&bw.ox[gi], ex);
      if (ey) atomicAdd(&bw.oy[gi], ey);
      if (ez) atomicAdd(&bw.oz[gi], ez);
    }
  }
}

// ---------------------------------------------------------------------------
// Resolve.  list holds the bricks whose cells can be non-zero this step or
// were non-zero last step: bricks next to an occupied brick, flagged bricks
// and bricks active last step.  One block (256 threads, two cells each) per
// listed brick, persistent over the list.
__device__ __forceinline__ bool brick_needed(const BrickWork &bw, uint32_t b) {
  if (b >= bw.nbricks) return false;
  int bx, by, bz;
  brick_coords(bw, b, bx, by, bz);
  // All 27 neighbour flags load at once (indexes clamped into the grid, the
  // outside ones masked), not one after another.
  uint32_t need = (uint32_t)(bw.act[b] & 1) | (uint32_t)bw.ofl[b];
#pragma unroll
  for (int n = 0; n < 27; n++) {
    const int x = bx + n % 3 - 1, y = by + (n / 3) % 3 - 1, z = bz + n / 9 - 1;
    const bool in = x >= 0 && y >= 0 && z >= 0 && x < (int)bw.nbx && y < (int)bw.nby &&
                    z < (int)bw.nbz;
    const int cx = min(max(x, 0), (int)bw.nbx - 1), cy = min(max(y, 0), (int)bw.nby - 1),
              cz = min(max(z, 0), (int)bw.nbz - 1);
    need |= in ? (uint32_t)bw.occ[brick_index(bw, cx, cy, cz)] : 0u;
  }
  return need != 0;
}


struct NbrItems {
  uint32_t first[27], count[27];
};

// (first item, item count) of the 27 bricks around b, by brick index (zero
// outside the grid); returns whether any of them has items.
__device__ __forceinline__ bool load_neighbours(const BrickWork &bw, uint32_t b, NbrItems &nb,
                                                int t) {
  __shared__ int any;
  if (t == 0) any = 0;
  __syncthreads();
  if (t < 27) {
    int bx, by, bz;
    brick_coords(bw, b, bx, by, bz);
    bx += t % 3 - 1;
    by += (t / 3) % 3 - 1;
    bz += t / 9 - 1;
    uint32_t cnt = 0, first = 0;
    if (bx >= 0 && by >= 0 && bz >= 0 && bx < (int)bw.nbx && by < (int)bw.nby &&
        bz < (int)bw.nbz) {
      const uint32_t nbi = brick_index(bw, bx, by, bz);
      cnt = bw.count[nbi];
      first = bw.first[nbi];
    }
    nb.first[t] = first;
    nb.count[t] = cnt;
    if (cnt) atomicOr(&any, 1);
  }
  __syncthreads();
  return any != 0;
}

// The ghost arrays a resolve sums (mass: 1, momentum: 3).
template <int R>
struct GhostSrc {
  const int32_t *p[R];
};

// Ghost sums of every cell of the brick into acc[R][512] (LDS), by the whole
// block.  The contributions are the brick's own items (all 512 cells) and,
// for each neighbour brick n, its items' halo cells that lie in this brick
// (a face of 64 cells, an edge of 8 or a corner).  They are numbered flat
// (neighbour, item, cell) through a per-neighbour prefix, and every thread
// takes four entries at a time: decode, issue all their loads, then add
// into LDS, so a block keeps many independent loads in flight instead of
// walking the neighbours one after another.  Integer sums: the order is
// free.  Block-uniform; ends with a barrier.
template <int R>
__device__ __forceinline__ void gather_cells(const NbrItems &nb, const GhostSrc<R> &gs,
                                             int32_t (*acc)[512], int t, int nt) {
  __shared__ uint32_t pre[28];
  if (t < 64) {
    // work of neighbour t: items x cells of its region in this brick
    uint32_t w = 0;
    if (t < 27) {
      const int m = (t % 3 != 1) + ((t / 3) % 3 != 1) + (t / 9 != 1);
      w = nb.count[t] << (9 - 3 * m);
    }
    // inclusive scan over the wavefront, stored exclusive
    uint32_t x = w;
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) {
      const uint32_t y = __shfl_up(x, off, 64);
      if (t >= off) x += y;
    }
    if (t < 28) pre[t] = x - w;
  }
  for (int c = t; c < R * 512; c += nt) acc[c >> 9][c & 511] = 0;
  __syncthreads();
  const uint32_t total = pre[27];
  for (uint32_t f0 = t; f0 < total; f0 += 4u * nt) {
    size_t src[4];
    int tgt[4];
    bool ok[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint32_t f = f0 + u * nt;
      ok[u] = f < total;
      const uint32_t ff = ok[u] ? f : total - 1u;
      int lo = 0, hi = 27;  // pre[lo] <= ff < pre[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= ff) lo = mid;
        else hi = mid;
      }
      const int n = lo, d[3] = {n % 3, (n / 3) % 3, n / 9};
      const int m = (d[0] != 1) + (d[1] != 1) + (d[2] != 1), sh = 9 - 3 * m;
      const uint32_t r = ff - pre[n], k = r >> sh, q = r & ((1u << sh) - 1u);
      int l[3], cc[3], j = 0;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        // free axis: coordinate from q; neighbour below: our low face, its
        // core's high halo (9); neighbour above: our high face, its core's 0
        if (d[a] == 1) {
          l[a] = (int)(q >> (3 * j)) & 7;
          cc[a] = l[a] + 1;
          j++;
        } else {
          l[a] = d[a] == 0 ? 0 : 7;
          cc[a] = d[a] == 0 ? 9 : 0;
        }
      }
      tgt[u] = (l[2] * 8 + l[1]) * 8 + l[0];
      src[u] = (size_t)(nb.first[n] + k) * kCoreV + (size_t)ghost_off(cc[0], cc[1], cc[2]);
    }
    int32_t v[4][R];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int r = 0; r < R; r++) v[u][r] = gs.p[r][src[u]];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (ok[u])
#pragma unroll
        for (int r = 0; r < R; r++) atomicAdd(&acc[r][tgt[u]], v[u][r]);
  }
  __syncthreads();
}

// Listed brick b, one block of 512 / C threads, C cells per thread;
// block-uniform.
template <int C>
__device__ __forceinline__ void resolve_mass_entry(const Params &P, const GState &g,
                                                   const BrickWork &bw, uint32_t b) {
  constexpr int NT = 512 / C;
  __shared__ NbrItems nb;
  const int t = threadIdx.x;
  const bool ofl = bw.ofl[b] != 0;
  bool any = load_neighbours(bw, b, nb, t);
  int bx, by, bz;
  brick_coords(bw, b, bx, by, bz);
  __shared__ int32_t acc[1][512];
  if (any) {
    const GhostSrc<1> gs = {{bw.gm}};
    gather_cells<1>(nb, gs, acc, t, NT);
  }
#pragma unroll
  for (int j = 0; j < C; j++) {
    const int c = t + j * NT;
    const int x = bx * 8 + (c & 7), y = by * 8 + ((c >> 3) & 7), z = bz * 8 + (c >> 6);
    if (x < (int)P.gx && y < (int)P.gy && z < (int)P.gz) {
      const uint32_t cell = ((uint32_t)z * P.gy + (uint32_t)y) * P.gx + (uint32_t)x;
      uint32_t sum = any ? (uint32_t)acc[0][c] : 0u;
      if (ofl) {
        sum += (uint32_t)bw.om[cell];
        bw.om[cell] = 0;
      }
      g.m[cell] = (int32_t)sum;
    }
  }
  if (t == 0) {
    uint8_t a = bw.act[b];
    bw.act[b] = (uint8_t)(((a & 1) << 1) | ((any || ofl) ? 1 : 0));
  }
  __syncthreads();
}

// Momentum resolve fused with the grid update (WGSL_UPDATE_GRID,
// shaders.js:443-512): (vx, vy, vz, decoded mass) per cell.  Listed brick
// b, one block of 512 / C threads, C cells per thread; block-uniform.
template <int C>
__device__ __forceinline__ void resolve_update_entry(const Params &P, const GState &g,
                                                     const BrickWork &bw, const Walls &W,
                                                     int keep, uint32_t b) {
  constexpr int NT = 512 / C;
  __shared__ NbrItems nb;
  const int t = threadIdx.x;
  const bool ofl = bw.ofl[b] != 0;
  bool any = load_neighbours(bw, b, nb, t);
  int bx, by, bz;
  brick_coords(bw, b, bx, by, bz);
  __shared__ int32_t acc[3][512];
  if (any) {
    const GhostSrc<3> gs = {{bw.gx, bw.gy, bw.gz}};
    gather_cells<3>(nb, gs, acc, t, NT);
  }
#pragma unroll
  for (int j = 0; j < C; j++) {
    const int c = t + j * NT;
    const int x = bx * 8 + (c & 7), y = by * 8 + ((c >> 3) & 7), z = bz * 8 + (c >> 6);
    if (x < (int)P.gx && y < (int)P.gy && z < (int)P.gz) {
      const uint32_t cell = ((uint32_t)z * P.gy + (uint32_t)y) * P.gx + (uint32_t)x;
      uint32_t sx = any ? (uint32_t)acc[0][c] : 0u, sy = any ? (uint32_t)acc[1][c] : 0u,
               sz = any ? (uint32_t)acc[2][c] : 0u;
      if (ofl) {
        sx += (uint32_t)bw.ox[cell];
        sy += (uint32_t)bw.oy[cell];
        sz += (uint32_t)bw.oz[cell];
        bw.ox[cell] = 0;
        bw.oy[cell] = 0;
        bw.oz[cell] = 0;
      }
      g.v[cell] = update_cell(P, W, g.m[cell], (int32_t)sx, (int32_t)sy, (int32_t)sz, x, y, z);
      if (keep) {
        g.mx[cell] = (int32_t)sx;
        g.my[cell] = (int32_t)sy;
        g.mz[cell] = (int32_t)sz;
      }
    }
  }
  if (t == 0) bw.ofl[b] = 0;
  __syncthreads();
}

// ---------------------------------------------------------------------------
// G2P over brick work items (WGSL_G2P, shaders.js:589-704).  With src_of set,
// slot d of the destination state takes the particle at old slot src_of[d]
// (re-sort; id, density and a non-unit mass move along) and every particle lies in the
// item's brick (EXACT: no particle can leave the window).  Otherwise G2P runs
// in place over the current items with the drift margin.
//
// GEN: the pass also takes the diffuse spawn decision of every particle it
// writes, on its new position and velocity, its density and the grid's
// decoded mass (what the separate decide pass would read after it), marking
// the spawners in go.
// DMOVED (with EXACT): the momentum pass already wrote every density to its
// new slot of D, so the re-sort does not move it.
template <int M, bool EXACT, bool GEN = false, bool DMOVED = false>
__device__ __forceinline__ void g2p_item(const Params &P, const PState &S, const PState &D,
                                         const GState &g, const Walls &W, const BrickWork &bw,
                                         const uint32_t *src_of, uint32_t blk,
                                         const GenOut &go = GenOut{}) {
  constexpr int E = 10 + 2 * M, V = E * E * E;
  __shared__ float4 lv[V];
  // Re-sort: the item's source slots, staged with the window.
  __shared__ uint32_t lsrc[EXACT ? kTB * kMaxPpt : 1];
  // A re-sort writes the next work-item set's order; it also empties the
  // resolve list for the next step (no resolve runs after G2P).
  uint4 it;
  const bool live =
      load_item(EXACT ? bw.items_w : bw.items, EXACT ? bw.nitems_w : bw.nitems, blk, it);
  if (EXACT && blk == 0 && threadIdx.x == 0) *bw.nlist = 0u;
  if (!live) return;
  int ibx, iby, ibz;
  brick_coords(bw, it.w, ibx, iby, ibz);
  const int ox = ibx * 8 - 1 - M, oy = iby * 8 - 1 - M, oz = ibz * 8 - 1 - M;
  const int t = threadIdx.x, nt = blockDim.x;
  // The window and the first kPerS rows of source slots are all loaded
  // before the first is stored (blocks of kTB threads): the staging costs
  // one memory latency, not one per round.  Items of more than kPerS rows
  // (16 particles per thread) stage the rest afterwards.
  constexpr int kPerV = (V + kTB - 1) / kTB, kPerS = EXACT ? 8 : 1;
  float4 v4[kPerV];
  uint32_t s4[kPerS];
  const int tw = opaque_lane(t);
#pragma unroll
  for (int u = 0; u < kPerV; u++) {
    const int c = tw + u * kTB;
    const int gi = c < V ? window_cell<E>(P, ox, oy, oz, c) : -1;
    v4[u] = gi >= 0 ? g.v[gi] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  if (EXACT)
#pragma unroll
    for (int u = 0; u < kPerS; u++) {
      const uint32_t i = (uint32_t)tw + u * kTB;
      s4[u] = i < it.z ? src_of[it.y + i] : 0u;
    }
#pragma unroll
  for (int u = 0; u < kPerV; u++)
    if (tw + u * kTB < V) lv[tw + u * kTB] = v4[u];
  if (EXACT) {
#pragma unroll
    for (int u = 0; u < kPerS; u++)
      if ((uint32_t)tw + u * kTB < it.z) lsrc[tw + u * kTB] = s4[u];
    for (uint32_t i = t + kPerS * kTB; i < it.z; i += kTB) lsrc[i] = src_of[it.y + i];
  }
  __syncthreads();
  uint32_t idx = t;
  if (idx >= it.z) return;
  const uint32_t gx = P.gx, slice = P.gx * P.gy;
  const uint32_t gen_seed = GEN ? *go.seed + 1u : 0u;
  // Software pipeline: the next particle's inputs (for a re-sort also the
  // mass, id and density it moves) load while this one computes.  Prefetch
  // indexes past the item's end are clamped to its last particle, so every
  // load is in bounds and unconditional.  The first particle is peeled off
  // the loop: vmcnt counts loads and stores in issue order, and with both
  // paths into the loop alike (loads, then the previous particle's stores)
  // the wait for a particle's inputs does not cover the stores before it.
  struct In {
    uint32_t s, id;
    float px, py, pz, m, dens;
  };
  auto load_in = [&](uint32_t i, In &in) {
    const uint32_t j = i < it.z ? i : it.z - 1;
    in.s = EXACT ? lsrc[j] : it.y + j;
    const float4 q = S.rec(in.s);  // position and id: one load
    in.px = q.x;
    in.py = q.y;
    in.pz = q.z;
    in.id = __float_as_uint(q.w);
    if (EXACT) {
      in.m = P.unit_mass ? 1.0f : S.at(PM, in.s);  // unit mass: the plane is never read
      in.dens = DMOVED ? 0.0f : S.dens(in.s);
    }
  };
  auto body = [&](uint32_t i, const In &cur) {
    const uint32_t d = it.y + i;
    const float px = cur.px, py = cur.py, pz = cur.pz;
    PW pw;
    particle_weights(px, py, pz, pw);
    WinPos wp = win_pos<M>(pw, ox, oy, oz);
    G2POut o;
    if (EXACT || wp.inside) {
      const int lb = wp.lb;
      g2p_particle(P, W, px, py, pz, pw, [&](int kk, int j, int ii) {
        float4 v = lv[lb + (kk * E + j) * E + ii];
        return make_float3(v.x, v.y, v.z);
      }, o);
    } else {
      uint32_t base = (uint32_t)((pw.b[2] * (int)P.gy + pw.b[1]) * (int)gx + pw.b[0]);
      g2p_particle(P, W, px, py, pz, pw, [&](int kk, int j, int ii) {
        float4 v = g.v[base + kk * slice + j * gx + ii];
        return make_float3(v.x, v.y, v.z);
      }, o);
    }
    store_g2p(D, d, o, cur.id);
    if (EXACT) {
      if (!P.unit_mass) D.at(PM, d) = cur.m;
      if (!DMOVED) D.dens(d) = cur.dens;
    }
    if (GEN) {
      const float dens = EXACT ? (DMOVED ? D.dens(d) : cur.dens) : S.dens(cur.s);
      const float m = P.unit_mass ? 1.0f : (EXACT ? cur.m : S.at(PM, cur.s));
      if (dens > 0.0f && dens <= P.density_bubble_min) {
        const uint32_t id = cur.id;
        GenEval e;
        if (gen_decide(P, g, dens, o.vel.x, o.vel.y, o.vel.z, o.pos.x, o.pos.y, o.pos.z, m, id,
                       gen_seed, e))
          gen_mark(go, id, d, e);
      }
    }
  };
  In cur, nxt;
  load_in(idx, cur);
  load_in(idx + nt, nxt);
  body(idx, cur);
  for (idx += nt; idx < it.z; idx += nt) {
    cur = nxt;
    load_in(idx + nt, nxt);
    body(idx, cur);
  }
}

// ---------------------------------------------------------------------------
// The re-sort's scan inside the step's kernels (ScanArgs): groups of 64
// brick ranks, kScanStride words per group: particles, work items, then the
// items of each size class (class k: R - k rows of kTB particles, R the rows
// of a full item; dispatch order is class 0 first).
static_assert(kScanStride == 2u + (uint32_t)kMaxPpt, "scan group stride");

__device__ __forceinline__ uint32_t wave_incl(uint32_t v, uint32_t lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t u = __shfl_up(v, off, 64);
    if (lane >= (uint32_t)off) v += u;
  }
  return v;
}

// A brick's particles in items: full ones, the rest, and the rest's class
// (kMaxPpt: no partial item).
struct BrickPieces {
  uint32_t c, full, rest, cls;
};
__device__ __forceinline__ BrickPieces brick_pieces(uint32_t c, uint32_t chs) {
  const uint32_t ch = 1u << chs, R = ch / kTB;
  BrickPieces b;
  b.c = c;
  b.full = c >> chs;
  b.rest = c & (ch - 1u);
  b.cls = b.rest ? R - (b.rest + kTB - 1u) / kTB : (uint32_t)kMaxPpt;
  return b;
}

// Phase 1, one wavefront (the brick-list kernel's block g): the group's sums,
// and its bases reserved on the running totals (lane q adds quantity q).
__device__ __forceinline__ void scan_group_sums(const BrickWork &bw, const ScanArgs &sa,
                                                uint32_t g) {
  const uint32_t lane = threadIdx.x & 63u, nb = bw.nbricks;
  const uint32_t R = (1u << sa.chs) / kTB;
  const uint32_t r = g * 64u + lane;
  const BrickPieces b = brick_pieces(r < nb ? sa.counts[r] : 0u, sa.chs);
  uint32_t sum[kScanStride];
  sum[0] = __shfl(wave_incl(b.c, lane), 63, 64);
  sum[1] = __shfl(wave_incl(b.full + (b.rest ? 1u : 0u), lane), 63, 64);
  sum[2] = __shfl(wave_incl(b.full + (b.cls == 0u ? 1u : 0u), lane), 63, 64);
#pragma unroll
  for (uint32_t k = 1; k < (uint32_t)kMaxPpt; k++)
    sum[2 + k] = k < R ? (uint32_t)__popcll(__ballot(b.cls == k)) : 0u;
  // lane q reserves quantity q: one atomic instruction for the group
  uint32_t mine = 0;
#pragma unroll
  for (uint32_t q = 0; q < kScanStride; q++) mine = lane == q ? sum[q] : mine;
  if (lane < kScanStride)
    sa.gsum[(size_t)g * kScanStride + lane] = atomicAdd(&sa.acc[lane * kAccStride], mine);
}

// Phase 2, one wavefront per group (the mass resolve kernel): the group's
// bases (its reserved ranges; class ranges also offset by the totals of the
// larger classes), then the brick spans, each brick's first item / item
// count / occupancy in the next set, and the work items at their dispatch
// positions (class range of the group plus the rank among the group's items
// of that class); the counts are cleared, and group 0 writes the item count.
__device__ __forceinline__ void scan_group_write(const BrickWork &bw, const ScanArgs &sa,
                                                 uint32_t g) {
  const uint32_t lane = threadIdx.x & 63u, nb = bw.nbricks;
  const uint32_t ch = 1u << sa.chs, R = ch / kTB;
  const uint32_t *gs = sa.gsum + (size_t)g * kScanStride;
  if (g == 0u && lane == 0) *bw.nitems_w = sa.acc[kAccStride];
  const uint32_t r = g * 64u + lane;
  const bool in = r < nb;
  const BrickPieces b = brick_pieces(in ? sa.counts[r] : 0u, sa.chs);
  if (in) sa.counts[r] = 0u;
  const uint32_t pieces = b.full + (b.rest ? 1u : 0u);
  const uint32_t ep = gs[0] + wave_incl(b.c, lane) - b.c;
  const uint32_t ei = gs[1] + wave_incl(pieces, lane) - pieces;
  // class k's range: the larger classes' totals, then the group's base
  const uint32_t n0 = b.full + (b.cls == 0u ? 1u : 0u);
  const uint32_t p0 = gs[2] + wave_incl(n0, lane) - n0;
  uint32_t pp = p0 + b.full;  // the partial item's position (class 0: after the full ones)
  uint32_t cb = sa.acc[2u * kAccStride];
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t k = 1; k < R; k++) {
    const uint64_t m = __ballot(b.cls == k);
    if (b.cls == k) pp = cb + gs[2 + k] + (uint32_t)__popcll(m & below);
    cb += sa.acc[(2u + k) * kAccStride];
  }
  if (in) {
    sa.spans[r] = make_uint2(ep, ep + b.c);
    const uint32_t bk = bw.brick_of_rank[r];
    bw.first_w[bk] = b.c ? ei : 0u;
    bw.count_w[bk] = pieces;
    bw.occ_w[bk] = b.c ? 1 : 0;
    for (uint32_t p = 0; p < b.full; p++)
      bw.items_w[p0 + p] = make_uint4(ei + p, ep + p * ch, ch, bk);
    if (b.rest) bw.items_w[pp] = make_uint4(ei + b.full, ep + b.full * ch, b.rest, bk);
  }
}

// ---------------------------------------------------------------------------
// Sort plumbing.

// Work items of a brick (index b): its range cut into pieces of at most
// ch = 1 << chs particles, numbered item0.. in brick-rank order.  That number
// (the item's ghost core) goes into the item's first field; the item itself is
// stored at its dispatch position: items of more rows of kTB particles come
// first (cur[k] is the next free position among items of R - k rows, R the
// rows of a full item), so the last blocks of a launch are the shortest.  The
// order among items of one size is irrelevant to every result.
__device__ __forceinline__ void write_items(const BrickWork &bw, uint32_t b, uint32_t start,
                                            uint32_t cnt, uint32_t chs, uint32_t item0,
                                            uint32_t *cur, uint32_t R) {
  const uint32_t ch = 1u << chs;
  const uint32_t full = cnt >> chs, rest = cnt & (ch - 1u);
  bw.first_w[b] = cnt ? item0 : 0u;
  bw.count_w[b] = full + (rest ? 1u : 0u);
  bw.occ_w[b] = cnt ? 1 : 0;
  if (full) {
    const uint32_t pos = atomicAdd(&cur[0], full);
    for (uint32_t p = 0; p < full; p++)
      bw.items_w[pos + p] = make_uint4(item0 + p, start + p * ch, ch, b);
  }
  if (rest) {
    const uint32_t pos = atomicAdd(&cur[R - (rest + kTB - 1u) / kTB], 1u);
    bw.items_w[pos] = make_uint4(item0 + full, start + full * ch, rest, b);
  }
}

// One block: exclusive scans of the per-brick particle counts (the new
// brick offsets) and of the per-brick work-item counts, the work items of
// the new order (numbered in brick-rank order, stored largest first: a first
// pass over the counts sizes each class) and their number; the counts are
// cleared for the next step.  Writes the next work-item set.  A round
// covers K * blockDim bricks, row k of it bricks base + k * blockDim + t, so
// every load and store of a row is coalesced; the rows are scanned in turn.
// The block size is a multiple of 64, at most 1024.
template <int K>
__device__ __forceinline__ void scan_items_block(const BrickWork &bw, uint32_t chs,
                                                 uint32_t *counts, uint32_t *offsets,
                                                 uint2 *spans = nullptr) {
  __shared__ uint32_t wp[16], wi[16];
  // items per size class (class k: R - k rows), then the next free dispatch
  // position of each class
  __shared__ uint32_t cur[kMaxPpt];
  const uint32_t ch = 1u << chs, R = ch / kTB;
  const uint32_t t = threadIdx.x, nt = blockDim.x, lane = t & 63u, wave = t >> 6;
  const uint32_t nw = nt >> 6, nb = bw.nbricks;
  if (t < (uint32_t)kMaxPpt) cur[t] = 0u;
  __syncthreads();
  for (uint32_t r = t; r < nb; r += nt) {
    const uint32_t c = counts[r], full = c >> chs, rest = c & (ch - 1u);
    if (full) atomicAdd(&cur[0], full);
    if (rest) atomicAdd(&cur[R - (rest + kTB - 1u) / kTB], 1u);
  }
  __syncthreads();
  if (t == 0) {
    uint32_t run = 0;
    for (uint32_t k = 0; k < R; k++) {
      const uint32_t v = cur[k];
      cur[k] = run;
      run += v;
    }
  }
  __syncthreads();
  uint32_t carry_p = 0, carry_i = 0;  // running totals, the same in every thread
  for (uint32_t base = 0; base < nb; base += nt * K) {
    uint32_t c[K], bk[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const uint32_t r = base + k * nt + t;
      c[k] = r < nb ? counts[r] : 0u;
      bk[k] = r < nb ? bw.brick_of_rank[r] : 0u;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      const uint32_t r = base + k * nt + t;
      const uint32_t pc = (c[k] + ch - 1) >> chs;
      uint32_t ip = c[k], ii = pc;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t vp = __shfl_up(ip, off, 64), vi = __shfl_up(ii, off, 64);
        if (lane >= (uint32_t)off) {
          ip += vp;
          ii += vi;
        }
      }
      if (lane == 63u) {
        wp[wave] = ip;
        wi[wave] = ii;
      }
      __syncthreads();
      uint32_t ep = carry_p, ei = carry_i, tp = 0, ti = 0;
      for (uint32_t w = 0; w < nw; w++) {
        if (w < wave) {
          ep += wp[w];
          ei += wi[w];
        }
        tp += wp[w];
        ti += wi[w];
      }
      __syncthreads();  // the wave sums are read
      ep += ip - c[k];
      ei += ii - pc;
      if (r < nb) {
        offsets[r] = ep;
        if (spans) spans[r] = make_uint2(ep, ep + c[k]);
        counts[r] = 0u;
        write_items(bw, bk[k], ep, c[k], chs, ei, cur, R);
      }
      carry_p += tp;
      carry_i += ti;
    }
  }
  if (t == 0) *bw.nitems_w = carry_i;
}




}  // namespace mpm
