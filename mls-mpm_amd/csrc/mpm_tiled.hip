// mpm_tiled.hip -- the brick-ordered, LDS-staged P2G / G2P and the particle
// re-sort.
//
// Layout.  Particles are kept grouped by the 8x8x8-cell brick holding them,
// bricks in Morton order, so every brick's particles are one contiguous range
// of the particle arrays.  A work item is a piece of one brick's range (at
// most 256 * ppt particles); inside an item the particles are stored
// column-major over rows of 256 slots (item_slot), so that a P2G block
// reading slots k*256 + t (coalesced) holds, in the 64 lanes of one
// instruction, particles that sit K = ceil(count/256) apart in the item,
// i.e. mostly in different cells.
//
// Re-sort (every `sort_period` steps, by default every step).  The P2G mass
// pass counts, per block in LDS, how many of its particles now lie in each of
// the 27 bricks around its own, reserves ranges of those bricks with one
// global atomic per brick and stores each particle's (brick, rank).  A scan
// over the per-brick counts gives brick offsets, the work items of the new
// order and the inverse permutation; G2P then writes every particle's new
// state straight into its new slot, so the re-sort moves no extra data.  Only
// loading a scene (caller order), or changing its particles, runs a separate
// scatter pass.
//
// P2G.  An item's block owns an LDS window: its brick, the 1-cell stencil
// halo and a drift margin of M cells ((10+2M)^3 cells).  The block adds the
// encoded i32 contributions of its particles into the window with LDS
// atomics, then stores the window's core (brick + halo, 10^3 cells) to the
// item's ghost buffer with plain stores.  Non-zero margin cells, and
// particles that drifted out of the window altogether, go to an overflow grid
// through global atomics and flag the bricks they touch.  A resolve pass (one
// block per brick that can hold mass) sums, for every cell, the ghost cores
// of the items of the surrounding bricks plus flagged overflow; the momentum
// resolve also performs the grid update.  Integer addition is associative,
// so every grid value is bit-identical to per-particle global atomics.
#include "mpm_tiled_dev.h"

namespace mpm {

// ---------------------------------------------------------------------------
// Kernels: one block / thread per index of the device bodies.
template <int M, bool UNIT>
__global__ __launch_bounds__(256) void k_p2g_mass_brick(Params P, PState S, BrickWork bw,
                                                        KeyOut ko) {
  p2g_mass_item<M, UNIT>(P, S, bw, ko, blockIdx.x);
}

template <int M, bool UNIT, bool SRC>
__global__ __launch_bounds__(256) void k_p2g_mom_brick(Params P, PState S, GState g,
                                                       BrickWork bw, SrcOut so) {
  p2g_mom_item<M, UNIT, SRC>(P, S, g, bw, blockIdx.x, so);
}

// One wavefront per 64 bricks: its needed bricks go to the list together,
// in brick order, with one atomic (the list follows the grid).  With
// sa.counts set, the same wavefront also sums the group of 64 brick ranks
// of the re-sort's scan (first phase).
__global__ __launch_bounds__(64) void k_brick_list(BrickWork bw, ScanArgs sa) {
  const uint32_t lane = threadIdx.x, b = blockIdx.x * 64u + lane;
  const bool need = brick_needed(bw, b);
  const uint64_t m = __ballot(need);
  if (m) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(bw.nlist, (uint32_t)__popcll(m));
    base = __shfl(base, 0, 64);
    if (need) bw.list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = b;
  }
  if (sa.counts) scan_group_sums(bw, sa, blockIdx.x);
}


template <int M, bool EXACT, bool GEN>
__global__ __launch_bounds__(256) void k_g2p_brick(Params P, PState S, PState D, GState g,
                                                   Walls W, BrickWork bw, const uint32_t *src_of,
                                                   GenOut go) {
  // a re-sort here follows a momentum pass that wrote the densities to
  // their new slots
  g2p_item<M, EXACT, GEN, EXACT>(P, S, D, g, W, bw, src_of, blockIdx.x, go);
}

// Offsets and work items of the new order, one block.
__global__ __launch_bounds__(1024) void k_scan_items(BrickWork bw, uint32_t chs, uint32_t *counts,
                                                     uint32_t *offsets) {
  scan_items_block<8>(bw, chs, counts, offsets);
}

// With sa.counts set, the first wavefront of each block first writes groups of
// the re-sort's scan (second phase), one group per block and round.
__global__ __launch_bounds__(kResolveThreads) void k_resolve_mass(Params P, GState g,
                                                                  BrickWork bw, ScanArgs sa) {
  if (sa.counts && threadIdx.x < 64u)
    for (uint32_t q = blockIdx.x; q < (bw.nbricks + 63u) >> 6; q += gridDim.x)
      scan_group_write(bw, sa, q);
  const uint32_t nl = *bw.nlist;
  for (uint32_t i = blockIdx.x; i < nl; i += gridDim.x)
    resolve_mass_entry<kResolveCells>(P, g, bw, bw.list[i]);
}

__global__ __launch_bounds__(kResolveThreads) void k_resolve_update(Params P, GState g,
                                                                    BrickWork bw, Walls W,
                                                                    int keep) {
  const uint32_t nl = *bw.nlist;
  for (uint32_t i = blockIdx.x; i < nl; i += gridDim.x)
    resolve_update_entry<kResolveCells>(P, g, bw, W, keep, bw.list[i]);
}

// First-step keys: brick and rank of every particle.  A block takes 4096
// consecutive slots (16 per thread, wavefronts over consecutive slots).  A
// reset fills bricks row by row, so a block holds large runs of a few dozen
// bricks: its particles are counted per brick in an LDS table (open
// addressing), one LDS atomic per brick per wavefront, and the block then
// reserves each brick's range with one global atomic.  Ranks inside a
// wavefront's group follow lane order, so consecutive ranks stay consecutive
// source slots.  A brick that finds no table entry within kProbe probes
// (more distinct bricks than the table holds) takes the global counter
// directly, one atomic per wavefront group.
constexpr uint32_t kKeysPerThread = 16, kKeyTable = 1024, kProbe = 32;
constexpr uint32_t kNoEntry = 0xffffffffu;
__global__ __launch_bounds__(256) void k_sort_keys(Params P, PState S, BrickWork bw, KeyOut ko) {
  __shared__ uint32_t tkey[kKeyTable], tcnt[kKeyTable];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < kKeyTable; i += 256u) {
    tkey[i] = kNoEntry;
    tcnt[i] = 0u;
  }
  __syncthreads();
  const uint32_t lane = __lane_id();
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t base = blockIdx.x * (256u * kKeysPerThread);
  uint32_t ent[kKeysPerThread], pre[kKeysPerThread];
#pragma unroll
  for (uint32_t j = 0; j < kKeysPerThread; j++) {
    const uint32_t s = base + j * 256u + t;
    const bool live = s < P.n;
    uint32_t dest = 0u, e = kNoEntry;
    if (live) {
      int bx, by, bz;
      cell_brick(P, S.at(PX, s), S.at(PY, s), S.at(PZ, s), bx, by, bz);
      dest = bw.brick_rank[brick_index(bw, bx, by, bz)];
      ko.keys[s] = dest;
      uint32_t h = (dest * 2654435761u) >> 22;
      for (uint32_t k = 0; k < kProbe; k++) {
        const uint32_t prev = atomicCAS(&tkey[h], kNoEntry, dest);
        if (prev == kNoEntry || prev == dest) {
          e = h;
          break;
        }
        h = (h + 1u) & (kKeyTable - 1u);
      }
    }
    // group the lanes by table entry (or by brick, without one); the group's
    // first lane counts it with one atomic
    uint64_t pending = __ballot(live);
    uint32_t leader_of = lane, prefix = 0u, group = 0u;
    while (pending) {
      const int leader = __ffsll((unsigned long long)pending) - 1;
      const uint32_t le = __shfl(e, leader, 64), ld = __shfl(dest, leader, 64);
      const bool mine = live && e == le && dest == ld && ((pending >> lane) & 1ull);
      const uint64_t same = __ballot(mine);
      if (mine) {
        leader_of = (uint32_t)leader;
        prefix = (uint32_t)__popcll(same & below);
      }
      if (lane == (uint32_t)leader) group = (uint32_t)__popcll(same);
      pending &= ~same;
    }
    uint32_t gb = 0u;
    if (group) gb = e != kNoEntry ? atomicAdd(&tcnt[e], group) : atomicAdd(&ko.counts[dest], group);
    gb = __shfl(gb, (int)leader_of, 64);
    ent[j] = e;
    pre[j] = gb + prefix;  // rank in the block's share (table) or final rank (global)
  }
  __syncthreads();
  for (uint32_t i = t; i < kKeyTable; i += 256u)
    if (tkey[i] != kNoEntry) tcnt[i] = atomicAdd(&ko.counts[tkey[i]], tcnt[i]);
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < kKeysPerThread; j++) {
    const uint32_t s = base + j * 256u + t;
    if (s < P.n) ko.rank[s] = ent[j] != kNoEntry ? tcnt[ent[j]] + pre[j] : pre[j];
  }
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

// The first step's data move: every particle to slot offsets[brick] + rank.
// The order inside a brick's range is free (every pass reads an item's slots
// as a set), so this keeps the order the keys pass handed out: a reset fills
// bricks row by row, consecutive source slots mostly take consecutive ranks,
// and the loads and the stores both stay contiguous.  (A block-shaped form,
// 16 slots per thread as in the keys pass, measured slower: 227 against
// 187 us on the 4.2M-particle dam break, at 4 blocks per CU.)
__global__ __launch_bounds__(256) void k_sort_scatter(uint32_t n, PState src, PState dst,
                                                      const uint32_t *keys, const uint32_t *rank,
                                                      const uint32_t *offsets) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t d = offsets[keys[s]] + rank[s];
  dst.rec(d) = src.rec(s);
#pragma unroll
  for (int k = VX; k < NPLANES; k++) dst.at(k, d) = src.at(k, s);
  dst.density[d] = src.density[s];
}

// ---------------------------------------------------------------------------
// Host launchers.
static inline uint32_t nblk(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }
static inline uint32_t log2_pow2(uint32_t v) { return (uint32_t)__builtin_ctz(v); }

bool tiled_ppt_supported(uint32_t ppt) { return ppt == 4 || ppt == 8 || ppt == 16; }
bool tiled_margin_supported(uint32_t m) { return m == 1 || m == 2; }
bool tiled_threads_supported(uint32_t t) { return t == 256; }

uint32_t max_items(uint32_t n, uint32_t nbricks, uint32_t ppt) {
  return nbricks + n / (kTB * ppt) + 1;
}

static uint32_t resolve_grid(const BrickWork &bw) { return bw.nbricks < 2048 ? bw.nbricks : 2048; }

void launch_p2g_mass_tiled(const Params &P, const PState &S, const GState &g, const TileCfg &tc,
                           const BrickWork &bw, const KeyOut &ko, hipStream_t st) {
  if (!P.n) return;
  uint32_t grid = max_items(P.n, bw.nbricks, tc.ppt), tb = tc.threads;
  if (tc.margin == 1) {
    if (P.unit_mass) k_p2g_mass_brick<1, true><<<grid, tb, 0, st>>>(P, S, bw, ko);
    else k_p2g_mass_brick<1, false><<<grid, tb, 0, st>>>(P, S, bw, ko);
  } else {
    if (P.unit_mass) k_p2g_mass_brick<2, true><<<grid, tb, 0, st>>>(P, S, bw, ko);
    else k_p2g_mass_brick<2, false><<<grid, tb, 0, st>>>(P, S, bw, ko);
  }
  (void)g;
}

void launch_resolve_mass(const Params &P, const GState &g, const BrickWork &bw, hipStream_t st,
                         const ScanArgs &sa) {
  if (!P.n) return;
  // 64-thread blocks: one brick per thread spread over more CUs
  k_brick_list<<<nblk(bw.nbricks, 64), 64, 0, st>>>(bw, sa);
  k_resolve_mass<<<resolve_grid(bw), kResolveThreads, 0, st>>>(P, g, bw, sa);
}

ScanArgs scan_args(const SortBufs &sb, const TileCfg &tc) {
  ScanArgs sa;
  sa.counts = sb.counts;
  sa.spans = sb.spans;
  sa.gsum = sb.gsum;
  sa.acc = sb.gacc;
  sa.chs = log2_pow2(kTB * tc.ppt);
  return sa;
}

SrcOut src_out(const SortBufs &sb, const TileCfg &tc, float *dens_w) {
  SrcOut so;
  so.keys = sb.keys;
  so.rank = sb.rank;
  so.spans = sb.spans;
  so.src_of = sb.src_of;
  so.dens_w = dens_w;
  so.chs = log2_pow2(kTB * tc.ppt);
  return so;
}

void launch_p2g_mom_tiled(const Params &P, const PState &S, const GState &g, const TileCfg &tc,
                          const BrickWork &bw, hipStream_t st, const SrcOut &so) {
  if (!P.n) return;
  uint32_t grid = max_items(P.n, bw.nbricks, tc.ppt), tb = tc.threads;
#define MOM_LAUNCH(M_, U_)                                                                  \
  do {                                                                                      \
    if (so.src_of) k_p2g_mom_brick<M_, U_, true><<<grid, tb, 0, st>>>(P, S, g, bw, so);     \
    else k_p2g_mom_brick<M_, U_, false><<<grid, tb, 0, st>>>(P, S, g, bw, so);              \
  } while (0)
  if (tc.margin == 1) {
    if (P.unit_mass) MOM_LAUNCH(1, true);
    else MOM_LAUNCH(1, false);
  } else {
    if (P.unit_mass) MOM_LAUNCH(2, true);
    else MOM_LAUNCH(2, false);
  }
#undef MOM_LAUNCH
}

void launch_resolve_update(const Params &P, const GState &g, const TileCfg &tc,
                           const BrickWork &bw, const Walls &W, bool keep, hipStream_t st) {
  (void)tc;
  k_resolve_update<<<resolve_grid(bw), kResolveThreads, 0, st>>>(P, g, bw, W, keep ? 1 : 0);
}

// G2P over the current work items.  Re-sort (src_of set): items of the new
// order, exact windows.  In place: margin as configured.
void launch_g2p_tiled(const Params &P, const PState &S, const PState &D, const GState &g,
                      const Walls &W, const TileCfg &tc, const BrickWork &bw,
                      const uint32_t *src_of, hipStream_t st, const GenOut &go) {
  if (!P.n) return;
  uint32_t grid = max_items(P.n, bw.nbricks, tc.ppt), tb = tc.g2p_threads;
#define G2P_LAUNCH(M_, EX_, SRC_)                                                             \
  do {                                                                                        \
    if (go.bits) k_g2p_brick<M_, EX_, true><<<grid, tb, 0, st>>>(P, S, D, g, W, bw, SRC_, go);  \
    else k_g2p_brick<M_, EX_, false><<<grid, tb, 0, st>>>(P, S, D, g, W, bw, SRC_, go);         \
  } while (0)
  if (src_of) G2P_LAUNCH(0, true, src_of);
  else if (tc.margin == 1) G2P_LAUNCH(1, false, nullptr);
  else G2P_LAUNCH(2, false, nullptr);
#undef G2P_LAUNCH
}

KeyOut key_out(const SortBufs &sb, const BrickWork &bw) {
  (void)bw;
  KeyOut ko;
  ko.counts = sb.counts;
  ko.keys = sb.keys;
  ko.rank = sb.rank;
  ko.acc = nullptr;
  ko.cols = 0;
  return ko;
}

// Full sort of the caller's order (at a reset on the brick path; before the
// next step after particles were appended or deleted): brick keys, offsets, work items
// and a scatter of every particle into its slot.  Leaves the counts zeroed.
void launch_sort(const Params &P, const PState &src, const PState &dst, const SortBufs &sb,
                 const BrickWork &bw, const TileCfg &tc, hipStream_t st) {
  if (!P.n) return;
  k_sort_keys<<<nblk(P.n, 256 * kKeysPerThread), 256, 0, st>>>(P, src, bw, key_out(sb, bw));
  k_scan_items<<<1, 1024, 0, st>>>(bw, log2_pow2(kTB * tc.ppt), sb.counts, sb.offsets);
  k_sort_scatter<<<nblk(P.n, 256), 256, 0, st>>>(P.n, src, dst, sb.keys, sb.rank, sb.offsets);
}

void launch_clear_list(const BrickWork &bw, hipStream_t st) {
  k_zero_u32<<<1, 1, 0, st>>>(bw.nlist);
}

}  // namespace mpm
