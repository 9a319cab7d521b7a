// mpm_tiled.h -- launchers of the sorted / LDS-staged kernels (internal).
#pragma once
#include "mpm_kernels.h"
#include "mpm_ops.h"

namespace mpm {

// Shape of the LDS-staged P2G: work items of at most 256 * ppt particles of
// one 8x8x8 brick, LDS windows of (10 + 2 * margin)^3 cells.
struct TileCfg {
  uint32_t ppt;          // item size in units of 256 particles: 4, 8 or 16
  uint32_t margin;       // drift margin in cells: 1 or 2
  uint32_t threads;      // P2G threads per item: 256, 512 or 1024
  uint32_t g2p_threads;  // G2P threads per item: 256, 512 or 1024
};

// Work items of the brick-ordered particle arrays.
// Two sets of work items alternate with the particle buffers: the current
// set (the particle order the P2G passes read) and the next set, written by
// a re-sort and read by the re-sorting G2P.
struct BrickWork {
  const uint4 *items;            // current, in dispatch order: (ghost core, first particle,
                                 // count, brick index)
  uint4 *items_w;                // next
  uint32_t *nitems, *nitems_w;   // device counts of items (current, next)
  uint32_t *first, *count;       // current, per brick index: first item and number of items
  uint32_t *first_w, *count_w;   // next
  const uint32_t *brick_of_rank; // brick index of each Morton rank
  const uint32_t *brick_rank;    // Morton rank of each brick index
  int32_t *gm;                   // ghost mass cores (10^3 cells), one per item
  int32_t *gx, *gy, *gz;         // ghost momentum cores, one per item
  int32_t *om, *ox, *oy, *oz;    // overflow grids (zero outside flagged bricks)
  uint8_t *ofl;                  // per brick: overflow grid may hold values here
  uint8_t *act;                  // per brick: bit 0 active this step, bit 1 last step
  uint8_t *occ, *occ_w;          // per brick: has particles in the current / next order
  uint32_t *list, *nlist;        // bricks to resolve this step
  uint32_t nbx, nby, nbz, nbricks;
};

// The re-sort's scan of the per-brick counts, spread over the step's own
// kernels: brick ranks in groups of 64 (one wavefront).  The brick-list
// kernel sums each group (particles, work items, items per size class) and
// reserves the group's ranges with one atomic per quantity on running
// totals (zeroed by the mass pass); the mass resolve kernel then writes the
// group's brick spans and work items.  Groups take their ranges in arrival
// order: each brick's particles stay contiguous, the bricks of a group
// consecutive.  counts == nullptr: no re-sort this step.
constexpr uint32_t kScanStride = 18;  // quantities: 2 + the largest particles-per-thread
constexpr uint32_t kAccStride = 32;   // words between running totals (one cache line each)
struct ScanArgs {
  uint32_t *counts;  // per brick rank, cleared by the second phase
  uint2 *spans;      // per brick rank: [first, end) slots of the brick in the new order
  uint32_t *gsum;    // per group, kScanStride words: the group's reserved bases
  uint32_t *acc;     // kScanStride running totals, kAccStride words apart
  uint32_t chs;      // log2 of the work-item size
};
// The inverse permutation of the re-sort, written by the momentum pass (it
// reads every particle of the current order after the scan has run).
struct SrcOut {
  const uint32_t *keys, *rank;
  const uint2 *spans;
  uint32_t *src_of;
  // set: the particle densities go to their new slots in this plane (the
  // destination state's), and the re-sorting G2P leaves them there
  float *dens_w;
  uint32_t chs;
};

// Where the P2G mass pass records each particle's new brick (Morton rank) and
// its rank among that brick's particles (counts == nullptr: do not record).
struct KeyOut {
  uint32_t *counts, *keys, *rank;
  uint32_t *acc;  // set: the pass zeroes the re-sort scan's running totals (ScanArgs)
  // Rows of a full work item (its particles / 256).  Above 1, the particles
  // of a wavefront that stay in their brick take their block of consecutive
  // ranks column by column (rank_in_block), so that the slots the lanes of
  // one row read in G2P come from consecutive slots of the old order.
  uint32_t cols;
};

bool tiled_ppt_supported(uint32_t ppt);
bool tiled_margin_supported(uint32_t m);
bool tiled_threads_supported(uint32_t t);
uint32_t max_items(uint32_t n, uint32_t nbricks, uint32_t ppt);
KeyOut key_out(const SortBufs &sb, const BrickWork &bw);
// The resolve list is empty when a step starts: the re-sorting G2P empties
// it, and a step without a re-sort ends with launch_clear_list.
void launch_p2g_mass_tiled(const Params &P, const PState &S, const GState &g, const TileCfg &tc,
                           const BrickWork &bw, const KeyOut &ko, hipStream_t st);
// The resolve list (bricks to resolve) and the mass resolve; with sa.counts
// set, also the re-sort's scan (offsets, next work items and their count).
void launch_resolve_mass(const Params &P, const GState &g, const BrickWork &bw, hipStream_t st,
                         const ScanArgs &sa = ScanArgs{});
ScanArgs scan_args(const SortBufs &sb, const TileCfg &tc);
SrcOut src_out(const SortBufs &sb, const TileCfg &tc, float *dens_w);
void launch_clear_list(const BrickWork &bw, hipStream_t st);
// so.src_of set: the pass also writes the re-sort's inverse permutation.
void launch_p2g_mom_tiled(const Params &P, const PState &S, const GState &g, const TileCfg &tc,
                          const BrickWork &bw, hipStream_t st, const SrcOut &so = SrcOut{});
void launch_resolve_update(const Params &P, const GState &g, const TileCfg &tc,
                           const BrickWork &bw, const Walls &W, bool keep, hipStream_t st);
// go.bits set: the pass also decides the diffuse spawners (launch_diffuse
// with decided = true skips its own decide pass).
void launch_g2p_tiled(const Params &P, const PState &S, const PState &D, const GState &g,
                      const Walls &W, const TileCfg &tc, const BrickWork &bw,
                      const uint32_t *src_of, hipStream_t st, const GenOut &go = GenOut{});
void launch_sort(const Params &P, const PState &src, const PState &dst, const SortBufs &sb,
                 const BrickWork &bw, const TileCfg &tc, hipStream_t st);

}  // namespace mpm
