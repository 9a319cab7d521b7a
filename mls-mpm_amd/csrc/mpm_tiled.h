// mpm_tiled.h -- launchers of the sorted / LDS-staged kernels (internal).
#pragma once
#include "mpm_kernels.h"

namespace mpm {

// Tile shape of the LDS-staged P2G: chunks of 64*slices particles, windows of
// at most wmax cells (LDS: 4 B per cell for mass, 16 B per cell for momentum).
struct TileCfg {
  uint32_t slices;  // 4, 8, 16 or 32
  uint32_t wmax;    // window cells
};

// Where the P2G mass pass records each particle's cell key and its rank among
// the particles of that cell (counts == nullptr: do not record).
struct KeyOut {
  uint32_t *counts, *keys, *rank;
  const uint32_t *tile_rank;
  uint32_t tdx, tdy;
};

bool tiled_slices_supported(uint32_t sl);
KeyOut key_out(const SortBufs &sb, uint32_t tdx, uint32_t tdy);
void launch_p2g_mass_tiled(const Params &P, const PState &S, const GState &g, const TileCfg &tc,
                           int32_t *boxes, const KeyOut &ko, hipStream_t st);
void launch_p2g_mom_tiled(const Params &P, const PState &S, const GState &g, const TileCfg &tc,
                          const int32_t *boxes, hipStream_t st);
void launch_sort(const Params &P, const PState &src, const PState &dst, const SortBufs &sb,
                 uint32_t tdx, uint32_t tdy, uint32_t slices, hipStream_t st);
void launch_relayout_prep(const Params &P, const SortBufs &sb, uint32_t slices, hipStream_t st);

}  // namespace mpm
