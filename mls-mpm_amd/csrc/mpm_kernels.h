// mpm_kernels.h -- host-side launchers of the MLS-MPM kernels (internal).
#pragma once
#include "mpm_device.h"

namespace mpm {

// Device buffers of the diffuse pool and its per-step scratch.
struct DiffuseBufs {
  DRec *pool;          // dmax records (DiffuseGPU.particleBuffer layout)
  uint32_t *free_list; // dmax slot indices (stack)
  int32_t *top;        // stack top
  uint32_t *seed;      // last-used DiffuseGPU._seed
  uint32_t *genbits;   // spawner bitmask by reset-order id, ceil(cap/32) words
  float4 *genval;      // per spawner id: (slot bits, score, speed, density), cap entries
  uint32_t *killbits;  // killed-slot bitmask, ceil(dmax/32) words
  uint32_t *alivebits; // alive flag of every slot (a copy of the records'), ceil(dmax/32) words
  uint32_t *list;      // dmax words: the spawners in rank order
  uint32_t *scratch;   // scan tile sums
  uint32_t *total;     // [0] spawners, [1] stack top after the spawns
};

// RayParams (shaders.js:78): origin, cylinder radius, direction.
struct Ray {
  float ox, oy, oz, radius, dx, dy, dz, pad;
};

// The hand push on the device (HandParams, shaders.js:394-402): the box of
// cells to visit (bmin > bmax: none) was computed on the host.
struct HandDev {
  float px, py, pz, radius, vx, vy, vz, strength, ex, ey, ez;
  uint32_t on;
  int32_t bmin[3], bmax[3];
};

// Sort buffers (mpm_tiled.hip).
struct SortBufs {
  uint32_t *keys;       // per particle: Morton rank of its new brick
  uint32_t *rank;       // per particle: rank among its brick's particles
  uint32_t *counts;     // per brick rank
  uint32_t *offsets;    // per brick rank, exclusive scan of counts
  uint32_t *scratch;    // scan scratch
  uint32_t *brick_rank; // 8x8x8 grid brick -> position in Morton order
  uint32_t *src_of;     // per new slot: old slot (G2P re-layout)
  uint32_t *gsum;       // re-sort scan: per group of 64 brick ranks, its bases
  uint32_t *gacc;       // re-sort scan: running totals
  uint2 *spans;         // re-sort scan: per brick rank, [first, end) slots
  uint32_t nkeys;       // number of bricks
};

void launch_clear_grid(const GState &g, uint32_t G, hipStream_t st);
void launch_p2g_mass_global(const Params &P, const PState &S, const GState &g, hipStream_t st);
void launch_p2g_mom_global(const Params &P, const PState &S, const GState &g, hipStream_t st);
void launch_update_dense(const Params &P, const GState &g, const Walls &W, bool clear,
                         hipStream_t st);
void launch_g2p_global(const Params &P, const PState &S, const GState &g, const Walls &W,
                       hipStream_t st);
void launch_g2p_relayout(const Params &P, const PState &S, const PState &D, const GState &g,
                         const Walls &W, const uint32_t *src_of, hipStream_t st);
// Rebuilds the alive bitmask from the pool records (after a host write).
void launch_alive_bits(const DiffuseBufs &D, uint32_t dmax, hipStream_t st);
// decided: the G2P pass already marked the spawners (GenOut).
void launch_diffuse(const Params &P, const PState &S, const GState &g, const Walls &W,
                    const DiffuseBufs &D, hipStream_t st, bool decided = false);
uint32_t scan_scratch_words(uint32_t n);
void launch_scan(const uint32_t *in, uint32_t n, bool popc, uint32_t *out, uint32_t *scratch,
                 uint32_t *total, hipStream_t st, uint32_t *clear = nullptr);
void launch_export_particles(const PState &S, uint32_t n, float4 *pos4, float4 *vel4, float *C9,
                             float *dens, hipStream_t st, bool reset_order = true);
// Copies bytes (a multiple of 16) with a kernel: works on any device pointer
// of this process, whichever runtime allocated it.
void launch_copy16(const void *src, void *dst, size_t bytes, hipStream_t st);
void launch_export_grid(const GState &g, uint32_t G, int32_t *m, int32_t *mv3, float *mf,
                        float *v3, hipStream_t st);
// Slots first .. first+n-1 (ids alike) from host-layout arrays.
void launch_import_particles(const PState &S, uint32_t n, const float *pos3, const float *vel3,
                             const float *C9, const float *mass, hipStream_t st,
                             uint32_t first = 0);

// Interaction (mpm_interact.hip).  Hand: between the grid update and G2P;
// act (brick-sparse update) receives the touched bricks.
void launch_apply_hand(const Params &P, const GState &g, uint8_t *act, uint32_t nbx,
                       uint32_t nby, const HandDev *h, hipStream_t st);
// Survivors of the cylinder r, from S into D in reset order (slot = new id);
// *total receives their number.  bits/prefix: nwords = ceil(n/32) words.
void launch_delete_mark_move(const PState &S, const PState &D, uint32_t n, const Ray &r,
                             uint32_t *bits, uint32_t *prefix, uint32_t *scratch,
                             uint32_t *total, uint32_t nwords, hipStream_t st);
// *hit = bits of the smallest t > 0 (0xffffffff: none).
void launch_raycast(const PState &S, uint32_t n, const Ray &r, uint32_t *hit, hipStream_t st);

}  // namespace mpm
