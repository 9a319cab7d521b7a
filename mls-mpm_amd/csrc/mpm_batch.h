// mpm_batch.h -- many scenes stepped together: per-scene device descriptors
// and the batched stage launchers (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/mlsmpm.h"
#include "mpm_kernels.h"
#include "mpm_tiled.h"

namespace mpm {

// Everything one scene's stage bodies need, for one buffer parity: the
// batched kernels read it from a device table indexed by scene.
struct SceneDev {
  Params P;
  PState S;            // current particle buffer
  PState D;            // re-sort destination buffer
  GState g;
  BrickWork bw;
  KeyOut ko;           // brick counts, keys and ranks recorded by P2G mass
  uint32_t *offsets;   // per brick rank: first particle of the new order
  uint32_t *src_of;    // per new slot: old slot
  uint2 *spans;        // per brick rank: [first, end) slots of the new order
  SrcOut so;           // the momentum pass's re-sort outputs (src_of, densities)
  uint32_t ch;         // work-item size in particles (256 * ppt)
  int32_t keep;        // keep the i32 momentum grids readable
  Walls W;
};

// The work units of one stage over all scenes: unit u belongs to the scene
// si with pre[si] <= u < pre[si + 1].
struct BatchView {
  const SceneDev *sd;
  const uint32_t *pre;
  uint32_t ns, total;
};

// Persistent blocks of each batched kernel class (0 = one block per unit).
struct BatchGrid {
  uint32_t items, resolve, aux;
};

// Per-stage unit prefixes (device, ns + 1 words each) of a batch table.
struct BatchTables {
  const SceneDev *sd;
  const uint32_t *pre_items;    // work items (upper bound per scene)
  const uint32_t *pre_bricks;   // bricks (upper bound of the resolve list)
  const uint32_t *pre_bblocks;  // 256-thread blocks over bricks
  uint32_t ns;
  uint32_t tot_items, tot_bricks, tot_bblocks;
  uint32_t all_unit;            // every scene has unit particle mass
  // The step's resolve list over all scenes, (scene, brick) pairs
  // (capacity tot_bricks), and its length; empty at every step start.
  uint2 *list;
  uint32_t *nlist;
  // Dispatch order over all scenes of the current work-item set and of the
  // set the step's re-sort writes: (scene, item) pairs, items with more rows
  // of 256 particles first (capacity tot_items); with the totals per row
  // class and in all (kMaxPpt + 1 counters, kAccStride words apart, the
  // item count last).  ocls: per scene, the bases reserved per class.
  const uint2 *order_cur;
  const uint32_t *tot_cur;
  uint2 *order_next;
  uint32_t *tot_next;
  uint32_t *ocls;
};

// The stream the batch's members run on.
hipStream_t batch_stream(const mpm_batch *b);

// Batched steps replayed from captured graphs, and graphs captured, over the
// batch's life.
void batch_graph_counts(const mpm_batch *b, uint64_t *graph_steps, uint64_t *captures);

// Count of the items of an order (its last counter).
constexpr uint32_t kOrderTotal = 16u;  // = kMaxPpt

// The dispatch order of a work-item set (the current one, or the one the
// re-sort writes: next) from the members' item lists.
void launch_batch_order(const BatchTables &t, bool next, hipStream_t st);

// One batched step (sort period 1: every step re-sorts inside G2P) of every
// scene of the table, all with drift margin `margin`.
// The re-sort preparation runs on `side`, forked after the mass pass with
// event `fork` and joined before G2P with `join`.
int launch_batch_step(const BatchTables &t, uint32_t margin, const BatchGrid &grid,
                      hipStream_t st, hipStream_t side, hipEvent_t fork, hipEvent_t join);

}  // namespace mpm
