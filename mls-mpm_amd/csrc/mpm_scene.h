// mpm_scene.h -- scene object behind the C ABI (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/mlsmpm.h"
#include "mpm_arena.h"
#include "mpm_batch.h"
#include "mpm_kernels.h"
#include "mpm_tiled.h"

namespace mpm {
// Slabs of device memory, or of pinned host memory, for a SlabArena.
struct HipSlabs {
  int device;
  bool pinned;
  static constexpr int kNoMemory = (int)hipErrorOutOfMemory;
  int get(size_t bytes, void **p) {
    hipError_t e = pinned ? hipHostMalloc(p, bytes, hipHostMallocDefault) : hipMalloc(p, bytes);
    if (e != hipSuccess) (void)hipGetLastError();
    return (int)e;
  }
  void put(void *p) {
    (void)hipSetDevice(device);
    (void)(pinned ? hipHostFree(p) : hipFree(p));
  }
};

// Device (or pinned host) memory from slabs kept until the arena is
// destroyed (mpm_arena.h).  A scene created over a device arena takes its
// buffers from it and gives them back when it is freed, so a loop that
// creates and frees scenes calls hipMalloc only while the slabs grow, and
// never hipFree (which waits for the whole device).
class Arena : public SlabArena<HipSlabs> {
 public:
  Arena(int device, size_t slab_bytes, bool pinned = false, size_t cap = 0)
      : SlabArena<HipSlabs>(HipSlabs{device, pinned}, slab_bytes, cap) {}
};
}  // namespace mpm

struct mpm_scene {
  mpm_config cfg{};
  mpm::Arena *arena = nullptr;  // device memory source (nullptr: hipMalloc)
  hipStream_t stream = nullptr;
  uint32_t n = 0;        // active particles
  uint32_t G = 0;        // cells
  uint64_t steps = 0;    // steps since reset
  int cur = 0;           // which particle buffer is current (sorting ping-pongs)
  float *planes[2] = {nullptr, nullptr};  // (x, y, z, id) records, then the planes VX .. PM
  float *dens[2] = {nullptr, nullptr};  // particle_density, per layout buffer
  int unit_mass = 1;
  mpm::GState g{};
  mpm::DiffuseBufs d{};
  mpm::BrickWork bw{};
  // the two work-item sets (current = index cur, alternating with the
  // particle buffers)
  uint4 *items_buf[2] = {nullptr, nullptr};
  uint32_t *nitems_buf[2] = {nullptr, nullptr};
  uint32_t *first_buf[2] = {nullptr, nullptr}, *count_buf[2] = {nullptr, nullptr};
  uint8_t *occ_buf[2] = {nullptr, nullptr};
  uint32_t *brick_of_rank = nullptr;
  mpm::SortBufs sort{};
  mpm::Walls walls{};
  uint32_t nwalls = 0;
  uint32_t since_sort = 0;  // steps since the particles were last sorted
  bool sorted_once = false; // particles sorted at least once since reset
  mpm::TileCfg tc{8, 2, 256, 256};
  void *stage = nullptr;    // device staging for reads/writes
  size_t stage_bytes = 0;
  // Stage timing: per step, events bracket sort, P2G mass, P2G momentum,
  // grid update, G2P and diffuse; folded into sums when the pool fills.
  bool timing = false;
  static constexpr int kEvPerStep = 8;
  static constexpr int kEvSteps = 64;
  hipEvent_t ev[kEvPerStep * kEvSteps] = {};
  int ev_used = 0;               // steps recorded in the pool, not yet folded
  double stage_sum[8] = {};      // ms: p2g_mass, p2g_mom, update, g2p, diffuse, sort, total, relayout
  uint64_t timed_steps = 0;
  std::vector<float> step_ms;    // per-step totals since timing was enabled
  // Captured multi-step graphs, one set per start state (buffer parity and
  // steps since the last re-sort), least recently used replaced.  Untimed:
  // exec[0] only.  Timed: exec[0] and exec[1] record their steps' events into
  // the two halves of ev[] and are replayed alternately, so one replay's
  // events are read while the other runs.
  struct GraphSet {
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    bool timed = false;
    int cur = 0;
    uint32_t since = 0;
    int post_cur = 0;
    uint32_t post_since = 0;
    uint64_t used = 0;
  };
  static constexpr int kGraphSets = 4;
  GraphSet gset[kGraphSets];
  uint32_t graph_len = 0;        // steps per graph
  uint64_t graph_clock = 0;
  int graph_next = 0;            // timed: event half the next replay uses
  int graph_pending = -1;        // timed: event half not yet folded
  // Interaction, allocated at first use: the hand push on the device (the
  // step applies it once it was switched on), and the scratch of deleteNear
  // and raycastFluid.
  mpm::HandDev *hand_dev = nullptr;
  bool hand_on = false;
  uint32_t *ibits = nullptr, *iprefix = nullptr, *iscratch = nullptr, *itotal = nullptr,
           *ihit = nullptr;
  // Renderer handoff: two frames, each written to device staging on the
  // scene's stream and copied into pinned host memory on copy_stream.
  struct Frame {
    char *dev = nullptr;          // device staging (freed with owned[])
    char *host = nullptr;         // pinned, frame_bytes
    hipEvent_t ready = nullptr;   // staging written
    hipEvent_t done = nullptr;    // host copy finished
    uint64_t ticket = 0;          // 0: never used
    uint64_t steps = 0;
    uint32_t n = 0, what = 0;
    size_t off[5] = {};           // pos, vel, density, diffuse records, top
  };
  Frame frames[2];
  size_t frame_bytes = 0;
  hipStream_t copy_stream = nullptr;
  uint64_t frame_next = 1;
  // Batch membership: a member's work runs on the batch's stream.
  mpm_batch *batch = nullptr;
  hipStream_t own_stream = nullptr;
  // Bumped by every change a batch's descriptor table depends on (steps
  // outside the batch, reset, parameters, walls).
  uint64_t version = 0;
  size_t bytes = 0;         // device bytes held by the scene
  void *owned[64] = {};     // allocations to free
  int nowned = 0;
};

namespace mpm {
int set_error(int code, const char *msg);
int hip_fail(hipError_t e, const char *what);
int validate_config(const mpm_config &c);
// arena: where the scene's device buffers come from (nullptr: hipMalloc).
int scene_init(mpm_scene *s, const mpm_config &c, Arena *arena = nullptr);
void scene_free(mpm_scene *s);
int scene_reset(mpm_scene *s, uint32_t n, const float *pos3, const float *vel3, const float *C9,
                const float *mass, bool device_ptrs);
int scene_step(mpm_scene *s, uint32_t nsteps);
int scene_prepare(mpm_scene *s);
int scene_read_particles(mpm_scene *s, float *pos4, float *vel4, float *C9, float *density);
int scene_export_async(mpm_scene *s, hipStream_t st, hipStream_t copy, void *dst,
                       hipEvent_t exported, hipEvent_t done);
int scene_read_grid(mpm_scene *s, int32_t *m, int32_t *mv3, float *mf, float *v3);
int scene_read_diffuse(mpm_scene *s, void *rec48, uint32_t *free_list, int32_t *top);
int scene_write_diffuse(mpm_scene *s, const void *rec48, const uint32_t *free_list, int32_t top);
int scene_set_seed(mpm_scene *s, uint32_t seed);
int scene_set_walls(mpm_scene *s, const float *wallp);
int scene_set_hand(mpm_scene *s, const mpm_hand *hand);
int scene_delete_near(mpm_scene *s, const float *origin, const float *dir, float radius,
                      uint32_t *remaining);
int scene_raycast(mpm_scene *s, const float *origin, const float *dir, float radius, float *t,
                  int *hit);
int scene_raycast_noalloc(mpm_scene *s, uint32_t nrays, const float *origins3, const float *dirs3,
                          const float *radii, float *t, int *hit);
int scene_append(mpm_scene *s, uint32_t n, const float *pos3, const float *vel3);
int scene_export_device(mpm_scene *s, const mpm_export *dst, uint32_t flags);
int scene_frame_begin(mpm_scene *s, uint32_t what, uint64_t *ticket);
int scene_frame_poll(mpm_scene *s, uint64_t ticket, int wait, int *ready, mpm_frame *frame);
int scene_enable_timing(mpm_scene *s, bool on);
int scene_stage_times(mpm_scene *s, float *ms8);
int scene_tile_stats(mpm_scene *s, uint64_t out[4]);
void scene_drop_graph(mpm_scene *s);
int scene_step_times(mpm_scene *s, float *ms, uint32_t cap, uint32_t *count);
Params scene_params(const mpm_scene *s);
size_t scene_footprint(const mpm_config &c);
void scene_dev(const mpm_scene *s, int phase, SceneDev *d);
BrickWork bw_of(const mpm_scene *s, int c);
int batch_detach(mpm_scene *s);
}  // namespace mpm
