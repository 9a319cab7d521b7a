// mpm_capi.hip -- the C ABI of include/mlsmpm.h: scene objects, buffers and
// the per-step launch sequence (FluidGPU / DiffuseGPU host logic).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mlsmpm.h"
#include "mpm_kernels.h"
#include "mpm_scene.h"

using namespace mpm;

namespace {
thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
}  // namespace

namespace mpm {
int set_error(int code, const char *msg) { return fail(code, "%s", msg); }
int hip_fail(hipError_t e, const char *what) {
  return fail(MPM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}
}  // namespace mpm

#define HIPCHK(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

extern "C" {

const char *mpm_version(void) { return "mlsmpm-gfx950 0.1"; }
const char *mpm_last_error(void) { return g_err.c_str(); }

int mpm_config_default(mpm_config *c, uint32_t gx, uint32_t gy, uint32_t gz, uint32_t capacity) {
  if (!c) return fail(MPM_ERR_INVALID, "mpm_config_default: cfg is NULL");
  memset(c, 0, sizeof *c);
  c->grid_x = gx;
  c->grid_y = gy;
  c->grid_z = gz;
  c->capacity = capacity;
  c->diffuse_capacity = 0;
  c->sort_period = 1;
  c->flags = 0;
  c->device = 0;
  // FluidGPU constructor (fluid-gpu.js:24-45)
  c->dt = 0.3f;
  c->gravity = -0.98f;
  c->rest_density = 8.0f;
  c->stiffness = 100.0f;
  c->eos_power = 1.0f;
  c->viscosity = 0.01f;
  c->hard_min = 2.0f;
  c->hard_max_x = (float)gx - 2.0f;
  c->hard_max_y = (float)gy - 2.0f;
  c->hard_max_z = (float)gz - 2.0f;
  c->wall_min = 3.0f;
  c->wall_stiffness = 1.0f;
  c->lookahead_k = 2.0f;
  // DiffuseGPU tunables (diffuse-gpu.js:20-28)
  c->ke_threshold = 0.5f;
  c->pressure_threshold = 1.0f;
  c->crest_dot_threshold = 0.3f;
  c->spawn_rate_k = 3.0f;
  c->lifetime_foam = 3.0f;
  c->buoyancy = 1.2f;
  c->drag = 4.0f;
  c->density_spray_max = 1.5f;
  c->density_bubble_min = 6.0f;
  c->seed = 1;
  return MPM_OK;
}

int mpm_config_from_aspect(mpm_config *c, float ax, float ay, float az, float r,
                           uint32_t capacity) {
  if (!(r > 0.0f) || !(ax > 0.0f) || !(ay > 0.0f) || !(az > 0.0f))
    return fail(MPM_ERR_INVALID, "mpm_config_from_aspect: aspects and radius must be > 0");
  // fluid-gpu.js:11-14 computes in double precision.
  double normToGrid = 1.0 / ((double)r * 4.0);
  return mpm_config_default(c, (uint32_t)std::ceil((double)ax * normToGrid),
                            (uint32_t)std::ceil((double)ay * normToGrid),
                            (uint32_t)std::ceil((double)az * normToGrid), capacity);
}

int mpm_create(const mpm_config *cfg, mpm_scene **out) {
  if (!cfg || !out) return fail(MPM_ERR_INVALID, "mpm_create: NULL argument");
  *out = nullptr;
  int rc = validate_config(*cfg);
  if (rc) return rc;
  mpm_scene *s = new mpm_scene();
  rc = scene_init(s, *cfg);
  if (rc) {
    scene_free(s);
    delete s;
    return rc;
  }
  *out = s;
  return MPM_OK;
}

int mpm_destroy(mpm_scene *s) {
  if (!s) return MPM_OK;
  (void)batch_detach(s);
  scene_free(s);
  delete s;
  return MPM_OK;
}

int mpm_reset(mpm_scene *s, uint32_t n, const float *pos3, const float *vel3, const float *C9,
              const float *mass) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_reset: scene is NULL");
  return scene_reset(s, n, pos3, vel3, C9, mass, /*device_ptrs=*/false);
}

int mpm_step(mpm_scene *s, uint32_t nsteps) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_step: scene is NULL");
  return scene_step(s, nsteps);
}

int mpm_prepare(mpm_scene *s) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_prepare: scene is NULL");
  return scene_prepare(s);
}

int mpm_sync(mpm_scene *s) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_sync: scene is NULL");
  HIPCHK(hipSetDevice(s->cfg.device));
  HIPCHK(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

int mpm_get_counts(const mpm_scene *s, uint32_t *n, uint64_t *steps) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_get_counts: scene is NULL");
  if (n) *n = s->n;
  if (steps) *steps = s->steps;
  return MPM_OK;
}

int mpm_read_particles(mpm_scene *s, float *pos4, float *vel4, float *C9, float *density) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_read_particles: scene is NULL");
  return scene_read_particles(s, pos4, vel4, C9, density);
}

int mpm_read_grid(mpm_scene *s, int32_t *cell_mass, int32_t *cell_mv3, float *cell_mass_f,
                  float *cell_v3) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_read_grid: scene is NULL");
  return scene_read_grid(s, cell_mass, cell_mv3, cell_mass_f, cell_v3);
}

int mpm_read_diffuse(mpm_scene *s, void *rec48, uint32_t *free_list, int32_t *top) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_read_diffuse: scene is NULL");
  return scene_read_diffuse(s, rec48, free_list, top);
}

int mpm_write_diffuse(mpm_scene *s, const void *rec48, const uint32_t *free_list, int32_t top) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_write_diffuse: scene is NULL");
  return scene_write_diffuse(s, rec48, free_list, top);
}

int mpm_diffuse_alive(mpm_scene *s, uint32_t *alive) {
  if (!s || !alive) return fail(MPM_ERR_INVALID, "mpm_diffuse_alive: NULL argument");
  if (!s->cfg.diffuse_capacity) {
    *alive = 0;
    return MPM_OK;
  }
  int32_t top = 0;
  int rc = scene_read_diffuse(s, nullptr, nullptr, &top);
  if (rc) return rc;
  *alive = s->cfg.diffuse_capacity - (uint32_t)top;
  return MPM_OK;
}

int mpm_set_params(mpm_scene *s, const mpm_config *cfg) {
  if (!s || !cfg) return fail(MPM_ERR_INVALID, "mpm_set_params: NULL argument");
  const mpm_config &o = s->cfg;
  if (cfg->grid_x != o.grid_x || cfg->grid_y != o.grid_y || cfg->grid_z != o.grid_z ||
      cfg->capacity != o.capacity || cfg->diffuse_capacity != o.diffuse_capacity ||
      cfg->device != o.device || cfg->flags != o.flags || cfg->sort_period != o.sort_period ||
      cfg->tile_ppt != o.tile_ppt || cfg->tile_margin != o.tile_margin ||
      cfg->tile_threads != o.tile_threads || cfg->g2p_threads != o.g2p_threads)
    return fail(MPM_ERR_INVALID,
                "mpm_set_params: grid, capacities, device, flags, sort period and tile shape "
                "are fixed at create");
  int rc = validate_config(*cfg);
  if (rc) return rc;
  uint32_t seed = s->cfg.seed;
  s->cfg = *cfg;
  s->cfg.seed = seed;
  scene_drop_graph(s);
  s->version++;
  return MPM_OK;
}

int mpm_get_config(const mpm_scene *s, mpm_config *cfg) {
  if (!s || !cfg) return fail(MPM_ERR_INVALID, "mpm_get_config: NULL argument");
  *cfg = s->cfg;
  return MPM_OK;
}

int mpm_set_seed(mpm_scene *s, uint32_t seed) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_set_seed: scene is NULL");
  return scene_set_seed(s, seed);
}

int mpm_set_walls(mpm_scene *s, const float *wallp) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_set_walls: scene is NULL");
  return scene_set_walls(s, wallp);
}

int mpm_set_hand(mpm_scene *s, const mpm_hand *hand) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_set_hand: scene is NULL");
  return scene_set_hand(s, hand);
}

int mpm_delete_near(mpm_scene *s, const float origin[3], const float dir[3], float radius,
                    uint32_t *remaining) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_delete_near: scene is NULL");
  return scene_delete_near(s, origin, dir, radius, remaining);
}

int mpm_raycast(mpm_scene *s, const float origin[3], const float dir[3], float radius, float *t,
                int *hit) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_raycast: scene is NULL");
  return scene_raycast(s, origin, dir, radius, t, hit);
}

int mpm_raycast_noalloc(mpm_scene *s, uint32_t nrays, const float *origins3, const float *dirs3,
                        const float *radii, float *t, int *hit) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_raycast_noalloc: scene is NULL");
  return scene_raycast_noalloc(s, nrays, origins3, dirs3, radii, t, hit);
}

int mpm_append_particles(mpm_scene *s, uint32_t n, const float *pos3, const float *vel3) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_append_particles: scene is NULL");
  return scene_append(s, n, pos3, vel3);
}

int mpm_export_device(mpm_scene *s, const mpm_export *dst, uint32_t flags) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_export_device: scene is NULL");
  return scene_export_device(s, dst, flags);
}

int mpm_frame_begin(mpm_scene *s, uint32_t what, uint64_t *ticket) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_frame_begin: scene is NULL");
  return scene_frame_begin(s, what, ticket);
}

int mpm_frame_poll(mpm_scene *s, uint64_t ticket, int wait, int *ready, mpm_frame *frame) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_frame_poll: scene is NULL");
  return scene_frame_poll(s, ticket, wait, ready, frame);
}

int mpm_get_stream(mpm_scene *s, void **st) {
  if (!s || !st) return fail(MPM_ERR_INVALID, "mpm_get_stream: NULL argument");
  *st = (void *)s->stream;
  return MPM_OK;
}

int mpm_enable_timing(mpm_scene *s, int enable) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_enable_timing: scene is NULL");
  return scene_enable_timing(s, enable != 0);
}

int mpm_stage_times(mpm_scene *s, float *ms8) {
  if (!s || !ms8) return fail(MPM_ERR_INVALID, "mpm_stage_times: NULL argument");
  return scene_stage_times(s, ms8);
}

int mpm_step_times(mpm_scene *s, float *ms, uint32_t cap, uint32_t *count) {
  if (!s) return fail(MPM_ERR_INVALID, "mpm_step_times: scene is NULL");
  return scene_step_times(s, ms, cap, count);
}

int mpm_tile_stats(mpm_scene *s, uint64_t out[4]) {
  if (!s || !out) return fail(MPM_ERR_INVALID, "mpm_tile_stats: NULL argument");
  return scene_tile_stats(s, out);
}

}  // extern "C"
