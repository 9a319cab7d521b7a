/*
 * mlsmpm_napi.c -- Node-API (v8) addon: thin marshalling of JavaScript values
 * and TypedArrays onto the C ABI of include/mlsmpm.h.  The JavaScript classes
 * FluidHIP / DiffuseHIP (../fluid-hip.mjs, ../diffuse-hip.mjs) sit on top and
 * keep the method and field names of the reference's FluidGPU / DiffuseGPU.
 *
 * Every function throws a JavaScript Error (with .code = the MPM_ERR_* value)
 * when the C call fails; there is no fallback path.
 */
#define NAPI_VERSION 8
#include <node_api.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mlsmpm.h"

#define NAPI_CALL(env, call)                                             \
  do {                                                                   \
    if ((call) != napi_ok) {                                             \
      napi_throw_error((env), NULL, "N-API call failed: " #call);         \
      return NULL;                                                       \
    }                                                                    \
  } while (0)

static napi_value throw_mpm(napi_env env, int rc) {
  char code[16];
  snprintf(code, sizeof code, "%d", rc);
  napi_value msg, err, c;
  const char *m = mpm_last_error();
  napi_create_string_utf8(env, m && *m ? m : "mlsmpm call failed", NAPI_AUTO_LENGTH, &msg);
  napi_create_string_utf8(env, code, NAPI_AUTO_LENGTH, &c);
  napi_create_error(env, c, msg, &err);
  napi_throw(env, err);
  return NULL;
}

#define MPM_CALL(env, call)               \
  do {                                    \
    int rc_ = (call);                     \
    if (rc_ != MPM_OK) return throw_mpm((env), rc_); \
  } while (0)

/* ---- argument helpers ---------------------------------------------------- */

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  for (size_t i = argc; i < want; i++) napi_get_undefined(env, &argv[i]);
  return 1;
}

static int is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_null || t == napi_undefined;
}

/* Data pointer and element count of a Float32Array (NULL for null/undefined). */
static int f32_array(napi_env env, napi_value v, float **data, size_t *len) {
  *data = NULL;
  *len = 0;
  if (is_nullish(env, v)) return 1;
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_ta) return 0;
  napi_typedarray_type type;
  napi_value ab;
  size_t off;
  void *p;
  napi_get_typedarray_info(env, v, &type, len, &p, &ab, &off);
  if (type != napi_float32_array) return 0;
  *data = (float *)p;
  return 1;
}

/* Data pointer and byte length of an ArrayBuffer or any TypedArray. */
static int bytes_of(napi_env env, napi_value v, void **data, size_t *bytes) {
  *data = NULL;
  *bytes = 0;
  if (is_nullish(env, v)) return 1;
  bool is_ab = false, is_ta = false;
  napi_is_arraybuffer(env, v, &is_ab);
  if (is_ab) return napi_get_arraybuffer_info(env, v, data, bytes) == napi_ok;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_ta) return 0;
  napi_typedarray_type type;
  napi_value ab;
  size_t off, len;
  napi_get_typedarray_info(env, v, &type, &len, data, &ab, &off);
  size_t esz = 1;
  switch (type) {
    case napi_int16_array: case napi_uint16_array: esz = 2; break;
    case napi_int32_array: case napi_uint32_array: case napi_float32_array: esz = 4; break;
    case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: esz = 8; break;
    default: esz = 1;
  }
  *bytes = len * esz;
  return 1;
}

static mpm_scene *scene_of(napi_env env, napi_value v) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected a scene handle");
    return NULL;
  }
  return *(mpm_scene **)p;
}

static uint32_t u32_of(napi_env env, napi_value v) {
  uint32_t x = 0;
  napi_get_value_uint32(env, v, &x);
  return x;
}

static double f64_of(napi_env env, napi_value v) {
  double x = 0;
  napi_get_value_double(env, v, &x);
  return x;
}

/* ---- mpm_config <-> plain object ------------------------------------------ */

typedef struct { const char *name; size_t off; int is_float; } field_t;
#define FU(n) {#n, offsetof(mpm_config, n), 0}
#define FF(n) {#n, offsetof(mpm_config, n), 1}
static const field_t kFields[] = {
    FU(grid_x), FU(grid_y), FU(grid_z), FU(capacity), FU(diffuse_capacity), FU(sort_period),
    FU(flags), FU(device), FF(dt), FF(gravity), FF(rest_density), FF(stiffness), FF(eos_power),
    FF(viscosity), FF(hard_min), FF(hard_max_x), FF(hard_max_y), FF(hard_max_z), FF(wall_min),
    FF(wall_stiffness), FF(lookahead_k), FF(ke_threshold), FF(pressure_threshold),
    FF(crest_dot_threshold), FF(spawn_rate_k), FF(lifetime_foam), FF(buoyancy), FF(drag),
    FF(density_spray_max), FF(density_bubble_min), FU(seed), FU(tile_ppt), FU(tile_margin),
    FU(tile_threads), FU(g2p_threads), FU(lifetime_mode),
};

static napi_value config_to_js(napi_env env, const mpm_config *c) {
  napi_value o, v;
  NAPI_CALL(env, napi_create_object(env, &o));
  for (size_t i = 0; i < sizeof kFields / sizeof kFields[0]; i++) {
    const char *p = (const char *)c + kFields[i].off;
    if (kFields[i].is_float) napi_create_double(env, *(const float *)p, &v);
    else if (!strcmp(kFields[i].name, "device")) napi_create_int32(env, *(const int32_t *)p, &v);
    else napi_create_uint32(env, *(const uint32_t *)p, &v);
    napi_set_named_property(env, o, kFields[i].name, v);
  }
  return o;
}

/* Fields present on the object override *c. */
static void config_from_js(napi_env env, napi_value o, mpm_config *c) {
  for (size_t i = 0; i < sizeof kFields / sizeof kFields[0]; i++) {
    bool has = false;
    napi_has_named_property(env, o, kFields[i].name, &has);
    if (!has) continue;
    napi_value v;
    napi_get_named_property(env, o, kFields[i].name, &v);
    if (is_nullish(env, v)) continue;
    char *p = (char *)c + kFields[i].off;
    if (kFields[i].is_float) *(float *)p = (float)f64_of(env, v);
    else if (!strcmp(kFields[i].name, "device")) napi_get_value_int32(env, v, (int32_t *)p);
    else *(uint32_t *)p = u32_of(env, v);
  }
}

/* ---- exported functions ---------------------------------------------------- */

/* configDefault(gx, gy, gz, capacity) -> config object (fluid-gpu.js:24-45) */
static napi_value js_config_default(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!get_args(env, info, 4, a)) return NULL;
  mpm_config c;
  MPM_CALL(env, mpm_config_default(&c, u32_of(env, a[0]), u32_of(env, a[1]), u32_of(env, a[2]),
                                   u32_of(env, a[3])));
  return config_to_js(env, &c);
}

/* configFromAspect(ax, ay, az, radius, capacity) -> config (fluid-gpu.js:10-15) */
static napi_value js_config_from_aspect(napi_env env, napi_callback_info info) {
  napi_value a[5];
  if (!get_args(env, info, 5, a)) return NULL;
  mpm_config c;
  MPM_CALL(env, mpm_config_from_aspect(&c, (float)f64_of(env, a[0]), (float)f64_of(env, a[1]),
                                       (float)f64_of(env, a[2]), (float)f64_of(env, a[3]),
                                       u32_of(env, a[4])));
  return config_to_js(env, &c);
}

static void finalize_scene(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  mpm_scene **h = (mpm_scene **)data;
  if (*h) mpm_destroy(*h);
  free(h);
}

/* create(config) -> scene handle (FluidGPU.initGPU + DiffuseGPU.initGPU) */
static napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return NULL;
  mpm_config c;
  memset(&c, 0, sizeof c);
  config_from_js(env, a[0], &c);
  mpm_scene *s = NULL;
  MPM_CALL(env, mpm_create(&c, &s));
  mpm_scene **h = (mpm_scene **)malloc(sizeof *h);
  *h = s;
  napi_value ext;
  if (napi_create_external(env, h, finalize_scene, NULL, &ext) != napi_ok) {
    mpm_destroy(s);
    free(h);
    napi_throw_error(env, NULL, "napi_create_external failed");
    return NULL;
  }
  return ext;
}

/* destroy(handle): frees now (the finalizer then does nothing) */
static napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return NULL;
  void *p = NULL;
  if (napi_get_value_external(env, a[0], &p) == napi_ok && p) {
    mpm_scene **h = (mpm_scene **)p;
    if (*h) mpm_destroy(*h);
    *h = NULL;
  }
  return NULL;
}

/* reset(handle, pos3 Float32Array, vel3?, C9?, mass?) (fillBlock + upload) */
static napi_value js_reset(napi_env env, napi_callback_info info) {
  napi_value a[5];
  if (!get_args(env, info, 5, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  float *pos, *vel, *C9, *mass;
  size_t np, nv, nc, nm;
  if (!f32_array(env, a[1], &pos, &np) || !f32_array(env, a[2], &vel, &nv) ||
      !f32_array(env, a[3], &C9, &nc) || !f32_array(env, a[4], &mass, &nm)) {
    napi_throw_type_error(env, NULL, "reset: arrays must be Float32Array or null");
    return NULL;
  }
  if (!pos || np % 3) {
    napi_throw_type_error(env, NULL, "reset: pos3 must hold 3 floats per particle");
    return NULL;
  }
  size_t n = np / 3;
  if ((vel && nv != 3 * n) || (C9 && nc != 9 * n) || (mass && nm != n)) {
    napi_throw_range_error(env, NULL, "reset: vel3 / C9 / mass lengths do not match pos3");
    return NULL;
  }
  MPM_CALL(env, mpm_reset(s, (uint32_t)n, pos, vel, C9, mass));
  return NULL;
}

/* step(handle, nsteps) (simFrame) */
static napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  MPM_CALL(env, mpm_step(s, u32_of(env, a[1])));
  return NULL;
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  MPM_CALL(env, mpm_sync(s));
  return NULL;
}

/* counts(handle) -> {n, steps} (active_particle_num) */
static napi_value js_counts(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  uint32_t n = 0;
  uint64_t steps = 0;
  MPM_CALL(env, mpm_get_counts(s, &n, &steps));
  napi_value o, v;
  napi_create_object(env, &o);
  napi_create_uint32(env, n, &v);
  napi_set_named_property(env, o, "n", v);
  napi_create_double(env, (double)steps, &v);
  napi_set_named_property(env, o, "steps", v);
  return o;
}

/* readParticles(handle, pos4?, vel4?, C9?, density?) into Float32Arrays */
static napi_value js_read_particles(napi_env env, napi_callback_info info) {
  napi_value a[5];
  if (!get_args(env, info, 5, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  uint32_t n = 0;
  MPM_CALL(env, mpm_get_counts(s, &n, NULL));
  float *p[4];
  size_t len[4];
  const size_t want[4] = {4, 4, 9, 1};
  for (int i = 0; i < 4; i++) {
    if (!f32_array(env, a[1 + i], &p[i], &len[i])) {
      napi_throw_type_error(env, NULL, "readParticles: outputs must be Float32Array or null");
      return NULL;
    }
    if (p[i] && len[i] < want[i] * (size_t)n) {
      napi_throw_range_error(env, NULL, "readParticles: output array too short");
      return NULL;
    }
  }
  MPM_CALL(env, mpm_read_particles(s, p[0], p[1], p[2], p[3]));
  return NULL;
}

/* readDiffuse(handle, rec48 ArrayBuffer?) -> {top, alive} (particleBuffer) */
static napi_value js_read_diffuse(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  mpm_config c;
  MPM_CALL(env, mpm_get_config(s, &c));
  void *rec;
  size_t bytes;
  if (!bytes_of(env, a[1], &rec, &bytes)) {
    napi_throw_type_error(env, NULL, "readDiffuse: output must be an ArrayBuffer / TypedArray");
    return NULL;
  }
  if (rec && bytes < (size_t)c.diffuse_capacity * 48) {
    napi_throw_range_error(env, NULL, "readDiffuse: output shorter than 48 * maxCount bytes");
    return NULL;
  }
  int32_t top = 0;
  if (c.diffuse_capacity) MPM_CALL(env, mpm_read_diffuse(s, rec, NULL, &top));
  napi_value o, v;
  napi_create_object(env, &o);
  napi_create_int32(env, top, &v);
  napi_set_named_property(env, o, "top", v);
  napi_create_uint32(env, c.diffuse_capacity - (uint32_t)top, &v);
  napi_set_named_property(env, o, "alive", v);
  return o;
}

/* setParams(handle, partialConfig) (assignments to the public tunables) */
static napi_value js_set_params(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  mpm_config c;
  MPM_CALL(env, mpm_get_config(s, &c));
  config_from_js(env, a[1], &c);
  MPM_CALL(env, mpm_set_params(s, &c));
  return NULL;
}

static napi_value js_get_config(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  mpm_config c;
  MPM_CALL(env, mpm_get_config(s, &c));
  return config_to_js(env, &c);
}

static napi_value js_set_seed(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  MPM_CALL(env, mpm_set_seed(s, u32_of(env, a[1])));
  return NULL;
}

/* setWalls(handle, Float32Array(84) | null) (WallParams, shaders.js:123-133) */
static napi_value js_set_walls(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  float *w;
  size_t len;
  if (!f32_array(env, a[1], &w, &len) || (w && len < 84)) {
    napi_throw_type_error(env, NULL, "setWalls: expected Float32Array(84) or null");
    return NULL;
  }
  MPM_CALL(env, mpm_set_walls(s, w));
  return NULL;
}

/* Three numbers from a JavaScript array or TypedArray. */
static int vec3_of(napi_env env, napi_value v, float out[3]) {
  bool is_arr = false, is_ta = false;
  napi_is_array(env, v, &is_arr);
  napi_is_typedarray(env, v, &is_ta);
  if (!is_arr && !is_ta) return 0;
  for (uint32_t i = 0; i < 3; i++) {
    napi_value e;
    double x;
    if (napi_get_element(env, v, i, &e) != napi_ok) return 0;
    if (napi_get_value_double(env, e, &x) != napi_ok) return 0;
    out[i] = (float)x;
  }
  return 1;
}

/* setHand(handle, pos, vel, radius, strength, active, eye) (updateHand,
 * fluid-gpu.js:479-529); setHand(handle, null) switches the push off. */
static napi_value js_set_hand(napi_env env, napi_callback_info info) {
  napi_value a[7];
  if (!get_args(env, info, 7, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  if (is_nullish(env, a[1])) {
    MPM_CALL(env, mpm_set_hand(s, NULL));
    return NULL;
  }
  mpm_hand h;
  memset(&h, 0, sizeof h);
  if (!vec3_of(env, a[1], h.pos) || !vec3_of(env, a[2], h.vel) || !vec3_of(env, a[6], h.eye)) {
    napi_throw_type_error(env, NULL, "setHand: pos, vel and eye must hold three numbers");
    return NULL;
  }
  h.radius = (float)f64_of(env, a[3]);
  h.strength = (float)f64_of(env, a[4]);
  bool on = false;
  napi_coerce_to_bool(env, a[5], &a[5]);
  napi_get_value_bool(env, a[5], &on);
  h.active = on ? 1u : 0u;
  MPM_CALL(env, mpm_set_hand(s, &h));
  return NULL;
}

/* deleteNear(handle, origin, dir, radius) -> particles left (fluid-gpu.js:547-575) */
static napi_value js_delete_near(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!get_args(env, info, 4, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  float o[3], d[3];
  if (!vec3_of(env, a[1], o) || !vec3_of(env, a[2], d)) {
    napi_throw_type_error(env, NULL, "deleteNear: origin and dir must hold three numbers");
    return NULL;
  }
  uint32_t left = 0;
  MPM_CALL(env, mpm_delete_near(s, o, d, (float)f64_of(env, a[3]), &left));
  napi_value v;
  napi_create_uint32(env, left, &v);
  return v;
}

/* raycast(handle, origin, dir, radius) -> t or null (raycastFluid + pollRaycast,
 * fluid-gpu.js:597-638) */
static napi_value js_raycast(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!get_args(env, info, 4, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  float o[3], d[3];
  if (!vec3_of(env, a[1], o) || !vec3_of(env, a[2], d)) {
    napi_throw_type_error(env, NULL, "raycast: origin and dir must hold three numbers");
    return NULL;
  }
  float t = 0.0f;
  int hit = 0;
  MPM_CALL(env, mpm_raycast(s, o, d, (float)f64_of(env, a[3]), &t, &hit));
  napi_value v;
  if (hit) napi_create_double(env, (double)t, &v);
  else napi_get_null(env, &v);
  return v;
}

/* append(handle, pos Float32Array(3k), vel Float32Array(3k) | null) -> k
 * (spawnParticles' upload, fluid-gpu.js:145-180) */
static napi_value js_append(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  float *pos, *vel;
  size_t np, nv;
  if (!f32_array(env, a[1], &pos, &np) || !f32_array(env, a[2], &vel, &nv) || !pos ||
      np % 3 || (vel && nv < np)) {
    napi_throw_type_error(env, NULL, "append: expected Float32Array(3k) pos and vel (or null)");
    return NULL;
  }
  MPM_CALL(env, mpm_append_particles(s, (uint32_t)(np / 3), pos, vel));
  napi_value v;
  napi_create_uint32(env, (uint32_t)(np / 3), &v);
  return v;
}

/* frameBegin(handle, what) -> ticket (renderer handoff through pinned memory;
 * the bindings of fluid-renderer.js:315-316, 850-856) */
static napi_value js_frame_begin(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  uint64_t t = 0;
  MPM_CALL(env, mpm_frame_begin(s, u32_of(env, a[1]), &t));
  napi_value v;
  napi_create_double(env, (double)t, &v);
  return v;
}

/* framePoll(handle, ticket, wait, pos4?, vel4?, density?, diffuse?) ->
 * null while the copy runs, else {n, top, steps}; the parts the frame holds
 * are copied into the outputs given (Float32Array n*4 / n*4 / n, ArrayBuffer
 * or TypedArray of 48 * maxCount bytes). */
static napi_value js_frame_poll(napi_env env, napi_callback_info info) {
  napi_value a[7];
  if (!get_args(env, info, 7, a)) return NULL;
  mpm_scene *s = scene_of(env, a[0]);
  if (!s) return NULL;
  const uint64_t ticket = (uint64_t)f64_of(env, a[1]);
  bool wait = false;
  napi_coerce_to_bool(env, a[2], &a[2]);
  napi_get_value_bool(env, a[2], &wait);
  int ready = 0;
  mpm_frame f;
  MPM_CALL(env, mpm_frame_poll(s, ticket, wait ? 1 : 0, &ready, &f));
  napi_value v;
  if (!ready) {
    napi_get_null(env, &v);
    return v;
  }
  const void *src[4] = {f.pos4, f.vel4, f.density, f.diffuse48};
  mpm_config c;
  MPM_CALL(env, mpm_get_config(s, &c));
  const size_t need[4] = {(size_t)f.n * 16, (size_t)f.n * 16, (size_t)f.n * 4,
                          (size_t)c.diffuse_capacity * 48};
  for (int i = 0; i < 4; i++) {
    void *dst;
    size_t bytes;
    if (!bytes_of(env, a[3 + i], &dst, &bytes)) {
      napi_throw_type_error(env, NULL, "framePoll: outputs must be TypedArray / ArrayBuffer or null");
      return NULL;
    }
    if (!dst || !src[i]) continue;
    if (bytes < need[i]) {
      napi_throw_range_error(env, NULL, "framePoll: output too short for the frame");
      return NULL;
    }
    memcpy(dst, src[i], need[i]);
  }
  napi_create_object(env, &v);
  napi_value x;
  napi_create_uint32(env, f.n, &x);
  napi_set_named_property(env, v, "n", x);
  napi_create_int32(env, f.top, &x);
  napi_set_named_property(env, v, "top", x);
  napi_create_double(env, (double)f.steps, &x);
  napi_set_named_property(env, v, "steps", x);
  return v;
}

static napi_value js_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value v;
  napi_create_string_utf8(env, mpm_version(), NAPI_AUTO_LENGTH, &v);
  return v;
}

static napi_value init(napi_env env, napi_value exports) {
  const struct { const char *name; napi_callback fn; } fns[] = {
      {"version", js_version},
      {"configDefault", js_config_default},
      {"configFromAspect", js_config_from_aspect},
      {"create", js_create},
      {"destroy", js_destroy},
      {"reset", js_reset},
      {"step", js_step},
      {"sync", js_sync},
      {"counts", js_counts},
      {"readParticles", js_read_particles},
      {"readDiffuse", js_read_diffuse},
      {"setParams", js_set_params},
      {"getConfig", js_get_config},
      {"setSeed", js_set_seed},
      {"setWalls", js_set_walls},
      {"setHand", js_set_hand},
      {"deleteNear", js_delete_near},
      {"raycast", js_raycast},
      {"append", js_append},
      {"frameBegin", js_frame_begin},
      {"framePoll", js_frame_poll},
  };
  for (size_t i = 0; i < sizeof fns / sizeof fns[0]; i++) {
    napi_value f;
    NAPI_CALL(env, napi_create_function(env, fns[i].name, NAPI_AUTO_LENGTH, fns[i].fn, NULL, &f));
    NAPI_CALL(env, napi_set_named_property(env, exports, fns[i].name, f));
  }
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
