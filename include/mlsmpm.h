/*
 * mlsmpm.h -- C ABI of the MI355X (gfx950) MLS-MPM fluid step.
 *
 * The reference drives its fluid through two JavaScript host classes whose
 * methods record WebGPU dispatches; this header is the flat C surface a
 * Node-API addon (or ctypes, cgo, JNI) binds in their place.  Each entry point
 * names the reference interface it replaces.
 *
 * Conventions
 *   - Every call returns MPM_OK (0) or a negative MPM_ERR_* code; the message
 *     of the last failure on the calling thread is mpm_last_error().
 *   - Host arrays passed in are copied before the call returns; host arrays
 *     passed out are filled before the call returns (the call synchronises
 *     the scene's stream).  mpm_step only enqueues work.
 *   - One scene = one HIP stream.  A handle is not thread-safe; distinct
 *     handles may be driven from distinct threads.
 *   - Layouts handed out are the reference's renderer layouts: positions and
 *     velocities as stride-16 vec4 (w = 0), the affine matrix C as 9 floats
 *     row-major per particle, diffuse particles as 48-byte records.
 */
#ifndef MLSMPM_H
#define MLSMPM_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_OK 0
#define MPM_ERR_INVALID -1     /* bad argument or configuration */
#define MPM_ERR_RANGE -2       /* particle outside the grid or not finite */
#define MPM_ERR_CAPACITY -3    /* more particles than the scene holds */
#define MPM_ERR_HIP -4         /* HIP runtime failure (message has details) */
#define MPM_ERR_NOMEM -5       /* device memory or HBM budget exhausted */
#define MPM_ERR_UNSUPPORTED -6 /* option this build does not implement */
#define MPM_ERR_STATE -7       /* call not valid in the object's state */
#define MPM_ERR_NOT_FOUND -8   /* unknown id */

/* mpm_config.flags */
#define MPM_FLAG_KEEP_GRID 1u  /* keep the i32 grids readable after a step
                                  (separate clear pass instead of the fused one) */
#define MPM_FLAG_NO_LDS 2u     /* force the global-atomic scatter path */
#define MPM_FLAG_NO_GRAPH 4u  /* enqueue every kernel eagerly (no captured HIP graph) */

/* Scene configuration.  Physics fields and defaults follow the FluidGPU
 * constructor (fluid-gpu.js:10-45) and the Params uniform (shaders.js:27-49);
 * diffuse fields follow DiffuseGPU (diffuse-gpu.js:16-31). */
typedef struct mpm_config {
    uint32_t grid_x, grid_y, grid_z; /* grid_X_num.. (fluid-gpu.js:11-14) */
    uint32_t capacity;               /* particleNum (fluid-gpu.js:16) */
    uint32_t diffuse_capacity;       /* DiffuseGPU maxCount; 0 = none attached */
    uint32_t sort_period;            /* re-sort particles by cell every k steps; 0 = never */
    uint32_t flags;                  /* MPM_FLAG_* */
    int32_t device;                  /* HIP device ordinal */
    float dt, gravity, rest_density, stiffness, eos_power, viscosity;
    float hard_min, hard_max_x, hard_max_y, hard_max_z;
    float wall_min, wall_stiffness, lookahead_k;
    float ke_threshold, pressure_threshold, crest_dot_threshold, spawn_rate_k;
    float lifetime_foam, buoyancy, drag, density_spray_max, density_bubble_min;
    uint32_t seed; /* DiffuseGPU._seed initial value; advanced before each generate */
    uint32_t tile_ppt;     /* work-item size in units of 256 particles (4/8/16); 0 = 8 */
    uint32_t tile_margin;  /* P2G LDS window drift margin in cells (1/2); 0 = 2 */
    uint32_t tile_threads; /* P2G threads per work item (256); 0 = 256 */
    uint32_t g2p_threads;  /* G2P threads per work item (256); 0 = 256 */
    uint32_t lifetime_mode; /* foam lifetime: 0 = the reference's LIFETIME_FOAM * min(1,
                               score / KE_THRESHOLD); 1 = uniform in [0.5, 3] s, seeded
                               (0.5 + 2.5 * diffuse_rand(spawn seed + 4)) */
    uint32_t reserved[1];
} mpm_config;

typedef struct mpm_scene mpm_scene;

/* Version string of this library. */
const char *mpm_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char *mpm_last_error(void);

/* Fill cfg with the reference defaults for an explicit grid: DT 0.3,
 * GRAVITY -0.98, REST_DENSITY 8, STIFFNESS 100, EOS_POWER 1, VISCOSITY 0.01,
 * HARD_MIN 2, HARD_MAX_* = grid - 2, WALL_MIN 3, WALL_STIFFNESS 1,
 * LOOKAHEAD_K 2 (fluid-gpu.js:24-45) and the DiffuseGPU tunables
 * (diffuse-gpu.js:20-28), seed 1, sort_period 1 (the re-sort rides on G2P). */
int mpm_config_default(mpm_config *cfg, uint32_t gx, uint32_t gy, uint32_t gz,
                       uint32_t capacity);

/* Replaces `new FluidGPU(aspectX, aspectY, aspectZ, particleRadius,
 * particleNum)` (fluid-gpu.js:10-15): grid = ceil(aspect / (4 r)). */
int mpm_config_from_aspect(mpm_config *cfg, float aspect_x, float aspect_y, float aspect_z,
                           float particle_radius, uint32_t capacity);

/* Replaces FluidGPU.initGPU + DiffuseGPU.initGPU (fluid-gpu.js:183-187,
 * diffuse-gpu.js:39-44): allocates every device buffer of the scene.  The
 * diffuse pool starts all-dead with free list [0..D-1], top = D
 * (diffuse-gpu.js:60-68). */
int mpm_create(const mpm_config *cfg, mpm_scene **out);

/* Frees the scene and its device memory. */
int mpm_destroy(mpm_scene *s);

/* Replaces fillBlock + the initial buffer upload (fluid-gpu.js:107-128,
 * 278-281): load n particles.  pos3 is n*3 floats (required), vel3 n*3,
 * C9 n*9, mass n (NULL = zero, zero, 1.0).  Every coordinate must be finite
 * and in [1, grid - 1) so the 3x3x3 stencil stays inside the grid; otherwise
 * MPM_ERR_RANGE and the scene is unchanged.  On the brick path (a sort
 * period, no MPM_FLAG_NO_LDS) the particles are also sorted into bricks
 * here, so the first step is an ordinary one; reads return them in the
 * caller's order whatever the layout. */
int mpm_reset(mpm_scene *s, uint32_t n, const float *pos3, const float *vel3, const float *C9,
              const float *mass);

/* Replaces FluidGPU.simFrame (fluid-gpu.js:648-684), SUBSTEPS = 1, including
 * DiffuseGPU.step (diffuse-gpu.js:148-166) when diffuse_capacity > 0:
 * enqueue nsteps whole steps on the scene's stream. */
int mpm_step(mpm_scene *s, uint32_t nsteps);

/* No counterpart in the reference (WebGPU compiles its pipelines at
 * creation): capture the HIP graphs that the next multi-step mpm_step will
 * replay, without running a step, so that call pays no capture.  Where the
 * next step must first sort the particles (after mpm_append_particles or
 * mpm_delete_near) they are the graphs of the state that step leaves.  A
 * no-op where steps run eagerly (MPM_FLAG_NO_GRAPH, sort period 0, no
 * particles, batch members). */
int mpm_prepare(mpm_scene *s);

/* Wait for all work enqueued on the scene's stream. */
int mpm_sync(mpm_scene *s);

/* active_particle_num (fluid-gpu.js:17) and steps taken since mpm_reset. */
int mpm_get_counts(const mpm_scene *s, uint32_t *n, uint64_t *steps);

/* Read particles in reset order, in the layouts the renderer binds
 * (particlePosBuffer / particleVelBuffer stride 16, particleAffineBuffer,
 * particleDensityBuffer: fluid-gpu.js:208-212, fluid-renderer.js:315-316).
 * Any output may be NULL.  pos4/vel4: n*4 floats, C9: n*9, density: n. */
int mpm_read_particles(mpm_scene *s, float *pos4, float *vel4, float *C9, float *density);

/* Read the grids as the reference leaves them after a step (cellMassBuffer,
 * cellMvBuffer interleaved idx*3+c, cellMassF32Buffer, cellMvF32Buffer;
 * fluid-gpu.js:216-219).  The i32 grids are only available with
 * MPM_FLAG_KEEP_GRID (MPM_ERR_STATE otherwise).  Any output may be NULL. */
int mpm_read_grid(mpm_scene *s, int32_t *cell_mass, int32_t *cell_mv3, float *cell_mass_f,
                  float *cell_v3);

/* Read the diffuse pool: D records of 48 bytes {pos vec4, vel vec4, ptype,
 * alive, lifetime, pad} (shaders.js:725-732; DiffuseGPU.particleBuffer),
 * the free-list stack and its top.  Any output may be NULL. */
int mpm_read_diffuse(mpm_scene *s, void *rec48, uint32_t *free_list, int32_t *top);

/* ---- Renderer handoff ------------------------------------------------------
 * The reference's renderer binds the simulation buffers directly, without a
 * read-back: particlePosBuffer and particleVelBuffer as stride-16 instance
 * vertex buffers (fluid-renderer.js:315-316), particleDensityBuffer as
 * storage (fluid-renderer.js:232, 256) and DiffuseGPU.particleBuffer as a
 * stride-48 vertex buffer over all maxCount slots, dead slots culled by
 * their alive flag (fluid-renderer.js:850-856).  Two handoffs replace that
 * binding: a device-to-device export into buffers the caller owns (a
 * renderer on the same GPU), and a double-buffered export into pinned host
 * memory whose copy overlaps the following steps (a renderer elsewhere; the
 * mapAsync pattern of pollDelete / pollDebug). */
#define MPM_EXPORT_POS 1u           /* n float4 (x, y, z, 0) */
#define MPM_EXPORT_VEL 2u           /* n float4 (vx, vy, vz, 0) */
#define MPM_EXPORT_DENSITY 4u       /* n floats */
#define MPM_EXPORT_DIFFUSE 8u       /* diffuse_capacity 48-byte records */
#define MPM_EXPORT_ALL 15u
/* Particles in reset order (as mpm_read_particles).  Without it they come in
 * the scene's current storage order: the same particles, pos, vel and
 * density row for row, read and written contiguously (what a renderer
 * drawing instances needs). */
#define MPM_EXPORT_RESET_ORDER 16u

typedef struct mpm_export {
  float *pos4;     /* device pointers, NULL to skip */
  float *vel4;
  float *density;
  void *diffuse48;
} mpm_export;

/* Enqueue on the scene's stream (mpm_get_stream) the renderer layouts of the
 * current state into the caller's device buffers: n = active particles at
 * enqueue time (mpm_get_counts).  flags: 0 or MPM_EXPORT_RESET_ORDER.
 * Returns at once; order later work after it with the scene's stream or
 * mpm_sync. */
int mpm_export_device(mpm_scene *s, const mpm_export *dst, uint32_t flags);

/* One exported frame in pinned host memory. */
typedef struct mpm_frame {
  const float *pos4;        /* NULL unless asked for */
  const float *vel4;
  const float *density;
  const void *diffuse48;
  uint32_t n;               /* particles in the frame */
  int32_t top;              /* diffuse free-list top (alive = capacity - top); 0 without a pool */
  uint64_t steps;           /* steps since reset when the frame was taken */
  uint64_t ticket;
} mpm_frame;

/* Take a frame of the current state (what: MPM_EXPORT_* bits, plus
 * MPM_EXPORT_RESET_ORDER): the layouts are written to device staging on the
 * scene's stream and copied to pinned host memory on a copy stream, so the
 * steps enqueued after this call run while the copy does.  Two frames are
 * kept: a frame's memory is reused by the second mpm_frame_begin after it. */
int mpm_frame_begin(mpm_scene *s, uint32_t what, uint64_t *ticket);

/* *ready = 1 and *frame filled once the frame's copy has finished (wait != 0
 * blocks until then); *ready = 0 otherwise.  MPM_ERR_NOT_FOUND when the
 * ticket's memory has been reused.  The pointers stay valid until then. */
int mpm_frame_poll(mpm_scene *s, uint64_t ticket, int wait, int *ready, mpm_frame *frame);

/* Overwrite the diffuse pool, free list and top (checkpoint restore, tests). */
int mpm_write_diffuse(mpm_scene *s, const void *rec48, const uint32_t *free_list, int32_t top);

/* Replaces DiffuseGPU.requestDebugReadback/pollDebug (diffuse-gpu.js:172-188):
 * alive = maxCount - top. */
int mpm_diffuse_alive(mpm_scene *s, uint32_t *alive);

/* Replaces assignments to the public tunables of FluidGPU / DiffuseGPU
 * (fluid-gpu.js:24-45, diffuse-gpu.js:20-28): physics and diffuse fields of
 * cfg take effect from the next step.  Grid size, capacities, device, flags,
 * sort period and tile shape are fixed at create (MPM_ERR_INVALID if they
 * differ). */
int mpm_set_params(mpm_scene *s, const mpm_config *cfg);
int mpm_get_config(const mpm_scene *s, mpm_config *cfg);

/* Set DiffuseGPU._seed (diffuse-gpu.js:31, 149); the next generate uses seed+1. */
int mpm_set_seed(mpm_scene *s, uint32_t seed);

/* Replaces FluidGPU._writeWallParams's upload (fluid-gpu.js:337-413): the
 * WallParams uniform, 4 floats header (count in [0]) + 4 walls x 20 floats
 * (shaders.js:123-133).  NULL clears all walls. */
int mpm_set_walls(mpm_scene *s, const float *wallp84);

/* ---- Interaction ---------------------------------------------------------
 * The hand push of FluidGPU.updateHand (fluid-gpu.js:479-527; HandParams,
 * shaders.js:394-402): `vel` is added to the velocity of every grid cell
 * within `radius` of the line through `pos` along the eye-to-pos ray, scaled
 * by (1 - d/radius)^2 * strength, between the grid update and G2P of every
 * step (WGSL_APPLY_HAND, shaders.js:410-432) until set again with active = 0.
 * 48 bytes. */
typedef struct mpm_hand {
    float pos[3];
    float radius;
    float vel[3];
    uint32_t active;
    float strength;
    float eye[3];
} mpm_hand;
int mpm_set_hand(mpm_scene *s, const mpm_hand *hand);

/* Replaces FluidGPU.deleteNear + pollDelete (fluid-gpu.js:549-586;
 * WGSL_COMPACT, shaders.js:521-547): removes the particles within `radius` of
 * the line through origin along dir.  The survivors keep their reset order
 * (their ids are renumbered 0..remaining-1) and carry their whole state;
 * synchronous. */
int mpm_delete_near(mpm_scene *s, const float origin[3], const float dir[3], float radius,
                    uint32_t *remaining);

/* Replaces FluidGPU.raycastFluid + pollRaycast (fluid-gpu.js:594-638;
 * WGSL_RAYCAST, shaders.js:557-577): the smallest t > 0 of dot(p - origin,
 * dir) over particles within `radius` of the ray axis; *hit = 0 when none.
 * Synchronous. */
int mpm_raycast(mpm_scene *s, const float origin[3], const float dir[3], float radius,
                float *t, int *hit);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight frame or serving loops, harnesses that forbid hipMalloc):
 * mpm_raycast for nrays rays at once (origins3, dirs3: nrays * 3 floats,
 * radii: nrays), t[i] and hit[i] exactly as mpm_raycast gives for ray i.  It
 * calls no allocator: its scratch is a library-static device array of
 * MPM_RAYCAST_NOALLOC_MAX rays (36 KiB on each GPU it runs on, held until
 * the process ends), so nrays above that is MPM_ERR_INVALID.  Calls that
 * overlap on one GPU share that scratch and race.  Synchronous. */
#define MPM_RAYCAST_NOALLOC_MAX 1024
int mpm_raycast_noalloc(mpm_scene *s, uint32_t nrays, const float *origins3, const float *dirs3,
                        const float *radii, float *t, int *hit);

/* Replaces the upload of FluidGPU.spawnParticles (fluid-gpu.js:145-180): n
 * particles (positions inside the hard bounds; vel3 may be NULL) appended
 * after the current ones with unit mass and C = 0, ids continuing the reset
 * order.  MPM_ERR_CAPACITY past the capacity. */
int mpm_append_particles(mpm_scene *s, uint32_t n, const float *pos3, const float *vel3);

/* The scene's hipStream_t, for callers that order their own work after it. */
int mpm_get_stream(mpm_scene *s, void **hip_stream);

/* Stage timing with HIP events recorded on the scene's stream around every
 * stage of every step (inside captured graphs as event-record nodes).
 * mpm_enable_timing(s, 1) clears the record and starts it; mpm_stage_times
 * gives the mean per step, in milliseconds, of
 * [p2g_mass, p2g_mom, update, g2p, diffuse, sort, whole step, relayout] over
 * the steps recorded since ("g2p" brackets the G2P kernel alone; a re-sort's
 * offsets, work items and permutation run inside the p2g_mass and p2g_mom
 * stages, so "relayout", the wait for them, is kept at zero); mpm_step_times
 * gives each recorded step's total (up to cap values; *count = number
 * recorded).  Both synchronise. */
int mpm_enable_timing(mpm_scene *s, int enable);
int mpm_stage_times(mpm_scene *s, float *ms8);
int mpm_step_times(mpm_scene *s, float *ms, uint32_t cap, uint32_t *count);

/* Diagnostics of the LDS-staged scatter: out[0] = work items of the current
 * particle order, out[1] = 8x8x8 bricks in the grid, out[2] = particles,
 * out[3] = largest item in particles.  Synchronises. */
int mpm_tile_stats(mpm_scene *s, uint64_t out[4]);

/* ------------------------------------------------------------------------
 * Batched form.  No counterpart in the reference, which steps one FluidGPU
 * per WebGPU device: a batch steps many scenes of uneven particle counts and
 * grid sizes with one persistent launch per stage, captured in a HIP graph.
 * Each member's result is bit-identical to stepping it alone.
 *   - Members re-sort every step (sort_period 1) on the LDS-staged path,
 *     carry no diffuse pool, share the tile margin and live on the batch's
 *     device; their work runs on the batch's stream while they are members
 *     (every single-scene call on a member stays valid and is ordered with
 *     the batch).
 *   - A member whose particles must first be sorted (after
 *     mpm_append_particles or mpm_delete_near) takes that step through the
 *     single-scene path, together with the other members.
 *   - mpm_destroy on a member removes it from its batch first.
 *   - The batch's descriptor tables grow with its member set; a replaced
 *     table buffer is freed by mpm_batch_destroy, not while the batch runs
 *     (hipFree synchronizes the device, which fails a stream capture that
 *     another thread has in progress).
 */
typedef struct mpm_batch mpm_batch;

#define MPM_BATCH_NO_GRAPH 1u  /* mpm_batch_tuning.flags: enqueue every step eagerly */
/* mpm_batch_tuning.flags: capture graphs only for descriptor tables that an
 * earlier mpm_batch_step call already stepped, so a member set that changes
 * at every call (a serving loop admitting and retiring scenes) pays no
 * capture; one that stays is captured at its second call. */
#define MPM_BATCH_GRAPH_REUSED 2u

/* Launch shapes of the batched kernels, as chosen by the autotuner
 * (mlsmpm.autotune): persistent blocks of the work-item kernels (P2G mass,
 * P2G momentum, G2P), of the resolve kernels and of the per-brick kernel
 * (the brick list; 0 = one block per unit), and steps per captured graph
 * (rounded up to even; 0 = 16). */
typedef struct mpm_batch_tuning {
    uint32_t item_blocks;
    uint32_t resolve_blocks;
    uint32_t aux_blocks;
    uint32_t graph_steps;
    uint32_t flags;
    uint32_t reserved[3];
} mpm_batch_tuning;

int mpm_batch_create(int device, mpm_batch **out);
/* Removes every member (they return to their own streams), then frees. */
int mpm_batch_destroy(mpm_batch *b);
int mpm_batch_add(mpm_batch *b, mpm_scene *s);
int mpm_batch_remove(mpm_batch *b, mpm_scene *s);
int mpm_batch_count(const mpm_batch *b, uint32_t *count);
int mpm_batch_set_tuning(mpm_batch *b, const mpm_batch_tuning *t);
int mpm_batch_get_tuning(const mpm_batch *b, mpm_batch_tuning *t);
/* Enqueue nsteps steps of every member (each member advances nsteps). */
int mpm_batch_step(mpm_batch *b, uint32_t nsteps);
int mpm_batch_sync(mpm_batch *b);
/* Per batched step GPU time from HIP events (timing disables graph capture);
 * mpm_batch_step_times synchronises and returns up to cap values. */
int mpm_batch_enable_timing(mpm_batch *b, int enable);
int mpm_batch_step_times(mpm_batch *b, float *ms, uint32_t cap, uint32_t *count);

/* ------------------------------------------------------------------------
 * Serving form.  No counterpart in the reference: a request loop with
 * continuous batching.  Requests (scene configuration, particles, step count)
 * are queued from any thread; one loop thread owns the device, admits queued
 * requests in arrival order while the HBM budget and the resident-scene
 * limit allow, keeps every admitted scene resident and advances all of them
 * together (one mpm_batch, chunks of chunk_steps batched steps), and retires
 * each request when its steps are done (final particles copied to host
 * memory held for mpm_server_result, device memory released).  A request's
 * result is bit-identical to stepping its scene alone.  Served scenes re-sort
 * every step, carry no diffuse pool and keep no i32 grids (sort_period,
 * flags and device of the submitted configuration are set accordingly), and
 * all use the default drift margin: tile_margin 0 or 2 is accepted (and set
 * to 2), any other value is MPM_ERR_UNSUPPORTED at submit.
 */
typedef struct mpm_server mpm_server;

#define MPM_REQ_QUEUED 0
#define MPM_REQ_RUNNING 1
#define MPM_REQ_DONE 2
#define MPM_REQ_FAILED 3
#define MPM_REQ_CANCELLED 4

typedef struct mpm_server_config {
    int32_t device;
    uint32_t max_active;      /* resident scenes at most (0 = no limit) */
    uint64_t hbm_budget;      /* device bytes the resident scenes may hold */
    uint32_t chunk_steps;     /* batched steps per loop iteration (0 = 8) */
    uint32_t item_blocks;     /* batch launch shapes (mpm_batch_tuning; 0 = default) */
    uint32_t resolve_blocks;
    uint32_t aux_blocks;
    /* Work-item size (mpm_config.tile_ppt: 4, 8 or 16; 0 = leave it) for a
     * request of n particles whose own tile_ppt is 0, by particle-count class
     * floor(log2(n)) / 2, as the autotuner (mlsmpm.autotune) tuned them for
     * batched steps. */
    uint8_t tile_ppt_by_class[16];
    /* mpm_batch_tuning.flags of the server's batch (0 = MPM_BATCH_GRAPH_REUSED:
     * a member set that lasts more than one chunk replays captured graphs;
     * MPM_BATCH_NO_GRAPH: every step enqueued eagerly). */
    uint32_t batch_flags;
    uint32_t pad;
} mpm_server_config;

typedef struct mpm_server_stats {
    uint64_t submitted, completed, failed, cancelled;
    uint32_t queued, active;
    uint64_t bytes_resident;  /* admission estimate of the resident scenes */
    uint64_t particle_steps;  /* particle-steps completed on the device */
    uint64_t batched_steps;
    uint64_t resident_steps;  /* sum over batched steps of the resident scenes */
    double busy_seconds;      /* loop time spent stepping resident scenes */
    double elapsed_seconds;   /* since mpm_server_create */
    /* Where the loop's time went (seconds, loop thread): */
    double enqueue_seconds;   /* enqueuing batched chunks (with descriptor table rebuilds) */
    double admit_seconds;     /* creating and loading admitted scenes */
    double wait_seconds;      /* waiting for a chunk to finish on the device */
    double retire_seconds;    /* reading back and releasing finished requests */
    double enroll_seconds;    /* joining prepared scenes to the batch */
    double gpu_step_seconds;  /* sum of the batched steps' device times (HIP events) */
    uint64_t chunks;          /* loop iterations that stepped the batch */
    uint64_t admitted;        /* requests admitted */
    uint64_t graph_steps;     /* batched steps replayed from captured graphs */
    uint64_t graph_captures;  /* graphs captured */
    /* admit_seconds and retire_seconds split (seconds, loop thread): */
    double admit_create_seconds;    /* creating admitted scenes (buffers, stream, zeroed state) */
    double admit_load_seconds;      /* uploading and sorting their particles */
    double retire_read_seconds;     /* reading finished requests' particles back */
    double retire_release_seconds;  /* releasing their scenes */
} mpm_server_stats;

typedef struct mpm_request_info {
    int32_t state;            /* MPM_REQ_* */
    uint32_t n, steps, steps_done;
    uint64_t footprint;       /* device bytes reserved at admission */
    double t_submit, t_admit, t_done;  /* seconds since mpm_server_create */
    uint32_t tile_ppt;        /* work-item size the request's scene runs with */
    uint32_t pad;
} mpm_request_info;

/* Starts the loop thread. */
int mpm_server_create(const mpm_server_config *cfg, mpm_server **out);
/* Cancels everything outstanding, joins the loop, frees. */
int mpm_server_destroy(mpm_server *srv);
/* Queue a scene: inputs are copied (pos3 n*3, vel3 n*3, C9 n*9, mass n; NULL
 * = zero / zero / 1.0).  MPM_ERR_NOMEM if the scene alone exceeds the
 * budget.  Never waits for the device. */
int mpm_server_submit(mpm_server *srv, const mpm_config *cfg, uint32_t n, const float *pos3,
                      const float *vel3, const float *C9, const float *mass, uint32_t steps,
                      uint64_t *id);
int mpm_server_poll(mpm_server *srv, uint64_t id, mpm_request_info *info);
/* Block until the request is done, failed or cancelled, or timeout_s passes
 * (negative = no limit); *state receives its state. */
int mpm_server_wait(mpm_server *srv, uint64_t id, double timeout_s, int *state);
/* Copy a finished request's final particles (reset order; layouts of
 * mpm_read_particles; any output may be NULL) and release the request. */
int mpm_server_result(mpm_server *srv, uint64_t id, float *pos4, float *vel4, float *C9,
                      float *density);
int mpm_server_cancel(mpm_server *srv, uint64_t id);
int mpm_server_get_stats(mpm_server *srv, mpm_server_stats *out);
/* GPU time of every batched step so far (HIP events), in milliseconds. */
int mpm_server_step_times(mpm_server *srv, float *ms, uint32_t cap, uint32_t *count);

/* The same steps with what was resident: for attributing the latency tail. */
typedef struct mpm_step_record {
    float ms;                    /* device time of the batched step */
    uint32_t resident;           /* scenes stepped */
    uint32_t admitting;          /* scenes being created, loaded and first-stepped meanwhile */
    uint32_t pad;
    uint64_t resident_particles; /* particles stepped */
} mpm_step_record;
int mpm_server_step_records(mpm_server *srv, mpm_step_record *rec, uint32_t cap,
                            uint32_t *count);

#ifdef __cplusplus
}
#endif
#endif
