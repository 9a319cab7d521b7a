/*
 * mpm_oracle.h -- CPU restatement of the reference MLS-MPM fluid step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against; it is linked or called only by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg.  The product library never includes it.
 *
 * The same source (mpm_oracle.c) is compiled twice:
 *   liborc32.so  -- ORC_REAL=float  : "exact semantics" restatement, built with
 *                   -ffp-contract=off so every WGSL expression is evaluated as
 *                   written, left to right, one IEEE-754 rounding per operator.
 *   liborc64.so  -- ORC_REAL=double : FP64 twin used to state the FP32 error.
 *
 * Every kernel below follows one WGSL compute entry point of the reference
 * (shaders.js) expression by expression; the citation is on each function in
 * mpm_oracle.c.  Buffers use the reference layouts: particle pos/vel are
 * stride-4 vec4 (w = 0), the affine matrix C is 9 reals row-major per
 * particle, cell_mv is 3 interleaved i32 per cell (idx*3 + c), cell_v is 3
 * interleaved reals per cell, grids are x-fastest: idx = (cz*gy + cy)*gx + cx.
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H
#include <stdint.h>

#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL orc_real;

/* Params uniform, 80 bytes (shaders.js:27-49; packed by fluid-gpu.js:295-313).
 * Kept as f32 in both builds: the uniform is f32 in the reference. */
typedef struct {
    uint32_t grid_x, grid_y, grid_z, particle_num;
    float dt, gravity, rest_density, stiffness, eos_power, viscosity;
    float hard_min, hard_max_x, hard_max_y, hard_max_z;
    float wall_min, wall_stiffness, lookahead_k;
    float pad0, pad1, pad2;
} orc_params;

/* BaseWeights (shaders.js:65-71): base cell, fractional offset, 3x3 weights. */
typedef struct {
    int32_t b[3];
    orc_real f[3];
    orc_real w[3][3]; /* w[axis][i] = wx0 wx1 wx2 wy0 .. wz2 */
} orc_bw;

/* DiffuseParticle (shaders.js:725-732): 48 bytes in the f32 build. */
typedef struct {
    orc_real pos[4];
    orc_real vel[4];
    uint32_t ptype;
    uint32_t alive;
    orc_real lifetime;
    orc_real pad;
} orc_drec;

/* GenParams (shaders.js:734-743; diffuse-gpu.js:84-94). */
typedef struct {
    float ke_threshold, spawn_rate_k, density_spray_max, density_bubble_min;
    float lifetime_foam, pressure_threshold, crest_dot_threshold;
    uint32_t seed;
    uint32_t lifetime_mode; /* 0: reference; 1: uniform [0.5, 3] (extension, not in the reference) */
} orc_gen;

/* AdvectParams (shaders.js:872-877; diffuse-gpu.js:70-76). */
typedef struct {
    uint32_t diffuse_max;
    float buoyancy, drag, pad;
} orc_adv;

/* Wall uniform: 4 floats header (count in [0]) + 4 walls x 20 floats
 * (shaders.js:123-133; fluid-gpu.js:337-413).  NULL means count = 0. */
#define ORC_MAX_WALLS 4
#define ORC_WALLP_FLOATS (4 + ORC_MAX_WALLS * 20)

/* HandParams uniform, 80 bytes (shaders.js:394-402), filled by
 * FluidGPU.updateHand (fluid-gpu.js:479-527) including the cell box of the
 * push cylinder inside the grid (bbox_min > bbox_max: nothing dispatched). */
typedef struct {
    float px, py, pz, radius;
    float vx, vy, vz;
    uint32_t is_on;
    float strength, ex, ey, ez;
    int32_t bbox_min[3], bbox_max[3];
    int32_t pad0, pad1;
} orc_hand;

/* RayParams (shaders.js:78): ray origin, cylinder radius, direction. */
typedef struct {
    float ox, oy, oz, radius, dx, dy, dz, pad;
} orc_ray;

/* Whole-scene state for orc_step: caller-owned buffers. */
typedef struct {
    orc_real *pos4, *vel4, *affine, *mass, *density; /* 4n, 4n, 9n, n, n */
    orc_bw *bw;                                      /* n */
    int32_t *cell_mass, *cell_mv;                    /* G, 3G */
    orc_real *cell_mass_f, *cell_v;                  /* G, 3G */
    const float *wallp;                              /* ORC_WALLP_FLOATS or NULL */
    /* diffuse (pool == NULL: no diffuse attached) */
    orc_drec *pool;
    uint32_t *free_list;
    int32_t *top;
    orc_gen gen; /* gen.seed is advanced by orc_step like DiffuseGPU._seed */
    orc_adv adv;
    const orc_hand *hand; /* NULL: no hand push */
} orc_state;

#ifdef __cplusplus
extern "C" {
#endif
int orc_real_bytes(void);
void orc_set_threads(int n);
int32_t orc_encode(orc_real v);
orc_real orc_decode(int32_t v);
float orc_diffuse_rand(uint32_t seed);
void orc_clear(const orc_params *P, int32_t *cell_mv, int32_t *cell_mass);
void orc_compute_weights(const orc_params *P, const orc_real *pos4, orc_bw *bw);
void orc_p2g_mass(const orc_params *P, const orc_real *mass, int32_t *cell_mass, const orc_bw *bw);
void orc_decode_mass(const orc_params *P, const int32_t *cell_mass, orc_real *cell_mass_f);
void orc_p2g_mom(const orc_params *P, const orc_real *vel4, const orc_real *affine,
                 const orc_real *mass, const orc_real *cell_mass_f, int32_t *cell_mv,
                 orc_real *density, const orc_bw *bw);
void orc_update_grid(const orc_params *P, const int32_t *cell_mv, const int32_t *cell_mass,
                     orc_real *cell_v, const float *wallp);
void orc_g2p(const orc_params *P, orc_real *pos4, orc_real *vel4, orc_real *affine,
             const orc_real *cell_v, const orc_bw *bw, const float *wallp);
void orc_diffuse_generate(const orc_params *P, const orc_gen *gen, const orc_real *pos4,
                          const orc_real *vel4, const orc_real *mass, const orc_real *density,
                          const orc_real *cell_mass_f, orc_drec *pool, uint32_t *free_list,
                          int32_t *top, uint32_t diffuse_max);
void orc_diffuse_advect(const orc_params *P, const orc_adv *adv, orc_drec *pool,
                        const orc_real *cell_v, uint32_t *free_list, int32_t *top,
                        const float *wallp);
void orc_hand_bbox(const orc_params *P, orc_hand *h);
void orc_apply_hand(const orc_params *P, const orc_hand *h, orc_real *cell_v);
orc_real orc_ray_perp_dist(const orc_real *pos4, const orc_ray *r);
int orc_raycast(const orc_real *pos4, uint32_t n, const orc_ray *r, orc_real *t);
uint32_t orc_delete_mask(const orc_real *pos4, uint32_t n, const orc_ray *r, uint8_t *keep);
void orc_step(const orc_params *P, orc_state *S, uint32_t nsteps);
/* Helpers of the kernels above, exported for the tests' closed-form pins
 * (wall = 20 floats of one wall, shaders.js:123-129). */
void orc_wall_to_local(const orc_real p[3], const float *wall, orc_real out[3]);
void orc_wall_to_world_dir(const orc_real v[3], const float *wall, orc_real out[3]);
void orc_wall_vel_at(const orc_real p[3], const float *wall, orc_real out[3]);
orc_real orc_wall_sdf_normal(const orc_real lp[3], const orc_real mn[3], const orc_real mx[3],
                             orc_real n[3]);
int orc_estimate_normal(const orc_params *P, const orc_real *cell_mass_f, const orc_real p[3],
                        orc_real n[3]);
/* terms = (ke, trappedAir, waveCrest, speed); returns the spawn score. */
orc_real orc_spawn_score(const orc_params *P, const orc_gen *gen, const orc_real *cell_mass_f,
                         const orc_real pos[3], const orc_real vel[3], orc_real mass,
                         orc_real density, orc_real terms[4]);
#ifdef __cplusplus
}
#endif
#endif
