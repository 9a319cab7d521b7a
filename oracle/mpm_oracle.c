/*
 * mpm_oracle.c -- CPU restatement of the reference MLS-MPM fluid step.
 *
 * TEST INFRASTRUCTURE ONLY (see mpm_oracle.h): the checker, never the product.
 *
 * Each function cites the WGSL entry point (shaders.js) or host method
 * (fluid-gpu.js / diffuse-gpu.js) whose arithmetic it restates.  Evaluation
 * rule: WGSL evaluates left to right, so `a*b*c` is `(a*b)*c` and
 * `x + y*2.0 + z*1.5` is `(x + (y*2.0)) + (z*1.5)`; each operator rounds once
 * (the library is built with -ffp-contract=off, no -ffast-math).  Integer grid
 * adds wrap (two's complement), done on uint32_t to stay defined in C.
 *
 * Two points the WGSL leaves to the WebGPU implementation are fixed here:
 *   pow(x, e): e == 1 returns x exactly; otherwise pow in double rounded once.
 *   length(v): sqrt((x*x + y*y) + z*z), correctly rounded sqrt.
 * Division and sqrt are IEEE correctly rounded.
 */
#include "mpm_oracle.h"
#include <math.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef orc_real real;

#ifdef ORC_DOUBLE
#define K(x) ((real)(x))
#define SQRTR sqrt
#define FLOORR floor
#define MAXR fmax
#define MINR fmin
#else
#define K(x) ((real)(x##f))
#define SQRTR sqrtf
#define FLOORR floorf
#define MAXR fmaxf
#define MINR fminf
#endif

int orc_real_bytes(void) { return (int)sizeof(real); }

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* f32 -> i32 as the GPU converts: truncate toward zero, saturate, NaN -> 0. */
static inline int32_t cvt_i32(real x) {
    if (!(x == x)) return 0;
    if (x >= K(2147483647.0)) return 2147483647;
    if (x <= K(-2147483648.0)) return (int32_t)(-2147483647 - 1);
    return (int32_t)x;
}

static inline void atomic_add_i32(int32_t *dst, int32_t v) {
    uint32_t *u = (uint32_t *)dst;
#pragma omp atomic
    *u += (uint32_t)v;
}

/* encode / decode, FIXED = 1e5 (shaders.js:15-25). */
int32_t orc_encode(real v) {
    real c = MINR(MAXR(v, K(-20000.0)), K(20000.0));
    return cvt_i32(c * K(100000.0));
}
real orc_decode(int32_t v) { return (real)v / K(100000.0); }

/* pow as fixed above (shaders.js:326, 813). */
static inline real eos_pow(real x, float e) {
    if (e == 1.0f) return x;
    return (real)pow((double)x, (double)e);
}

/* P2G_MOM / DIFFUSE_GENERATE pressure with the gated cohesion floor
 * (shaders.js:12-13, 326-328 and 813-815 are the same expression). */
static inline real eos_pressure(const orc_params *P, real density) {
    real rest = (real)P->rest_density;
    real pressure = (real)P->stiffness * (eos_pow(density / rest, P->eos_power) - K(1.0));
    real pfloor = (density > rest * K(0.7)) ? K(-0.1) : K(0.0);
    return MAXR(pressure, pfloor);
}

/* cell_idx_checked (shaders.js:93-98). */
static inline int32_t cell_idx_checked(int32_t cx, int32_t cy, int32_t cz, const orc_params *P) {
    if (cx < 0 || cx >= (int32_t)P->grid_x || cy < 0 || cy >= (int32_t)P->grid_y || cz < 0 ||
        cz >= (int32_t)P->grid_z)
        return -1;
    return (cz * (int32_t)P->grid_y + cy) * (int32_t)P->grid_x + cx;
}

/* quadratic_weights (shaders.js:53-59). */
static inline void quad_weights(real f, real w[3]) {
    w[0] = (K(0.5) * (K(0.5) - f)) * (K(0.5) - f);
    w[1] = K(0.75) - f * f;
    w[2] = (K(0.5) * (K(0.5) + f)) * (K(0.5) + f);
}

/* ---- wall helpers (shaders.js:123-193); wall = 20 floats ---------------- */
typedef struct { real x, y, z; } v3;
static inline v3 mk3(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline real dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

static inline uint32_t wall_count(const float *wallp) {
    if (!wallp) return 0;
    float c = wallp[0];
    if (!(c > 0.0f)) return 0;
    uint32_t n = (uint32_t)c;
    return n > ORC_MAX_WALLS ? ORC_MAX_WALLS : n;
}
static inline const float *wall_at(const float *wallp, uint32_t i) { return wallp + 4 + 20 * i; }

/* wall_to_local (shaders.js:137-146). */
static inline v3 wall_to_local(v3 p, const float *w) {
    v3 d = mk3(p.x - (real)w[0], p.y - (real)w[1], p.z - (real)w[2]);
    real cy = (real)w[3], sy = (real)w[7], cp = (real)w[11], sp = (real)w[15];
    real x1 = cy * d.x - sy * d.z;
    real z1 = sy * d.x + cy * d.z;
    return mk3(x1, cp * d.y + sp * z1, (-sp) * d.y + cp * z1);
}
/* wall_to_world_dir (shaders.js:149-157). */
static inline v3 wall_to_world_dir(v3 v, const float *w) {
    real cy = (real)w[3], sy = (real)w[7], cp = (real)w[11], sp = (real)w[15];
    real y1 = cp * v.y - sp * v.z;
    real z1 = sp * v.y + cp * v.z;
    return mk3(cy * v.x + sy * z1, y1, (-sy) * v.x + cy * z1);
}
/* wall_vel_at (shaders.js:161-164): vel + cross(omega, p - center). */
static inline v3 wall_vel_at(v3 p, const float *w) {
    v3 r = mk3(p.x - (real)w[0], p.y - (real)w[1], p.z - (real)w[2]);
    real ox = (real)w[16], oy = (real)w[17], oz = (real)w[18];
    v3 c = mk3(oy * r.z - oz * r.y, oz * r.x - ox * r.z, ox * r.y - oy * r.x);
    return mk3((real)w[12] + c.x, (real)w[13] + c.y, (real)w[14] + c.z);
}
/* wall_sdf_normal (shaders.js:174-193): returns normal in n, sdf as value. */
static inline real wall_sdf_normal(v3 lp, v3 mn, v3 mx, v3 *n) {
    v3 dmin = mk3(mn.x - lp.x, mn.y - lp.y, mn.z - lp.z);
    v3 dmax = mk3(lp.x - mx.x, lp.y - mx.y, lp.z - mx.z);
    v3 q = mk3(MAXR(dmin.x, dmax.x), MAXR(dmin.y, dmax.y), MAXR(dmin.z, dmax.z));
    real qmax = MAXR(q.x, MAXR(q.y, q.z));
    if (qmax > K(0.0)) {
        v3 nearest = mk3(MINR(MAXR(lp.x, mn.x), mx.x), MINR(MAXR(lp.y, mn.y), mx.y),
                         MINR(MAXR(lp.z, mn.z), mx.z));
        v3 diff = mk3(lp.x - nearest.x, lp.y - nearest.y, lp.z - nearest.z);
        real len = SQRTR(dot3(diff, diff));
        if (len > K(1e-6)) *n = mk3(diff.x / len, diff.y / len, diff.z / len);
        else *n = mk3(K(0.0), K(1.0), K(0.0));
        return len;
    }
    if (qmax == q.x) *n = mk3(dmax.x > dmin.x ? K(1.0) : K(-1.0), K(0.0), K(0.0));
    else if (qmax == q.y) *n = mk3(K(0.0), dmax.y > dmin.y ? K(1.0) : K(-1.0), K(0.0));
    else *n = mk3(K(0.0), K(0.0), dmax.z > dmin.z ? K(1.0) : K(-1.0));
    return qmax;
}
static inline v3 wall_min3(const float *w) { return mk3((real)w[4], (real)w[5], (real)w[6]); }
static inline v3 wall_max3(const float *w) { return mk3((real)w[8], (real)w[9], (real)w[10]); }
static inline int inside_open(v3 l, v3 mn, v3 mx) {
    return l.x > mn.x && l.y > mn.y && l.z > mn.z && l.x < mx.x && l.y < mx.y && l.z < mx.z;
}

/* ---- WGSL_CLEAR (shaders.js:196-210) ------------------------------------ */
void orc_clear(const orc_params *P, int32_t *cell_mv, int32_t *cell_mass) {
    size_t G = (size_t)P->grid_x * P->grid_y * P->grid_z;
    memset(cell_mass, 0, G * sizeof(int32_t));
    memset(cell_mv, 0, 3 * G * sizeof(int32_t));
}

/* ---- WGSL_COMPUTE_WEIGHTS (shaders.js:216-240) -------------------------- */
void orc_compute_weights(const orc_params *P, const real *pos4, orc_bw *bw) {
    long n = (long)P->particle_num;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < n; p++) {
        orc_bw o;
        for (int a = 0; a < 3; a++) {
            real x = pos4[4 * p + a];
            int32_t ci = cvt_i32(FLOORR(x));
            o.b[a] = ci - 1;
            o.f[a] = (x - (real)ci) - K(0.5);
            quad_weights(o.f[a], o.w[a]);
        }
        bw[p] = o;
    }
}

/* ---- WGSL_P2G_MASS (shaders.js:242-274) --------------------------------- */
void orc_p2g_mass(const orc_params *P, const real *mass, int32_t *cell_mass, const orc_bw *bw) {
    long n = (long)P->particle_num;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < n; p++) {
        real pm = mass[p];
        const orc_bw *b = &bw[p];
        for (int k = 0; k < 3; k++) {
            int32_t cz = b->b[2] + k;
            for (int j = 0; j < 3; j++) {
                int32_t cy = b->b[1] + j;
                for (int i = 0; i < 3; i++) {
                    int32_t cx = b->b[0] + i;
                    int32_t idx = cell_idx_checked(cx, cy, cz, P);
                    if (idx < 0) continue;
                    real w = (b->w[0][i] * b->w[1][j]) * b->w[2][k];
                    atomic_add_i32(&cell_mass[idx], orc_encode(w * pm));
                }
            }
        }
    }
}

/* ---- WGSL_DECODE_MASS (shaders.js:373-384) ------------------------------ */
void orc_decode_mass(const orc_params *P, const int32_t *cell_mass, real *cell_mass_f) {
    long G = (long)P->grid_x * P->grid_y * P->grid_z;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < G; c++) cell_mass_f[c] = orc_decode(cell_mass[c]);
}

/* ---- WGSL_P2G_MOM (shaders.js:276-371) ---------------------------------- */
void orc_p2g_mom(const orc_params *P, const real *vel4, const real *affine, const real *mass,
                 const real *cell_mass_f, int32_t *cell_mv, real *density_out, const orc_bw *bw) {
    long n = (long)P->particle_num;
    const uint32_t gx = P->grid_x, gy = P->grid_y;
    const uint32_t SLICE = gx * gy;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < n; p++) {
        real vx = vel4[4 * p], vy = vel4[4 * p + 1], vz = vel4[4 * p + 2];
        real m = mass[p];
        const real *a = &affine[9 * p];
        real a00 = a[0], a01 = a[1], a02 = a[2];
        real a10 = a[3], a11 = a[4], a12 = a[5];
        real a20 = a[6], a21 = a[7], a22 = a[8];
        const orc_bw *b = &bw[p];
        uint32_t baseIdx = (uint32_t)((b->b[2] * (int32_t)gy + b->b[1]) * (int32_t)gx + b->b[0]);

        real density = K(0.0);
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++)
                for (int i = 0; i < 3; i++) {
                    uint32_t cidx = baseIdx + (uint32_t)k * SLICE + (uint32_t)j * gx + (uint32_t)i;
                    density += ((b->w[0][i] * b->w[1][j]) * b->w[2][k]) * cell_mass_f[cidx];
                }
        density_out[p] = density;
        if (density <= K(1e-8)) continue;

        real rest = (real)P->rest_density;
        real vol = m / MAXR(density, rest * K(0.15));
        real pressure = eos_pressure(P, density);
        real visc = (real)P->viscosity;
        real s00 = -pressure + visc * (a00 + a00);
        real s11 = -pressure + visc * (a11 + a11);
        real s22 = -pressure + visc * (a22 + a22);
        real s01 = visc * (a01 + a10);
        real s02 = visc * (a02 + a20);
        real s12 = visc * (a12 + a21);
        real coef = ((-vol) * K(4.0)) * (real)P->dt;

        real fx = b->f[0], fy = b->f[1], fz = b->f[2];
        real dxs[3] = {(-fx) - K(1.0), -fx, K(1.0) - fx};
        real dys[3] = {(-fy) - K(1.0), -fy, K(1.0) - fy};
        real dzs[3] = {(-fz) - K(1.0), -fz, K(1.0) - fz};

        for (int k = 0; k < 3; k++) {
            real wzk = b->w[2][k], dzk = dzs[k];
            for (int j = 0; j < 3; j++) {
                real wyzk = b->w[1][j] * wzk, dyj = dys[j];
                for (int i = 0; i < 3; i++) {
                    real w = b->w[0][i] * wyzk;
                    real dxi = dxs[i];
                    uint32_t cidx = baseIdx + (uint32_t)k * SLICE + (uint32_t)j * gx + (uint32_t)i;
                    real Qx = (a00 * dxi + a01 * dyj) + a02 * dzk;
                    real Qy = (a10 * dxi + a11 * dyj) + a12 * dzk;
                    real Qz = (a20 * dxi + a21 * dyj) + a22 * dzk;
                    real fxf = coef * ((s00 * dxi + s01 * dyj) + s02 * dzk);
                    real fyf = coef * ((s01 * dxi + s11 * dyj) + s12 * dzk);
                    real fzf = coef * ((s02 * dxi + s12 * dyj) + s22 * dzk);
                    atomic_add_i32(&cell_mv[cidx * 3 + 0], orc_encode(w * (m * (vx + Qx) + fxf)));
                    atomic_add_i32(&cell_mv[cidx * 3 + 1], orc_encode(w * (m * (vy + Qy) + fyf)));
                    atomic_add_i32(&cell_mv[cidx * 3 + 2], orc_encode(w * (m * (vz + Qz) + fzf)));
                }
            }
        }
    }
}

/* ---- WGSL_UPDATE_GRID (shaders.js:435-513) ------------------------------ */
void orc_update_grid(const orc_params *P, const int32_t *cell_mv, const int32_t *cell_mass,
                     real *cell_v, const float *wallp) {
    long G = (long)P->grid_x * P->grid_y * P->grid_z;
    const uint32_t gx = P->grid_x, gy = P->grid_y, gz = P->grid_z;
    uint32_t nw = wall_count(wallp);
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < G; idx++) {
        real cm = orc_decode(cell_mass[idx]);
        if (cm <= K(0.0)) {
            cell_v[3 * idx] = K(0.0);
            cell_v[3 * idx + 1] = K(0.0);
            cell_v[3 * idx + 2] = K(0.0);
            continue;
        }
        real vx = orc_decode(cell_mv[3 * idx]) / cm;
        real vy = orc_decode(cell_mv[3 * idx + 1]) / cm;
        real vz = orc_decode(cell_mv[3 * idx + 2]) / cm;
        vy += (real)P->gravity * (real)P->dt;

        int32_t cx = (int32_t)((uint32_t)idx % gx);
        int32_t cy = (int32_t)(((uint32_t)idx / gx) % gy);
        int32_t cz = (int32_t)((uint32_t)idx / (gx * gy));
        if (cx < 2 || cx > (int32_t)gx - 3) vx = K(0.0);
        if (cy < 2 && vy < K(0.0)) vy = K(0.0);
        if (cy > (int32_t)gy - 3 && vy > K(0.0)) vy = K(0.0);
        if (cz < 2 || cz > (int32_t)gz - 3) vz = K(0.0);
        if (cy < 2) {
            vx *= K(0.85);
            vz *= K(0.9);
        }

        v3 cellPos = mk3((real)cx + K(0.5), (real)cy + K(0.5), (real)cz + K(0.5));
        v3 wv = mk3(vx, vy, vz);
        for (uint32_t wi = 0; wi < nw; wi++) {
            const float *w = wall_at(wallp, wi);
            v3 lp = wall_to_local(cellPos, w);
            v3 mn = wall_min3(w), mx = wall_max3(w);
            if (lp.x < mn.x - K(1.0) || lp.y < mn.y - K(1.0) || lp.z < mn.z - K(1.0) ||
                lp.x > mx.x + K(1.0) || lp.y > mx.y + K(1.0) || lp.z > mx.z + K(1.0))
                continue;
            v3 wvel = wall_vel_at(cellPos, w);
            v3 nl;
            real sdf = wall_sdf_normal(lp, mn, mx, &nl);
            if (sdf < K(0.0)) {
                wv = wvel;
            } else if (sdf < K(1.0)) {
                v3 nn = wall_to_world_dir(nl, w);
                v3 vrel = mk3(wv.x - wvel.x, wv.y - wvel.y, wv.z - wvel.z);
                real vn = dot3(vrel, nn);
                if (vn < K(0.0)) wv = mk3(wv.x - vn * nn.x, wv.y - vn * nn.y, wv.z - vn * nn.z);
            }
        }
        cell_v[3 * idx] = wv.x;
        cell_v[3 * idx + 1] = wv.y;
        cell_v[3 * idx + 2] = wv.z;
    }
}

/* hard_clamp_axis (shaders.js:103-108). */
static inline void hard_clamp_axis(real *p, real *v, real lo, real hi) {
    if (*p < lo) { *p = lo; *v = K(0.0); }
    if (*p > hi) { *p = hi; *v = K(0.0); }
}

/* ---- WGSL_G2P (shaders.js:579-705), FLOOR_SPRING_BAND = 3 (shaders.js:5) */
void orc_g2p(const orc_params *P, real *pos4, real *vel4, real *affine, const real *cell_v,
             const orc_bw *bw, const float *wallp) {
    long n = (long)P->particle_num;
    const uint32_t gx = P->grid_x, gy = P->grid_y;
    const uint32_t SLICE = gx * gy;
    uint32_t nw = wall_count(wallp);
    const real dt = (real)P->dt;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < n; p++) {
        real px = pos4[4 * p], py = pos4[4 * p + 1], pz = pos4[4 * p + 2];
        const orc_bw *b = &bw[p];
        uint32_t baseIdx = (uint32_t)((b->b[2] * (int32_t)gy + b->b[1]) * (int32_t)gx + b->b[0]);
        real gvx = K(0.0), gvy = K(0.0), gvz = K(0.0);
        real B00 = K(0.0), B01 = K(0.0), B02 = K(0.0);
        real B10 = K(0.0), B11 = K(0.0), B12 = K(0.0);
        real B20 = K(0.0), B21 = K(0.0), B22 = K(0.0);
        for (int k = 0; k < 3; k++) {
            real wzk = b->w[2][k];
            real dzk = ((real)(b->b[2] + k) - pz) + K(0.5);
            for (int j = 0; j < 3; j++) {
                real wyzk = b->w[1][j] * wzk;
                real dyj = ((real)(b->b[1] + j) - py) + K(0.5);
                for (int i = 0; i < 3; i++) {
                    real w = b->w[0][i] * wyzk;
                    real dxi = ((real)(b->b[0] + i) - px) + K(0.5);
                    uint32_t cidx = baseIdx + (uint32_t)k * SLICE + (uint32_t)j * gx + (uint32_t)i;
                    real cvx = cell_v[cidx * 3], cvy = cell_v[cidx * 3 + 1], cvz = cell_v[cidx * 3 + 2];
                    real wvx = w * cvx, wvy = w * cvy, wvz = w * cvz;
                    gvx += wvx; gvy += wvy; gvz += wvz;
                    B00 += wvx * dxi; B01 += wvx * dyj; B02 += wvx * dzk;
                    B10 += wvy * dxi; B11 += wvy * dyj; B12 += wvy * dzk;
                    B20 += wvz * dxi; B21 += wvz * dyj; B22 += wvz * dzk;
                }
            }
        }
        real *a = &affine[9 * p];
        a[0] = B00 * K(4.0); a[1] = B01 * K(4.0); a[2] = B02 * K(4.0);
        a[3] = B10 * K(4.0); a[4] = B11 * K(4.0); a[5] = B12 * K(4.0);
        a[6] = B20 * K(4.0); a[7] = B21 * K(4.0); a[8] = B22 * K(4.0);

        v3 np = mk3(px + dt * gvx, py + dt * gvy, pz + dt * gvz);
        v3 nv = mk3(gvx, gvy, gvz);

        real lk = (real)P->lookahead_k;
        real floorZoneY = (real)P->hard_min + K(3.0);
        real xny = np.y + (nv.y * dt) * lk;
        if (xny < floorZoneY) {
            real distBelowFloor = floorZoneY - xny;
            real blend = MINR(MAXR(distBelowFloor / K(3.0), K(0.0)), K(1.0));
            nv.y += ((real)P->wall_stiffness * blend) * distBelowFloor;
        }

        for (uint32_t wi = 0; wi < nw; wi++) {
            const float *w = wall_at(wallp, wi);
            v3 wmin = wall_min3(w), wmax = wall_max3(w);
            v3 xnw = mk3(np.x + (nv.x * dt) * lk, np.y + (nv.y * dt) * lk, np.z + (nv.z * dt) * lk);
            v3 lxn = wall_to_local(xnw, w);
            if (inside_open(lxn, wmin, wmax)) {
                v3 nl;
                real sxn = wall_sdf_normal(lxn, wmin, wmax, &nl);
                real dist = -sxn;
                real blend = MINR(MAXR(dist / (real)P->wall_min, K(0.0)), K(1.0));
                v3 nn = wall_to_world_dir(nl, w);
                real s = ((real)P->wall_stiffness * blend) * dist;
                nv = mk3(nv.x + nn.x * s, nv.y + nn.y * s, nv.z + nn.z * s);
            }
            v3 lnp = wall_to_local(np, w);
            if (inside_open(lnp, wmin, wmax)) {
                v3 nl;
                real snp = wall_sdf_normal(lnp, wmin, wmax, &nl);
                v3 nn = wall_to_world_dir(nl, w);
                real push = MAXR(snp, K(-0.5));
                np = mk3(np.x - nn.x * push, np.y - nn.y * push, np.z - nn.z * push);
                v3 wvel = wall_vel_at(np, w);
                real vn = dot3(mk3(nv.x - wvel.x, nv.y - wvel.y, nv.z - wvel.z), nn);
                if (vn < K(0.0)) nv = mk3(nv.x - vn * nn.x, nv.y - vn * nn.y, nv.z - vn * nn.z);
            }
        }

        hard_clamp_axis(&np.x, &nv.x, (real)P->hard_min, (real)P->hard_max_x);
        hard_clamp_axis(&np.y, &nv.y, (real)P->hard_min, (real)P->hard_max_y);
        hard_clamp_axis(&np.z, &nv.z, (real)P->hard_min, (real)P->hard_max_z);
        pos4[4 * p] = np.x; pos4[4 * p + 1] = np.y; pos4[4 * p + 2] = np.z; pos4[4 * p + 3] = K(0.0);
        vel4[4 * p] = nv.x; vel4[4 * p + 1] = nv.y; vel4[4 * p + 2] = nv.z; vel4[4 * p + 3] = K(0.0);
    }
}

/* ---- diffuse ------------------------------------------------------------ */
#define DIFFUSE_SPRAY 0u
#define DIFFUSE_FOAM 1u
#define DIFFUSE_BUBBLE 2u

static inline uint32_t xorshift32(uint32_t x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}
/* diffuse_rand (shaders.js:756-762): f32(x) / 4294967295.0 in f32. */
float orc_diffuse_rand(uint32_t seed) { return (float)xorshift32(seed) / 4294967295.0f; }
static inline real drand(uint32_t seed) { return (real)xorshift32(seed) / K(4294967295.0); }

/* estimate_normal (shaders.js:766-787); returns 1 and fills n if valid. */
static int estimate_normal(const orc_params *P, const real *cmf, real px, real py, real pz, v3 *n) {
    int32_t gxi = cvt_i32(FLOORR(px)), gyi = cvt_i32(FLOORR(py)), gzi = cvt_i32(FLOORR(pz));
    int32_t X = (int32_t)P->grid_x, Y = (int32_t)P->grid_y, Z = (int32_t)P->grid_z;
    if (gxi < 1 || gxi >= X - 1 || gyi < 1 || gyi >= Y - 1 || gzi < 1 || gzi >= Z - 1) return 0;
    uint32_t idxL = (uint32_t)((gzi * Y + gyi) * X + (gxi - 1));
    uint32_t idxR = (uint32_t)((gzi * Y + gyi) * X + (gxi + 1));
    uint32_t idxD = (uint32_t)((gzi * Y + (gyi - 1)) * X + gxi);
    uint32_t idxU = (uint32_t)((gzi * Y + (gyi + 1)) * X + gxi);
    uint32_t idxB = (uint32_t)(((gzi - 1) * Y + gyi) * X + gxi);
    uint32_t idxF = (uint32_t)(((gzi + 1) * Y + gyi) * X + gxi);
    v3 g = mk3(cmf[idxL] - cmf[idxR], cmf[idxD] - cmf[idxU], cmf[idxB] - cmf[idxF]);
    real len = SQRTR(dot3(g, g));
    if (len < K(1e-6)) return 0;
    *n = mk3(g.x / len, g.y / len, g.z / len);
    return 1;
}

/* The spawn potential of one fluid particle (WGSL_DIFFUSE_GENERATE,
 * shaders.js:803-826): kinetic energy, trapped air from the EOS pressure and
 * the wave crest from the mass-gradient normal.  terms = (ke, trappedAir,
 * waveCrest, speed). */
static real spawn_score(const orc_params *P, const orc_gen *gen, const real *cmf, v3 pos, v3 vel,
                        real mass, real density, real terms[4]) {
    real speed2 = (vel.x * vel.x + vel.y * vel.y) + vel.z * vel.z;
    real speed = SQRTR(speed2);
    real ke = (K(0.5) * mass) * speed2;
    real pressure = eos_pressure(P, density);
    real trappedAir = MAXR(K(0.0), pressure - (real)gen->pressure_threshold);
    v3 nrm;
    int valid = estimate_normal(P, cmf, pos.x, pos.y, pos.z, &nrm);
    real waveCrest = K(0.0);
    if (speed > K(1e-4) && valid) {
        real dotVal = ((vel.x / speed) * nrm.x + (vel.y / speed) * nrm.y) + (vel.z / speed) * nrm.z;
        if (dotVal > (real)gen->crest_dot_threshold) waveCrest = dotVal * speed;
    }
    terms[0] = ke;
    terms[1] = trappedAir;
    terms[2] = waveCrest;
    terms[3] = speed;
    return (ke + trappedAir * K(2.0)) + waveCrest * K(1.5);
}

/* WGSL_DIFFUSE_GENERATE (shaders.js:721-856).  Particles are visited in index
 * order, so the free-list pops (try_spawn, shaders.js:791-798) happen in
 * particle-index order: the serialization the HIP library also uses. */
void orc_diffuse_generate(const orc_params *P, const orc_gen *gen, const real *pos4,
                          const real *vel4, const real *mass, const real *density_in,
                          const real *cmf, orc_drec *pool, uint32_t *free_list, int32_t *top,
                          uint32_t gen_dmax) {
    const real dt = (real)P->dt;
    for (uint32_t p = 0; p < P->particle_num; p++) {
        real density = density_in[p];
        if (density <= K(0.0) || density > (real)gen->density_bubble_min) continue;
        real vx = vel4[4 * p], vy = vel4[4 * p + 1], vz = vel4[4 * p + 2];
        real px = pos4[4 * p], py = pos4[4 * p + 1], pz = pos4[4 * p + 2];
        real terms[4];
        real score = spawn_score(P, gen, cmf, mk3(px, py, pz), mk3(vx, vy, vz), mass[p], density,
                                 terms);
        real speed = terms[3];
        real keth = (real)gen->ke_threshold;
        if (score < keth) continue;

        uint32_t randSeed = p * 9781u + gen->seed;
        real spawnChance = MINR(K(1.0), ((score - keth) * (real)gen->spawn_rate_k) * dt);
        if (drand(randSeed) > spawnChance) continue;

        int32_t prevTop = *top;
        *top = prevTop - 1;
        if (prevTop <= 0) {
            *top = *top + 1;
            continue;
        }
        if ((uint32_t)(prevTop - 1) >= gen_dmax) continue; /* inconsistent pool only */
        uint32_t slot = free_list[prevTop - 1];
        if (slot >= gen_dmax) continue;

        real jitter = K(0.3);
        real jx = ((drand(randSeed + 1u) - K(0.5)) * jitter) * speed;
        real jy = ((drand(randSeed + 2u) - K(0.5)) * jitter) * speed;
        real jz = ((drand(randSeed + 3u) - K(0.5)) * jitter) * speed;
        uint32_t ptype = DIFFUSE_FOAM;
        if (density < (real)gen->density_spray_max) ptype = DIFFUSE_SPRAY;
        real lifetime = K(999.0);
        if (ptype == DIFFUSE_FOAM)
            lifetime = gen->lifetime_mode
                           ? K(0.5) + K(2.5) * drand(randSeed + 4u)
                           : (real)gen->lifetime_foam * MINR(K(1.0), score / keth);
        orc_drec r;
        r.pos[0] = px; r.pos[1] = py; r.pos[2] = pz; r.pos[3] = K(0.0);
        r.vel[0] = vx + jx; r.vel[1] = vy + jy; r.vel[2] = vz + jz; r.vel[3] = K(0.0);
        r.ptype = ptype;
        r.alive = 1u;
        r.lifetime = lifetime;
        r.pad = K(0.0);
        pool[slot] = r;
    }
}

/* sample_grid_velocity (shaders.js:889-912). */
static v3 sample_grid_velocity(const orc_params *P, const real *cell_v, v3 pos) {
    real pa[3] = {pos.x, pos.y, pos.z};
    int32_t base[3];
    real w[3][3];
    for (int a = 0; a < 3; a++) {
        int32_t ci = cvt_i32(FLOORR(pa[a]));
        base[a] = ci - 1;
        quad_weights((pa[a] - (real)ci) - K(0.5), w[a]);
    }
    v3 v = mk3(K(0.0), K(0.0), K(0.0));
    for (int k = 0; k < 3; k++) {
        int32_t cz = base[2] + k;
        for (int j = 0; j < 3; j++) {
            int32_t cy = base[1] + j;
            for (int i = 0; i < 3; i++) {
                int32_t cx = base[0] + i;
                int32_t idx = cell_idx_checked(cx, cy, cz, P);
                if (idx < 0) continue;
                real ww = (w[0][i] * w[1][j]) * w[2][k];
                v.x += ww * cell_v[(uint32_t)idx * 3];
                v.y += ww * cell_v[(uint32_t)idx * 3 + 1];
                v.z += ww * cell_v[(uint32_t)idx * 3 + 2];
            }
        }
    }
    return v;
}

/* kill_particle (shaders.js:914-918): slot pushes happen in slot order. */
static inline void kill_particle(orc_drec *pool, uint32_t i, uint32_t *free_list, int32_t *top,
                                 uint32_t dmax) {
    pool[i].alive = 0u;
    int32_t slot = *top;
    *top = slot + 1;
    if (slot >= 0 && (uint32_t)slot < dmax) free_list[slot] = i;
}

/* reflect_axis (shaders.js:921-927), DIFFUSE_RESTITUTION = 0.3. */
static inline void reflect_axis(real *p, real *v, real lo, real hi) {
    real np = *p, nv = *v;
    if (np < lo) {
        np = lo + (lo - np);
        nv = (-nv) * K(0.3);
    } else if (np > hi) {
        np = hi - (np - hi);
        nv = (-nv) * K(0.3);
    }
    *p = MINR(MAXR(np, lo), hi);
    *v = nv;
}

/* WGSL_DIFFUSE_ADVECT (shaders.js:858-997), slots visited in index order. */
void orc_diffuse_advect(const orc_params *P, const orc_adv *adv, orc_drec *pool,
                        const real *cell_v, uint32_t *free_list, int32_t *top,
                        const float *wallp) {
    const real dt = (real)P->dt;
    const real g = (real)P->gravity;
    uint32_t nw = wall_count(wallp);
    for (uint32_t i = 0; i < adv->diffuse_max; i++) {
        if (pool[i].alive == 0u) continue;
        orc_drec p = pool[i];
        v3 pos3 = mk3(p.pos[0], p.pos[1], p.pos[2]);
        if (p.ptype == DIFFUSE_SPRAY) {
            p.vel[1] += g * dt;
            p.pos[0] = pos3.x + dt * p.vel[0];
            p.pos[1] = pos3.y + dt * p.vel[1];
            p.pos[2] = pos3.z + dt * p.vel[2];
            p.pos[3] = K(0.0);
        } else if (p.ptype == DIFFUSE_FOAM) {
            v3 gv = sample_grid_velocity(P, cell_v, pos3);
            p.pos[0] = pos3.x + dt * gv.x;
            p.pos[1] = pos3.y + dt * gv.y;
            p.pos[2] = pos3.z + dt * gv.z;
            p.pos[3] = K(0.0);
            p.lifetime -= dt;
            if (p.lifetime <= K(0.0)) {
                kill_particle(pool, i, free_list, top, adv->diffuse_max);
                continue;
            }
        } else if (p.ptype == DIFFUSE_BUBBLE) {
            v3 gv = sample_grid_velocity(P, cell_v, pos3);
            real drag = (real)adv->drag;
            p.vel[0] += (dt * drag) * (gv.x - p.vel[0]);
            p.vel[1] += dt * ((real)adv->buoyancy * (-g) + drag * (gv.y - p.vel[1]));
            p.vel[2] += (dt * drag) * (gv.z - p.vel[2]);
            p.pos[0] = pos3.x + dt * p.vel[0];
            p.pos[1] = pos3.y + dt * p.vel[1];
            p.pos[2] = pos3.z + dt * p.vel[2];
            p.pos[3] = K(0.0);
        }

        for (uint32_t wi = 0; wi < nw; wi++) {
            const float *w = wall_at(wallp, wi);
            v3 mn = wall_min3(w), mx = wall_max3(w);
            v3 lp = wall_to_local(mk3(p.pos[0], p.pos[1], p.pos[2]), w);
            if (!inside_open(lp, mn, mx)) continue;
            v3 nl;
            real sdf = wall_sdf_normal(lp, mn, mx, &nl);
            v3 nn = wall_to_world_dir(nl, w);
            p.pos[0] = p.pos[0] - nn.x * sdf;
            p.pos[1] = p.pos[1] - nn.y * sdf;
            p.pos[2] = p.pos[2] - nn.z * sdf;
            p.pos[3] = K(0.0);
            v3 wvel = wall_vel_at(mk3(p.pos[0], p.pos[1], p.pos[2]), w);
            v3 vrel = mk3(p.vel[0] - wvel.x, p.vel[1] - wvel.y, p.vel[2] - wvel.z);
            real vn = dot3(vrel, nn);
            if (vn < K(0.0)) {
                real e = K(1.0) + K(0.3);
                p.vel[0] = wvel.x + (vrel.x - (e * vn) * nn.x);
                p.vel[1] = wvel.y + (vrel.y - (e * vn) * nn.y);
                p.vel[2] = wvel.z + (vrel.z - (e * vn) * nn.z);
                p.vel[3] = K(0.0);
            }
        }

        if (p.pos[1] < (real)P->hard_min) {
            kill_particle(pool, i, free_list, top, adv->diffuse_max);
            continue;
        }
        reflect_axis(&p.pos[0], &p.vel[0], (real)P->hard_min, (real)P->hard_max_x);
        reflect_axis(&p.pos[1], &p.vel[1], (real)P->hard_min, (real)P->hard_max_y);
        reflect_axis(&p.pos[2], &p.vel[2], (real)P->hard_min, (real)P->hard_max_z);
        p.pos[3] = K(0.0);
        p.vel[3] = K(0.0);
        pool[i] = p;
    }
}

/* ---------------------------------------------------------------------------
 * Interaction.  The hand box is host arithmetic (JS numbers are doubles):
 * FluidGPU.updateHand (fluid-gpu.js:479-527) with lineBoxIntersect
 * (math.js:149-164). */
void orc_hand_bbox(const orc_params *P, orc_hand *h) {
    for (int a = 0; a < 3; a++) {
        h->bbox_min[a] = 0;
        h->bbox_max[a] = -1;
    }
    if (!h->is_on) return;
    const double pos[3] = {h->px, h->py, h->pz}, eye[3] = {h->ex, h->ey, h->ez};
    const double g[3] = {(double)P->grid_x, (double)P->grid_y, (double)P->grid_z};
    double d[3] = {pos[0] - eye[0], pos[1] - eye[1], pos[2] - eye[2]};
    double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (dl == 0.0) dl = 1.0;
    for (int a = 0; a < 3; a++) d[a] /= dl;
    const double R = h->radius;
    double tmin = -INFINITY, tmax = INFINITY;
    for (int a = 0; a < 3; a++) {
        const double lo = -R, hi = g[a] + R;
        if (fabs(d[a]) < 1e-9) {
            if (pos[a] < lo || pos[a] > hi) return;
            continue;
        }
        double t1 = (lo - pos[a]) / d[a], t2 = (hi - pos[a]) / d[a];
        if (t1 > t2) {
            const double tt = t1;
            t1 = t2;
            t2 = tt;
        }
        tmin = fmax(tmin, t1);
        tmax = fmin(tmax, t2);
        if (tmin > tmax) return;
    }
    for (int a = 0; a < 3; a++) {
        const double p0 = pos[a] + tmin * d[a], p1 = pos[a] + tmax * d[a];
        const double lo = floor(fmin(p0, p1) - R), hi = ceil(fmax(p0, p1) + R);
        h->bbox_min[a] = (int32_t)fmax(0.0, lo);
        h->bbox_max[a] = (int32_t)fmin(g[a] - 1.0, hi);
    }
}

/* WGSL_APPLY_HAND (shaders.js:410-432): inside the box, cells within `radius`
 * of the line through the hand along the eye ray get `vel * (1 - d/r)^2 *
 * strength` added to their (post-update) velocity. */
void orc_apply_hand(const orc_params *P, const orc_hand *h, real *cell_v) {
    if (!h->is_on) return;
    const int X = (int)P->grid_x, Y = (int)P->grid_y, Z = (int)P->grid_z;
    const v3 hp = mk3((real)h->px, (real)h->py, (real)h->pz);
    const v3 e0 = mk3(hp.x - (real)h->ex, hp.y - (real)h->ey, hp.z - (real)h->ez);
    const real len = SQRTR(dot3(e0, e0));
    const v3 rd = mk3(e0.x / len, e0.y / len, e0.z / len);
    const real radius = (real)h->radius;
    for (int cz = h->bbox_min[2]; cz <= h->bbox_max[2]; cz++)
        for (int cy = h->bbox_min[1]; cy <= h->bbox_max[1]; cy++)
            for (int cx = h->bbox_min[0]; cx <= h->bbox_max[0]; cx++) {
                if (cx < 0 || cy < 0 || cz < 0 || cx >= X || cy >= Y || cz >= Z) continue;
                const size_t idx = ((size_t)cz * Y + cy) * X + cx;
                const v3 to = mk3(((real)cx + K(0.5)) - hp.x, ((real)cy + K(0.5)) - hp.y,
                                  ((real)cz + K(0.5)) - hp.z);
                const real t0 = dot3(to, rd);
                const v3 perp = mk3(to.x - t0 * rd.x, to.y - t0 * rd.y, to.z - t0 * rd.z);
                const real d = SQRTR(dot3(perp, perp));
                if (d >= radius) continue;
                const real t = K(1.0) - d / radius;
                const real w = (t * t) * (real)h->strength;
                cell_v[3 * idx + 0] += (real)h->vx * w;
                cell_v[3 * idx + 1] += (real)h->vy * w;
                cell_v[3 * idx + 2] += (real)h->vz * w;
            }
}

/* ray_perp_dist (shaders.js:80-84). */
real orc_ray_perp_dist(const real *p, const orc_ray *r) {
    const v3 to = mk3(p[0] - (real)r->ox, p[1] - (real)r->oy, p[2] - (real)r->oz);
    const v3 d = mk3((real)r->dx, (real)r->dy, (real)r->dz);
    const real t = dot3(to, d);
    const v3 perp = mk3(to.x - t * d.x, to.y - t * d.y, to.z - t * d.z);
    return SQRTR(dot3(perp, perp));
}

/* WGSL_RAYCAST (shaders.js:557-577): the smallest positive distance along the
 * ray among particles within `radius` of its axis.  Returns whether any. */
int orc_raycast(const real *pos4, uint32_t n, const orc_ray *r, real *t_out) {
    int hit = 0;
    real best = K(0.0);
    for (uint32_t p = 0; p < n; p++) {
        const real *q = pos4 + 4 * (size_t)p;
        const v3 to = mk3(q[0] - (real)r->ox, q[1] - (real)r->oy, q[2] - (real)r->oz);
        const real t = dot3(to, mk3((real)r->dx, (real)r->dy, (real)r->dz));
        if (t <= K(0.0)) continue;
        if (orc_ray_perp_dist(q, r) >= (real)r->radius) continue;
        if (!hit || t < best) best = t;
        hit = 1;
    }
    *t_out = best;
    return hit;
}

/* WGSL_COMPACT (shaders.js:521-547): keep[p] = 1 for particles outside the
 * cylinder.  Returns the number kept. */
uint32_t orc_delete_mask(const real *pos4, uint32_t n, const orc_ray *r, uint8_t *keep) {
    uint32_t k = 0;
    for (uint32_t p = 0; p < n; p++) {
        keep[p] = orc_ray_perp_dist(pos4 + 4 * (size_t)p, r) < (real)r->radius ? 0 : 1;
        k += keep[p];
    }
    return k;
}

/* FluidGPU.simFrame with SUBSTEPS = 1 (fluid-gpu.js:648-684) followed by
 * DiffuseGPU.step (diffuse-gpu.js:148-166): seed advances before generate. */
void orc_step(const orc_params *P, orc_state *S, uint32_t nsteps) {
    for (uint32_t s = 0; s < nsteps; s++) {
        orc_clear(P, S->cell_mv, S->cell_mass);
        orc_compute_weights(P, S->pos4, S->bw);
        orc_p2g_mass(P, S->mass, S->cell_mass, S->bw);
        orc_decode_mass(P, S->cell_mass, S->cell_mass_f);
        orc_p2g_mom(P, S->vel4, S->affine, S->mass, S->cell_mass_f, S->cell_mv, S->density, S->bw);
        orc_update_grid(P, S->cell_mv, S->cell_mass, S->cell_v, S->wallp);
        if (S->hand) orc_apply_hand(P, S->hand, S->cell_v);
        orc_g2p(P, S->pos4, S->vel4, S->affine, S->cell_v, S->bw, S->wallp);
        if (S->pool) {
            S->gen.seed = S->gen.seed + 1u;
            orc_diffuse_generate(P, &S->gen, S->pos4, S->vel4, S->mass, S->density,
                                 S->cell_mass_f, S->pool, S->free_list, S->top,
                                 S->adv.diffuse_max);
            orc_diffuse_advect(P, &S->adv, S->pool, S->cell_v, S->free_list, S->top, S->wallp);
        }
    }
}

/* ---------------------------------------------------------------------------
 * Single helpers exported for the closed-form pins of the tests: each wraps
 * the function the kernels above call, unchanged. */
void orc_wall_to_local(const real p[3], const float *wall, real out[3]) {
    v3 r = wall_to_local(mk3(p[0], p[1], p[2]), wall);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_wall_to_world_dir(const real v[3], const float *wall, real out[3]) {
    v3 r = wall_to_world_dir(mk3(v[0], v[1], v[2]), wall);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_wall_vel_at(const real p[3], const float *wall, real out[3]) {
    v3 r = wall_vel_at(mk3(p[0], p[1], p[2]), wall);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
real orc_wall_sdf_normal(const real lp[3], const real mn[3], const real mx[3], real n[3]) {
    v3 nn;
    real d = wall_sdf_normal(mk3(lp[0], lp[1], lp[2]), mk3(mn[0], mn[1], mn[2]),
                             mk3(mx[0], mx[1], mx[2]), &nn);
    n[0] = nn.x; n[1] = nn.y; n[2] = nn.z;
    return d;
}
int orc_estimate_normal(const orc_params *P, const real *cell_mass_f, const real p[3], real n[3]) {
    v3 nn = mk3(K(0.0), K(0.0), K(0.0));
    int ok = estimate_normal(P, cell_mass_f, p[0], p[1], p[2], &nn);
    n[0] = nn.x; n[1] = nn.y; n[2] = nn.z;
    return ok;
}
real orc_spawn_score(const orc_params *P, const orc_gen *gen, const real *cell_mass_f,
                     const real pos[3], const real vel[3], real mass, real density, real terms[4]) {
    return spawn_score(P, gen, cell_mass_f, mk3(pos[0], pos[1], pos[2]),
                       mk3(vel[0], vel[1], vel[2]), mass, density, terms);
}
