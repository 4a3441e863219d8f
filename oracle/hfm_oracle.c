/* hfm_oracle.c — CPU restatement of the Heatflow-MPM air/heat substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP library, not part of it: only
 * tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py load it.  It is a
 * serial C restatement of the reference's Warp kernels, built twice:
 *   liboracle_f32.so  (REAL = float)   the reference's own arithmetic (fp32 storage and math,
 *                                      no FMA contraction: compile with -ffp-contract=off)
 *   liboracle_f64.so  (REAL = double)  the same discretisation in fp64, used as the
 *                                      high-precision truth for error statements.
 *
 * Every function cites the reference file:line it restates.  The conjugate-gradient driver
 * restates warp.optim.linear.cg (warp-lang >= 1.7, README tested 1.17; not vendored in the
 * reference): absolute tolerance max(tol*||b||, tol), Jacobi preconditioner, iterations in
 * pairs, a device-side freeze (alpha = beta = 0 once ||r||^2 <= atol^2) and a host check
 * every check_every iterations, so the returned count is a multiple of check_every or the
 * iteration cap.  Dot products: fp32 products accumulated in double and rounded to REAL
 * (warp's fp32 reduction order is unspecified).
 *
 * The pressure rhs mean of grid.py:239 is numpy's float32 mean (pairwise float32 sum, then
 * float32(sum / intp(N))); it is restated exactly for the fp32 build.
 *
 * Loops over cells, faces and vector entries run with OpenMP (every iteration writes its own
 * element, so the result does not depend on the thread count); sums (dot products, the numpy
 * mean) and the particle scatter stay serial in their fixed order.  orc_set_threads sets the
 * thread count (1 for the single-core baseline).
 *
 * Storage rounding (orc_set_storage_rounding, off by default) is not part of the restatement:
 * it rounds the stored velocity and/or temperature to BF16 or FP16 after every write, or traces
 * departure points through a BF16 copy of the velocity, so tests can measure which
 * low-precision storage keeps a run within a tolerance of the fp64 truth.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/hfm.h"

#ifdef ORC_F64
typedef double real;
#define FLOORR floor
#define SQRTR sqrt
#else
typedef float real;
#define FLOORR floorf
#define SQRTR sqrtf
#endif
#define R(x) ((real)(x))

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------------------------
 * MAC lattice helpers (fields.py:5-116)
 * ------------------------------------------------------------------------------------------ */

static inline long lidx(int ny, int nz, int i, int j, int k) { return ((long)i * ny + j) * nz + k; }

/* fields.py:13-15 */
static inline int clampi(int i, int lo, int hi) {
    int a = i > lo ? i : lo;
    return a < hi ? a : hi;
}

static const real OFF_CELL[3] = {R(0.5), R(0.5), R(0.5)};
static const real OFF_X[3] = {R(0.0), R(0.5), R(0.5)}; /* fields.py:86-87 */
static const real OFF_Y[3] = {R(0.5), R(0.0), R(0.5)}; /* fields.py:88-89 */
static const real OFF_Z[3] = {R(0.5), R(0.5), R(0.0)}; /* fields.py:90 */

/* fields.py:5-10: g = pos/h - off; i = int(floor(g)); f = g - float(i) */
static inline void base_index(const real pos[3], const real off[3], real h, int bi[3], real f[3]) {
    for (int c = 0; c < 3; ++c) {
        real g = pos[c] / h - off[c];
        bi[c] = (int)FLOORR(g);
        f[c] = g - (real)bi[c];
    }
}

/* fields.py:18-22 */
static inline real weight(real f, int d) { return d == 1 ? f : R(1.0) - f; }

/* fields.py:25-36: trilinear sample, clamped indices, weights from the unclamped fraction */
static real sample(const real* a, const real pos[3], const real off[3], real h, int nx, int ny, int nz) {
    int b[3];
    real f[3];
    base_index(pos, off, h, b, f);
    real s = R(0.0);
    for (int dx = 0; dx < 2; ++dx) {
        int ix = clampi(b[0] + dx, 0, nx - 1);
        for (int dy = 0; dy < 2; ++dy) {
            int iy = clampi(b[1] + dy, 0, ny - 1);
            for (int dz = 0; dz < 2; ++dz) {
                int iz = clampi(b[2] + dz, 0, nz - 1);
                real w = weight(f[0], dx) * weight(f[1], dy);
                w = w * weight(f[2], dz);
                s += w * a[lidx(ny, nz, ix, iy, iz)];
            }
        }
    }
    return s;
}

/* fields.py:39-48: the adjoint of sample (the reference uses atomics; serial here) */
static void scatter(real* a, const real pos[3], const real off[3], real h, int nx, int ny, int nz, real val) {
    int b[3];
    real f[3];
    base_index(pos, off, h, b, f);
    for (int dx = 0; dx < 2; ++dx) {
        int ix = clampi(b[0] + dx, 0, nx - 1);
        for (int dy = 0; dy < 2; ++dy) {
            int iy = clampi(b[1] + dy, 0, ny - 1);
            for (int dz = 0; dz < 2; ++dz) {
                int iz = clampi(b[2] + dz, 0, nz - 1);
                real w = weight(f[0], dx) * weight(f[1], dy);
                w = w * weight(f[2], dz);
                a[lidx(ny, nz, ix, iy, iz)] += w * val;
            }
        }
    }
}

static void backtrace(const real* u, const real* v, const real* w, const real x[3], real h, int n, real dt,
                      real out[3]);

/* fields.py:51-65: min and max of the eight lattice values around pos (the MacCormack limiter) */
static void value_range(const real* a, const real pos[3], const real off[3], real h, int nx, int ny, int nz,
                        real* lo, real* hi) {
    int b[3];
    real f[3];
    base_index(pos, off, h, b, f);
    real l = R(1.0e30), u = R(-1.0e30);
    for (int dx = 0; dx < 2; ++dx) {
        int ix = clampi(b[0] + dx, 0, nx - 1);
        for (int dy = 0; dy < 2; ++dy) {
            int iy = clampi(b[1] + dy, 0, ny - 1);
            for (int dz = 0; dz < 2; ++dz) {
                int iz = clampi(b[2] + dz, 0, nz - 1);
                real v = a[lidx(ny, nz, ix, iy, iz)];
                l = v < l ? v : l;
                u = v > u ? v : u;
            }
        }
    }
    *lo = l;
    *hi = u;
}

/* Limited MacCormack value at a point x of a lattice (HFM_FLAG_MACCORMACK; the reference prepares
 * value_range, fields.py:51-65, but advects semi-Lagrangian only): with fwd the semi-Lagrangian
 * value at x (field `fwd_field`), start the start field's value at x, d the departure point
 * (RK2 with dt) and e the point traced forward (RK2 with -dt):
 *   clamp(fwd + (start - fwd_field(e)) / 2, value_range(start_field, d)). */
static real maccormack(const real* u, const real* v, const real* w, const real* start_field, const real* fwd_field,
                       real fwd, real start, const real x[3], const real off[3], real h, int n, real dt, int nx,
                       int ny, int nz) {
    real d[3], e[3];
    backtrace(u, v, w, x, h, n, dt, d);
    backtrace(u, v, w, x, h, n, -dt, e);
    real back = sample(fwd_field, e, off, h, nx, ny, nz);
    real val = fwd + R(0.5) * (start - back);
    real lo, hi;
    value_range(start_field, d, off, h, nx, ny, nz, &lo, &hi);
    return val < lo ? lo : (val > hi ? hi : val);
}

/* fields.py:68-72 */
static void sample_velocity(const real* u, const real* v, const real* w, const real pos[3], real h, int n,
                            real vel[3]) {
    vel[0] = sample(u, pos, OFF_X, h, n + 1, n, n);
    vel[1] = sample(v, pos, OFF_Y, h, n, n + 1, n);
    vel[2] = sample(w, pos, OFF_Z, h, n, n, n + 1);
}

/* fields.py:111-116: midpoint RK2 departure point (also inlined in heat.py:38-41) */
static void backtrace(const real* u, const real* v, const real* w, const real x[3], real h, int n, real dt,
                      real out[3]) {
    real v1[3], x1[3], v2[3];
    sample_velocity(u, v, w, x, h, n, v1);
    real hd = R(0.5) * dt;
    for (int c = 0; c < 3; ++c) x1[c] = x[c] - hd * v1[c];
    sample_velocity(u, v, w, x1, h, n, v2);
    for (int c = 0; c < 3; ++c) out[c] = x[c] - dt * v2[c];
}

/* fields.py:75-81 */
static void face_position(int axis, int i, int j, int k, real h, real x[3]) {
    if (axis == 0) {
        x[0] = (real)i * h;
        x[1] = ((real)j + R(0.5)) * h;
        x[2] = ((real)k + R(0.5)) * h;
    } else if (axis == 1) {
        x[0] = ((real)i + R(0.5)) * h;
        x[1] = (real)j * h;
        x[2] = ((real)k + R(0.5)) * h;
    } else {
        x[0] = ((real)i + R(0.5)) * h;
        x[1] = ((real)j + R(0.5)) * h;
        x[2] = (real)k * h;
    }
}

static const real* face_offset(int axis) { return axis == 0 ? OFF_X : (axis == 1 ? OFF_Y : OFF_Z); }

static void face_dims(int axis, int n, int d[3]) {
    d[0] = n + (axis == 0);
    d[1] = n + (axis == 1);
    d[2] = n + (axis == 2);
}

/* ------------------------------------------------------------------------------------------
 * Reductions
 * ------------------------------------------------------------------------------------------ */

static real dotr(const real* a, const real* b, long N) {
    double s = 0.0;
    for (long t = 0; t < N; ++t) s += (double)(a[t] * b[t]);
    return (real)s;
}

/* numpy's float32 pairwise sum (numpy/_core/src/umath/loops_utils.h.src, PW_BLOCKSIZE 128) */
static float pairwise_sum_f32(const float* a, long n) {
    if (n < 8) {
        float res = 0.f;
        for (long i = 0; i < n; ++i) res += a[i];
        return res;
    } else if (n <= 128) {
        float r[8];
        for (int q = 0; q < 8; ++q) r[q] = a[q];
        long i;
        for (i = 8; i < n - (n % 8); i += 8)
            for (int q = 0; q < 8; ++q) r[q] += a[i + q];
        float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res += a[i];
        return res;
    } else {
        long n2 = n / 2;
        n2 -= n2 % 8;
        return pairwise_sum_f32(a, n2) + pairwise_sum_f32(a + n2, n - n2);
    }
}

/* np.add.reduce of a contiguous float32 array: the reduction runs over buffer-sized chunks of
 * 8192 elements (numpy's default bufsize), each summed pairwise and added in order to a float32
 * accumulator that starts at 0. */
ORC_API float orc_numpy_sum_f32(const float* a, long n) {
    float s = 0.f;
    for (long c = 0; c < n; c += 8192) {
        long m = n - c < 8192 ? n - c : 8192;
        s += pairwise_sum_f32(a + c, m);
    }
    return s;
}

/* np.float32 array .mean(): float32(sum / intp(N)); the quotient is formed in double because
 * numpy promotes float32 / intp to float64. */
ORC_API float orc_numpy_mean_f32(const float* a, long n) {
    float s = orc_numpy_sum_f32(a, n);
    return (float)((double)s / (double)n);
}

/* ------------------------------------------------------------------------------------------
 * Scene state (AirGrid grid.py:181-196, HeatField heat.py:178-187, HeatflowMPM solver.py:41-61)
 * ------------------------------------------------------------------------------------------ */

typedef struct orc_scene {
    hfm_params prm;
    int n;
    long N, F[3];
    real h;
    /* air */
    real *u[3], *scr[3], *af[3], *q[3], *kf[3], *mf[3];
    real *p, *rhs, *alpha_s, *pdiag;
    int has_static;
    real *st_alpha, *st_k[3], *st_m[3];
    /* heat */
    real *T, *T2, *Tw, *kap, *halpha, *tsum, *tfix, *hrhs, *hdiag;
    int32_t *fixed, *src;
    double rate, t_cap;
    /* particles */
    long P;
    real *pos, *vel, *temp, *melt, *yield, *visc;
    /* CG scratch */
    real *cr, *cz, *cp, *cAp;
    hfm_stats st;
    /* storage rounding of the velocity and the temperature (0 off, 1 BF16, 2 FP16) and departure
       points traced through a BF16 copy of the velocity (orc_set_storage_rounding) */
    int round_u, round_t, bt_bf16;
} orc_scene;

/* x rounded to BF16 (round to nearest even on the fp32 bit pattern) or FP16 (11 significant
   bits, subnormal spacing 2^-24, overflow to infinity), returned as a double */
static double store_round(double v, int mode) {
    if (mode == 1) {
        float f = (float)v;
        uint32_t u;
        memcpy(&u, &f, 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        u &= 0xFFFF0000u;
        memcpy(&f, &u, 4);
        return f;
    }
    if (mode == 2) {
        double a = fabs((double)(float)v);
        if (a == 0.0) return 0.0;
        int e;
        frexp(a, &e);
        double q = (e - 1) >= -14 ? ldexp(1.0, e - 11) : ldexp(1.0, -24);
        double r = nearbyint(a / q) * q;
        if (r > 65504.0) r = INFINITY;
        return v < 0 ? -r : r;
    }
    return v;
}

static void store_round_array(real* a, long n, int mode) {
    if (!mode) return;
    for (long t = 0; t < n; ++t) a[t] = (real)store_round((double)a[t], mode);
}

static real* zalloc(long n) { return (real*)calloc((size_t)(n > 0 ? n : 1), sizeof(real)); }

ORC_API orc_scene* orc_create(const hfm_params* prm) {
    orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
    s->prm = *prm;
    int n = s->n = prm->n;
    s->N = (long)n * n * n;
    for (int a = 0; a < 3; ++a) s->F[a] = (long)(n + 1) * n * n;
    s->h = (real)prm->h;
    for (int a = 0; a < 3; ++a) {
        s->u[a] = zalloc(s->F[a]);
        s->scr[a] = zalloc(s->F[a]);
        s->af[a] = zalloc(s->F[a]);
        s->q[a] = zalloc(s->F[a]);
        s->kf[a] = zalloc(s->F[a]);
        s->mf[a] = zalloc(s->F[a]);
        s->st_k[a] = zalloc(s->F[a]);
        s->st_m[a] = zalloc(s->F[a]);
    }
    long N = s->N;
    s->p = zalloc(N);
    s->rhs = zalloc(N);
    s->alpha_s = zalloc(N);
    s->pdiag = zalloc(N);
    s->st_alpha = zalloc(N);
    s->T = zalloc(N);
    s->T2 = zalloc(N);
    s->Tw = zalloc(N);
    s->kap = zalloc(N);
    s->halpha = zalloc(N);
    s->tsum = zalloc(N);
    s->tfix = zalloc(N);
    s->hrhs = zalloc(N);
    s->hdiag = zalloc(N);
    s->fixed = (int32_t*)calloc((size_t)N, sizeof(int32_t));
    s->src = (int32_t*)calloc((size_t)N, sizeof(int32_t));
    s->cr = zalloc(N);
    s->cz = zalloc(N);
    s->cp = zalloc(N);
    s->cAp = zalloc(N);
    /* heat.py:182-183: T, T2, Tw filled with the room temperature */
    for (long t = 0; t < N; ++t) s->T[t] = s->T2[t] = s->Tw[t] = (real)prm->t_amb;
    s->P = 0;
    return s;
}

ORC_API void orc_destroy(orc_scene* s) {
    if (!s) return;
    for (int a = 0; a < 3; ++a) {
        free(s->u[a]); free(s->scr[a]); free(s->af[a]); free(s->q[a]);
        free(s->kf[a]); free(s->mf[a]); free(s->st_k[a]); free(s->st_m[a]);
    }
    free(s->p); free(s->rhs); free(s->alpha_s); free(s->pdiag); free(s->st_alpha);
    free(s->T); free(s->T2); free(s->Tw); free(s->kap); free(s->halpha); free(s->tsum);
    free(s->tfix); free(s->hrhs); free(s->hdiag); free(s->fixed); free(s->src);
    free(s->cr); free(s->cz); free(s->cp); free(s->cAp);
    free(s->pos); free(s->vel); free(s->temp); free(s->melt); free(s->yield); free(s->visc);
    free(s);
}

/* grid.py:198-200 */
ORC_API void orc_set_static(orc_scene* s, const float* alpha, const float* kx, const float* ky, const float* kz,
                            const float* mx, const float* my, const float* mz) {
    const float* k[3] = {kx, ky, kz};
    const float* m[3] = {mx, my, mz};
    for (long t = 0; t < s->N; ++t) s->st_alpha[t] = (real)alpha[t];
    for (int a = 0; a < 3; ++a)
        for (long t = 0; t < s->F[a]; ++t) {
            s->st_k[a][t] = (real)k[a][t];
            s->st_m[a][t] = (real)m[a][t];
        }
    s->has_static = 1;
}

/* heat.py:147-151 */
static void apply_fixed(orc_scene* s, real* T) {
    #pragma omp parallel for schedule(static)
    for (long t = 0; t < s->N; ++t)
        if (s->fixed[t] == 1) T[t] = s->tfix[t];
}

/* heat.py:189-192 */
ORC_API void orc_set_fixed(orc_scene* s, const int32_t* fixed, const float* tfix) {
    for (long t = 0; t < s->N; ++t) {
        s->fixed[t] = fixed[t];
        s->tfix[t] = (real)tfix[t];
    }
    apply_fixed(s, s->T);
}

/* heat.py:194-196 */
ORC_API void orc_set_source(orc_scene* s, const int32_t* mask, double rate, double t_cap) {
    memcpy(s->src, mask, (size_t)s->N * sizeof(int32_t));
    s->rate = rate;
    s->t_cap = t_cap;
}

/* solver.py:59-61 plus the body arrays of mpm.py:54-68 (static body) */
ORC_API void orc_set_particles(orc_scene* s, long P, const float* pos, const float* vel, const float* temp) {
    free(s->pos); free(s->vel); free(s->temp); free(s->melt); free(s->yield); free(s->visc);
    s->P = P;
    s->pos = zalloc(3 * P);
    s->vel = zalloc(3 * P);
    s->temp = zalloc(P);
    s->melt = zalloc(P);
    s->yield = zalloc(P);
    s->visc = zalloc(P);
    for (long t = 0; t < 3 * P; ++t) {
        s->pos[t] = (real)pos[t];
        s->vel[t] = (real)vel[t];
    }
    for (long t = 0; t < P; ++t) {
        s->temp[t] = (real)temp[t];
        s->yield[t] = (real)s->prm.yield_solid; /* mpm.py:42 */
        s->visc[t] = (real)s->prm.mu_solid;     /* mpm.py:43 */
    }
}

ORC_API void orc_update_particles(orc_scene* s, const float* pos, const float* vel) {
    for (long t = 0; t < 3 * s->P; ++t) {
        s->pos[t] = (real)pos[t];
        s->vel[t] = (real)vel[t];
    }
}

static real* field_ptr(orc_scene* s, int f, long* size) {
    long N = s->N;
    switch (f) {
        case HFM_FIELD_T: *size = N; return s->T;
        case HFM_FIELD_T2: *size = N; return s->T2;
        case HFM_FIELD_TW: *size = N; return s->Tw;
        case HFM_FIELD_U: *size = s->F[0]; return s->u[0];
        case HFM_FIELD_V: *size = s->F[1]; return s->u[1];
        case HFM_FIELD_W: *size = s->F[2]; return s->u[2];
        case HFM_FIELD_P: *size = N; return s->p;
        case HFM_FIELD_RHS: *size = N; return s->rhs;
        case HFM_FIELD_ALPHA_S: *size = N; return s->alpha_s;
        case HFM_FIELD_ALPHA_P: *size = N; return s->halpha;
        case HFM_FIELD_TSUM: *size = N; return s->tsum;
        case HFM_FIELD_KAPPA: *size = N; return s->kap;
        case HFM_FIELD_AX: case HFM_FIELD_AY: case HFM_FIELD_AZ:
            *size = s->F[f - HFM_FIELD_AX]; return s->af[f - HFM_FIELD_AX];
        case HFM_FIELD_QX: case HFM_FIELD_QY: case HFM_FIELD_QZ:
            *size = s->F[f - HFM_FIELD_QX]; return s->q[f - HFM_FIELD_QX];
        case HFM_FIELD_KX: case HFM_FIELD_KY: case HFM_FIELD_KZ:
            *size = s->F[f - HFM_FIELD_KX]; return s->kf[f - HFM_FIELD_KX];
        case HFM_FIELD_MX: case HFM_FIELD_MY: case HFM_FIELD_MZ:
            *size = s->F[f - HFM_FIELD_MX]; return s->mf[f - HFM_FIELD_MX];
        case HFM_FIELD_PDIAG: *size = N; return s->pdiag;
        case HFM_FIELD_HRHS: *size = N; return s->hrhs;
        case HFM_FIELD_HDIAG: *size = N; return s->hdiag;
        default: *size = 0; return NULL;
    }
}

ORC_API int orc_get_field(orc_scene* s, int f, double* out) {
    long size;
    real* a = field_ptr(s, f, &size);
    if (!a) return -1;
    for (long t = 0; t < size; ++t) out[t] = (double)a[t];
    return 0;
}

ORC_API int orc_set_field(orc_scene* s, int f, const double* in) {
    long size;
    real* a = field_ptr(s, f, &size);
    if (!a) return -1;
    for (long t = 0; t < size; ++t) a[t] = (real)in[t];
    return 0;
}

ORC_API void orc_get_particles(orc_scene* s, double* temp, double* melt, double* yield, double* visc) {
    for (long t = 0; t < s->P; ++t) {
        if (temp) temp[t] = s->temp[t];
        if (melt) melt[t] = s->melt[t];
        if (yield) yield[t] = s->yield[t];
        if (visc) visc[t] = s->visc[t];
    }
}

ORC_API void orc_get_stats(orc_scene* s, hfm_stats* out) { *out = s->st; }

/* ------------------------------------------------------------------------------------------
 * Linear operators
 * ------------------------------------------------------------------------------------------ */

/* grid.py:94-115: out = alpha * sum_f a_f (p_c - p_nb) / h^2 + beta * y, missing neighbours skipped */
static void laplacian_apply(int n, real h, real* const af[3], const real* p, const real* y, real* out, real alpha,
                            real beta) {
    long nn = (long)n * n;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                long t = ((long)i * n + j) * n + k;
                real pc = p[t];
                real acc = R(0.0);
                if (i > 0) acc += af[0][lidx(n, n, i, j, k)] * (pc - p[t - nn]);
                if (i < n - 1) acc += af[0][lidx(n, n, i + 1, j, k)] * (pc - p[t + nn]);
                if (j > 0) acc += af[1][lidx(n + 1, n, i, j, k)] * (pc - p[t - n]);
                if (j < n - 1) acc += af[1][lidx(n + 1, n, i, j + 1, k)] * (pc - p[t + n]);
                if (k > 0) acc += af[2][lidx(n, n + 1, i, j, k)] * (pc - p[t - 1]);
                if (k < n - 1) acc += af[2][lidx(n, n + 1, i, j, k + 1)] * (pc - p[t + 1]);
                real yt = y[t];
                out[t] = alpha * acc / (h * h) + beta * yt;
            }
}

/* grid.py:118-137 */
static void jacobi_diag(int n, real h, real* const af[3], real* d) {
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                long t = ((long)i * n + j) * n + k;
                real s = R(0.0);
                if (i > 0) s += af[0][lidx(n, n, i, j, k)];
                if (i < n - 1) s += af[0][lidx(n, n, i + 1, j, k)];
                if (j > 0) s += af[1][lidx(n + 1, n, i, j, k)];
                if (j < n - 1) s += af[1][lidx(n + 1, n, i, j + 1, k)];
                if (k > 0) s += af[2][lidx(n, n + 1, i, j, k)];
                if (k < n - 1) s += af[2][lidx(n, n + 1, i, j, k + 1)];
                real m = s > R(1.0e-12) ? s : R(1.0e-12);
                d[t] = (h * h) / m;
            }
}

/* heat.py:50-52 */
static inline real face_kappa(real ka, real kb) { return R(2.0) * ka * kb / (ka + kb + R(1.0e-12)); }

static const int NB_DI[6] = {-1, 1, 0, 0, 0, 0};
static const int NB_DJ[6] = {0, 0, -1, 1, 0, 0};
static const int NB_DK[6] = {0, 0, 0, 0, -1, 1};

/* heat.py:55-91 */
static real neighbour_terms(const real* x, const real* kap, const int32_t* fixed, const real* tfix, long t, int i,
                            int j, int k, int n, int use_fixed) {
    real kc = kap[t];
    real acc = R(0.0);
    real xc = x[t];
    for (int s = 0; s < 6; ++s) {
        int di = NB_DI[s], dj = NB_DJ[s], dk = NB_DK[s];
        int ii = i + di, jj = j + dj, kk = k + dk;
        if (ii < 0 || jj < 0 || kk < 0 || ii >= n || jj >= n || kk >= n) continue;
        long tn = t + (long)di * n * n + (long)dj * n + dk;
        real kf = face_kappa(kc, kap[tn]);
        if (fixed[tn] == 1) {
            if (use_fixed == 1)
                acc += kf * tfix[tn];
            else
                acc += kf * xc;
        } else if (use_fixed == 0) {
            acc += kf * (xc - x[tn]);
        }
    }
    return acc;
}

/* heat.py:94-105 */
static void heat_apply(int n, real h, real dt, const real* kap, const int32_t* fixed, const real* tfix, const real* x,
                       const real* y, real* out, real alpha, real beta) {
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                long t = ((long)i * n + j) * n + k;
                real yt = y[t];
                if (fixed[t] == 1) {
                    out[t] = alpha * x[t] + beta * yt;
                    continue;
                }
                real nt = neighbour_terms(x, kap, fixed, tfix, t, i, j, k, n, 0);
                out[t] = alpha * (x[t] + dt / (h * h) * nt) + beta * yt;
            }
}

/* heat.py:108-118 */
static void heat_rhs(int n, real h, real dt, const real* T, const real* kap, const int32_t* fixed, const real* tfix,
                     real* rhs) {
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                long t = ((long)i * n + j) * n + k;
                if (fixed[t] == 1) {
                    rhs[t] = tfix[t];
                    continue;
                }
                rhs[t] = T[t] + dt / (h * h) * neighbour_terms(T, kap, fixed, tfix, t, i, j, k, n, 1);
            }
}

/* heat.py:121-144 */
static void heat_diag(int n, real h, real dt, const real* kap, const int32_t* fixed, real* d) {
    long nn = (long)n * n;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                long t = ((long)i * n + j) * n + k;
                if (fixed[t] == 1) {
                    d[t] = R(1.0);
                    continue;
                }
                real kc = kap[t];
                real s = R(0.0);
                if (i > 0) s += face_kappa(kc, kap[t - nn]);
                if (i < n - 1) s += face_kappa(kc, kap[t + nn]);
                if (j > 0) s += face_kappa(kc, kap[t - n]);
                if (j < n - 1) s += face_kappa(kc, kap[t + n]);
                if (k > 0) s += face_kappa(kc, kap[t - 1]);
                if (k < n - 1) s += face_kappa(kc, kap[t + 1]);
                d[t] = R(1.0) / (R(1.0) + dt / (h * h) * s);
            }
}

/* ------------------------------------------------------------------------------------------
 * Jacobi-preconditioned CG with warp.optim.linear.cg semantics
 * ------------------------------------------------------------------------------------------ */

typedef struct opctx {
    int kind; /* 0 pressure, 1 heat */
    int n;
    real h, dt;
    real* const* af;
    const real *kap, *tfix;
    const int32_t* fixed;
} opctx;

static void op_apply(const opctx* o, const real* x, const real* y, real* out, real alpha, real beta) {
    if (o->kind == 0)
        laplacian_apply(o->n, o->h, o->af, x, y, out, alpha, beta);
    else
        heat_apply(o->n, o->h, o->dt, o->kap, o->fixed, o->tfix, x, y, out, alpha, beta);
}

typedef struct cgres {
    int exact, reported, converged;
    double rr, atol;
} cgres;

/* Count the warp driver returns (_run_solver_loop): cycles of 2 iterations, stop when the
 * count reaches maxiter, or at a check (count % check_every < 2) once converged. */
static int warp_reported_count(int exact, int converged, int maxiter, int check_every) {
    int cur = 0;
    for (;;) {
        cur += 2;
        if (cur >= maxiter) break;
        if ((cur % check_every) < 2 && converged && exact <= cur) break;
    }
    return cur;
}

static void pcg(const opctx* op, const real* diag, const real* b, real* x, long N, double tol, int maxiter,
                int check_every, real* r, real* z, real* p, real* Ap, cgres* res) {
    /* atol = max(tol * ||b||, tol) (warp: _get_tolerances + _initialize_residual_and_tolerance) */
    real bb = dotr(b, b, N);
    double atol = tol * sqrt((double)bb);
    if (atol < tol) atol = tol;
    real atol_sq = (real)(atol * atol);
    res->atol = atol;
    op_apply(op, x, b, r, R(-1.0), R(1.0)); /* r = b - A x */
    real rr = dotr(r, r, N);
    if (rr <= atol_sq) {
        res->exact = 0;
        res->reported = 0;
        res->converged = 1;
        res->rr = rr;
        return;
    }
    #pragma omp parallel for schedule(static)
    for (long t = 0; t < N; ++t) z[t] = diag[t] * r[t];
    real rz_new = dotr(r, z, N);
    memcpy(p, z, (size_t)N * sizeof(real));
    int maxeff = maxiter + (maxiter & 1);
    int it = 0, converged = 0;
    /* Breakdown stop of the build (not in warp, which runs on to maxiter): an iteration whose p.Ap
     * is not positive or whose alpha is not finite takes alpha = 0 (x and r keep the last iterate),
     * and the solve ends with converged = -1 there, as it does where r.z of the previous iteration is
     * not positive, beta is not finite or ||r||^2 is not finite.  A converging solve never meets
     * these conditions. */
    while (it < maxeff) {
        real rz_old = rz_new;
        op_apply(op, p, Ap, Ap, R(1.0), R(0.0));
        real pAp = dotr(p, Ap, N);
        int brk = 0;
        real alpha = R(0.0);
        if (rr > atol_sq) {
            alpha = rz_old / pAp;
            if (!(pAp > R(0.0)) || !isfinite(alpha)) {
                alpha = R(0.0);
                brk = 1;
            }
        }
        #pragma omp parallel for schedule(static)
        for (long t = 0; t < N; ++t) {
            x[t] = x[t] + alpha * p[t];
            r[t] = r[t] - alpha * Ap[t];
        }
        rr = dotr(r, r, N);
        #pragma omp parallel for schedule(static)
        for (long t = 0; t < N; ++t) z[t] = diag[t] * r[t];
        rz_new = dotr(r, z, N);
        real beta = R(0.0);
        if (rr > atol_sq) {
            beta = rz_new / rz_old;
            if (!(rz_old > R(0.0)) || !isfinite(beta)) brk = 1;
        }
        if (!isfinite(rr)) brk = 1;
        #pragma omp parallel for schedule(static)
        for (long t = 0; t < N; ++t) p[t] = z[t] + beta * p[t];
        ++it;
        /* past this point alpha = beta = 0: x and r are frozen, so stopping is exact */
        if (rr <= atol_sq) {
            converged = 1;
            break;
        }
        if (brk) {
            converged = -1;
            break;
        }
    }
    res->exact = it;
    res->converged = converged;
    res->reported = warp_reported_count(it, converged, maxiter, check_every);
    res->rr = rr;
}

/* ------------------------------------------------------------------------------------------
 * Substep stages (solver.py:63-86, grid.py:219-251, heat.py:12-47,147-174,198-210)
 * ------------------------------------------------------------------------------------------ */

enum {
    ORC_ST_CLEAR = 0,
    ORC_ST_P2G,
    ORC_ST_ADD_STATIC,
    ORC_ST_BLEND,
    ORC_ST_FIXED_TW,
    ORC_ST_SOURCE,
    ORC_ST_ADVECT_T,
    ORC_ST_CONDUCTIVITY,
    ORC_ST_DIFFUSE,
    ORC_ST_PARTICLE_UPDATE,
    ORC_ST_SWAP_T,
    ORC_ST_FACE_FRACTION,
    ORC_ST_ADVECT_FACES,
    ORC_ST_BUOYANCY,
    ORC_ST_RELAX,
    ORC_ST_DIVERGENCE,
    ORC_ST_MEAN,
    ORC_ST_JACOBI,
    ORC_ST_PRESSURE_CG,
    ORC_ST_PROJECT,
    ORC_ST_COUNT
};

/* grid.py:202-204 and solver.py:67-68 */
static void st_clear(orc_scene* s) {
    memset(s->alpha_s, 0, (size_t)s->N * sizeof(real));
    for (int a = 0; a < 3; ++a) {
        memset(s->q[a], 0, (size_t)s->F[a] * sizeof(real));
        memset(s->kf[a], 0, (size_t)s->F[a] * sizeof(real));
        memset(s->mf[a], 0, (size_t)s->F[a] * sizeof(real));
    }
    memset(s->halpha, 0, (size_t)s->N * sizeof(real));
    memset(s->tsum, 0, (size_t)s->N * sizeof(real));
}

/* solver.py:19-38 (k_scatter_particles) then heat.py:12-18 (k_scatter_heat) */
static void st_p2g(orc_scene* s) {
    int n = s->n;
    real h = s->h;
    real vol = (real)s->prm.vol, k0 = (real)s->prm.k_particle;
    for (long t = 0; t < s->P; ++t) {
        const real* x = s->pos + 3 * t;
        const real* vp = s->vel + 3 * t;
        real frac = vol / (h * h * h);
        scatter(s->alpha_s, x, OFF_CELL, h, n, n, n, frac);
        scatter(s->q[0], x, OFF_X, h, n + 1, n, n, frac * vp[0]);
        scatter(s->q[1], x, OFF_Y, h, n, n + 1, n, frac * vp[1]);
        scatter(s->q[2], x, OFF_Z, h, n, n, n + 1, frac * vp[2]);
        scatter(s->kf[0], x, OFF_X, h, n + 1, n, n, k0);
        scatter(s->kf[1], x, OFF_Y, h, n, n + 1, n, k0);
        scatter(s->kf[2], x, OFF_Z, h, n, n, n + 1, k0);
        scatter(s->mf[0], x, OFF_X, h, n + 1, n, n, k0 * vp[0]);
        scatter(s->mf[1], x, OFF_Y, h, n, n + 1, n, k0 * vp[1]);
        scatter(s->mf[2], x, OFF_Z, h, n, n, n + 1, k0 * vp[2]);
    }
    for (long t = 0; t < s->P; ++t) {
        const real* x = s->pos + 3 * t;
        real frac = vol / (h * h * h);
        scatter(s->halpha, x, OFF_CELL, h, n, n, n, frac);
        scatter(s->tsum, x, OFF_CELL, h, n, n, n, frac * s->temp[t]);
    }
}

/* grid.py:206-213 */
static void st_add_static(orc_scene* s) {
    if (!s->has_static) return;
    #pragma omp parallel for schedule(static)
    for (long t = 0; t < s->N; ++t) s->alpha_s[t] = s->alpha_s[t] + s->st_alpha[t];
    for (int a = 0; a < 3; ++a)
        #pragma omp parallel for schedule(static)
        for (long t = 0; t < s->F[a]; ++t) {
            s->kf[a][t] = s->kf[a][t] + s->st_k[a][t];
            s->mf[a][t] = s->mf[a][t] + s->st_m[a][t];
        }
}

/* heat.py:21-28 */
static void st_blend(orc_scene* s) {
    #pragma omp parallel for schedule(static)
    for (long t = 0; t < s->N; ++t) {
        if (s->halpha[t] > R(0.05))
            s->Tw[t] = s->tsum[t] / s->halpha[t];
        else
            s->Tw[t] = s->T[t];
    }
}

/* heat.py:154-158, launched only when rate > 0 (solver.py:76) */
/* Also the build's heater energy balance (hfm_stats.source_dT_sum): the rises T_new - T_old of the
 * source cells summed in fp64, serially in cell order. */
static void st_source(orc_scene* s, double dt) {
    s->st.source_dT_sum = 0.0;
    if (!(s->rate > 0)) return;
    real rate = (real)s->rate, tcap = (real)s->t_cap, d = (real)dt;
    double rise = 0.0;
    for (long t = 0; t < s->N; ++t) {
        if (s->src[t] == 1) {
            real T = s->Tw[t];
            real a = T + rate * d;
            real b = T > tcap ? T : tcap;
            s->Tw[t] = a < b ? a : b;
            rise += (double)(s->Tw[t] - T);
        }
    }
    s->st.source_dT_sum = rise;
}

static real* bf16_copy(const real* a, long n) {
    real* c = zalloc(n);
    for (long t = 0; t < n; ++t) c[t] = (real)store_round((double)a[t], 1);
    return c;
}

/* heat.py:31-41 */
static void st_advect_T(orc_scene* s, double dtd) {
    real* bu[3] = {s->u[0], s->u[1], s->u[2]};
    if (s->bt_bf16)
        for (int a = 0; a < 3; ++a) bu[a] = bf16_copy(s->u[a], s->F[a]);
    int n = s->n;
    real h = s->h, dt = (real)dtd;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                long t = ((long)i * n + j) * n + k;
                if (s->alpha_s[t] > R(0.5)) {
                    s->T2[t] = s->Tw[t];
                    continue;
                }
                real x[3] = {((real)i + R(0.5)) * h, ((real)j + R(0.5)) * h, ((real)k + R(0.5)) * h};
                real d[3];
                backtrace(bu[0], bu[1], bu[2], x, h, n, dt, d);
                s->T2[t] = sample(s->Tw, d, OFF_CELL, h, n, n, n);
            }
    if (s->bt_bf16)
        for (int a = 0; a < 3; ++a) free(bu[a]);
    if (s->prm.flags & HFM_FLAG_MACCORMACK) {   /* T2 holds the semi-Lagrangian field */
        real* mc = zalloc(s->N);
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                for (int k = 0; k < n; ++k) {
                    long t = ((long)i * n + j) * n + k;
                    if (s->alpha_s[t] > R(0.5)) {
                        mc[t] = s->T2[t];
                        continue;
                    }
                    real x[3] = {((real)i + R(0.5)) * h, ((real)j + R(0.5)) * h, ((real)k + R(0.5)) * h};
                    mc[t] = maccormack(s->u[0], s->u[1], s->u[2], s->Tw, s->T2, s->T2[t], s->Tw[t], x, OFF_CELL, h, n,
                                       dt, n, n, n);
                }
        memcpy(s->T2, mc, s->N * sizeof(real));
        free(mc);
    }
    store_round_array(s->T2, s->N, s->round_t);
}

/* heat.py:44-47 */
static void st_conductivity(orc_scene* s) {
    real ka = (real)s->prm.kappa_air, km = (real)s->prm.kappa_mat;
    #pragma omp parallel for schedule(static)
    for (long t = 0; t < s->N; ++t) {
        real c = s->alpha_s[t] / R(0.6);
        c = c < R(0.0) ? R(0.0) : (c > R(1.0) ? R(1.0) : c);
        s->kap[t] = ka + (km - ka) * c;
    }
}

/* heat.py:198-210 */
static void st_diffuse(orc_scene* s, double dtd) {
    int n = s->n;
    real h = s->h, dt = (real)dtd;
    heat_rhs(n, h, dt, s->T2, s->kap, s->fixed, s->tfix, s->hrhs);
    heat_diag(n, h, dt, s->kap, s->fixed, s->hdiag);
    opctx op = {1, n, h, dt, NULL, s->kap, s->tfix, s->fixed};
    cgres r;
    pcg(&op, s->hdiag, s->hrhs, s->T2, s->N, s->prm.heat_tol, s->prm.heat_max_iter, s->prm.heat_check_every, s->cr,
        s->cz, s->cp, s->cAp, &r);
    s->st.h_iters_total += r.exact;
    s->st.h_iters = r.reported;
    s->st.h_iters_exact = r.exact;
    s->st.h_converged = r.converged;
    s->st.h_rr = r.rr;
    s->st.h_atol = r.atol;
    store_round_array(s->T2, s->N, s->round_t);
}

/* heat.py:161-174 */
static void st_particle_update(orc_scene* s) {
    int n = s->n;
    real h = s->h;
    real t_sol = (real)s->prm.t_sol, t_liq = (real)s->prm.t_liq, ys = (real)s->prm.yield_solid;
    real mus = (real)s->prm.mu_solid, mul = (real)s->prm.mu_liquid, tmax = (real)s->prm.t_max;
    for (long t = 0; t < s->P; ++t) {
        const real* x = s->pos + 3 * t;
        real dT = sample(s->T2, x, OFF_CELL, h, n, n, n) - sample(s->Tw, x, OFF_CELL, h, n, n, n);
        real tp = s->temp[t] + dT;
        tp = tp < R(0.0) ? R(0.0) : (tp > tmax ? tmax : tp);
        s->temp[t] = tp;
        real f = (tp - t_sol) / (t_liq - t_sol);
        f = f < R(0.0) ? R(0.0) : (f > R(1.0) ? R(1.0) : f);
        s->melt[t] = f;
        s->yield[t] = ys * (R(1.0) - f);
        s->visc[t] = mus * (R(1.0) - f) + mul * f;
    }
}

/* solver.py:85 */
static void st_swap_T(orc_scene* s) {
    real* tmp = s->T;
    s->T = s->T2;
    s->T2 = tmp;
}

/* grid.py:46-67 (k_face_fraction, k_zero_wall) */
static void st_face_fraction(orc_scene* s) {
    int n = s->n;
    real amin = (real)s->prm.alpha_f_min;
    for (int axis = 0; axis < 3; ++axis) {
        int d[3];
        face_dims(axis, n, d);
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < d[0]; ++i)
            for (int j = 0; j < d[1]; ++j)
                for (int k = 0; k < d[2]; ++k) {
                    long f = lidx(d[1], d[2], i, j, k);
                    int idx = axis == 0 ? i : (axis == 1 ? j : k);
                    if (idx == 0 || idx == n) {
                        s->af[axis][f] = R(0.0);
                        continue;
                    }
                    long tc = ((long)i * n + j) * n + k;
                    long tl = axis == 0 ? tc - (long)n * n : (axis == 1 ? tc - n : tc - 1);
                    real a = R(0.5) * (s->alpha_s[tl] + s->alpha_s[tc]);
                    real v = R(1.0) - a;
                    v = v < amin ? amin : (v > R(1.0) ? R(1.0) : v);
                    s->af[axis][f] = v;
                }
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < d[0]; ++i)
            for (int j = 0; j < d[1]; ++j)
                for (int k = 0; k < d[2]; ++k) {
                    int idx = axis == 0 ? i : (axis == 1 ? j : k);
                    if (idx == 0 || idx == n) s->q[axis][lidx(d[1], d[2], i, j, k)] = R(0.0);
                }
    }
}

/* grid.py:25-33 and the swap of grid.py:224-227: all three axes read the old u, v, w */
static void st_advect_faces(orc_scene* s, double dtd) {
    real* bu[3] = {s->u[0], s->u[1], s->u[2]};
    if (s->bt_bf16)
        for (int a = 0; a < 3; ++a) bu[a] = bf16_copy(s->u[a], s->F[a]);
    int n = s->n;
    real h = s->h, dt = (real)dtd;
    for (int axis = 0; axis < 3; ++axis) {
        int d[3];
        face_dims(axis, n, d);
        const real* off = face_offset(axis);
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < d[0]; ++i)
            for (int j = 0; j < d[1]; ++j)
                for (int k = 0; k < d[2]; ++k) {
                    long f = lidx(d[1], d[2], i, j, k);
                    int idx = axis == 0 ? i : (axis == 1 ? j : k);
                    if (idx == 0 || idx == n) {
                        s->scr[axis][f] = R(0.0);
                        continue;
                    }
                    real x[3], dp[3];
                    face_position(axis, i, j, k, h, x);
                    backtrace(bu[0], bu[1], bu[2], x, h, n, dt, dp);
                    s->scr[axis][f] = sample(s->u[axis], dp, off, h, d[0], d[1], d[2]);
                }
    }
    if (s->bt_bf16)
        for (int a = 0; a < 3; ++a) free(bu[a]);
    if (s->prm.flags & HFM_FLAG_MACCORMACK) {   /* scr holds the semi-Lagrangian u, v, w; u the old ones */
        real* mc[3];
        for (int axis = 0; axis < 3; ++axis) {
            int d[3];
            face_dims(axis, n, d);
            const real* off = face_offset(axis);
            mc[axis] = zalloc(s->F[axis]);
            #pragma omp parallel for schedule(static)
            for (int i = 0; i < d[0]; ++i)
                for (int j = 0; j < d[1]; ++j)
                    for (int k = 0; k < d[2]; ++k) {
                        long f = lidx(d[1], d[2], i, j, k);
                        int idx = axis == 0 ? i : (axis == 1 ? j : k);
                        if (idx == 0 || idx == n) continue;   /* wall faces stay 0 */
                        real x[3];
                        face_position(axis, i, j, k, h, x);
                        mc[axis][f] = maccormack(s->u[0], s->u[1], s->u[2], s->u[axis], s->scr[axis], s->scr[axis][f],
                                                 s->u[axis][f], x, off, h, n, dt, d[0], d[1], d[2]);
                    }
        }
        for (int a = 0; a < 3; ++a) {
            memcpy(s->scr[a], mc[a], s->F[a] * sizeof(real));
            free(mc[a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        real* tmp = s->u[a];
        s->u[a] = s->scr[a];
        s->scr[a] = tmp;
        store_round_array(s->u[a], s->F[a], s->round_u);
    }
}

/* grid.py:36-43, launched only with buoyancy on (grid.py:228-229); reads the new T */
static void st_buoyancy(orc_scene* s, double dtd) {
    if (!s->prm.buoyancy) return;
    int n = s->n;
    real dt = (real)dtd, g = (real)s->prm.gravity, tamb = (real)s->prm.t_amb;
    real* w = s->u[2];
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 1; k < n; ++k) {
                long t = ((long)i * n + j) * n + k;
                real tf = R(0.5) * (s->T[t - 1] + s->T[t]);
                long f = lidx(n, n + 1, i, j, k);
                w[f] = w[f] + dt * g * (tf - tamb) / (tamb + R(273.15));
            }
}

/* grid.py:70-79 */
static void st_relax(orc_scene* s, double dtd) {
    int n = s->n;
    real h = s->h, dt = (real)dtd, rho = (real)s->prm.rho, amin = (real)s->prm.alpha_f_min;
    for (int axis = 0; axis < 3; ++axis) {
        int d[3];
        face_dims(axis, n, d);
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < d[0]; ++i)
            for (int j = 0; j < d[1]; ++j)
                for (int k = 0; k < d[2]; ++k) {
                    int idx = axis == 0 ? i : (axis == 1 ? j : k);
                    if (idx == 0 || idx == n) continue;
                    long f = lidx(d[1], d[2], i, j, k);
                    real a = s->af[axis][f];
                    real am = a > amin ? a : amin;
                    real c = R(1.0) / (rho * am * h * h * h);
                    s->u[axis][f] = (s->u[axis][f] + dt * c * s->mf[axis][f]) / (R(1.0) + dt * c * s->kf[axis][f]);
                }
    }
}

/* grid.py:82-91, scale rho/dt formed on the host in double (grid.py:237) */
static void divergence(orc_scene* s, real scale, real* rhs) {
    int n = s->n;
    real h = s->h;
    real *u = s->u[0], *v = s->u[1], *w = s->u[2];
    real *ax = s->af[0], *ay = s->af[1], *az = s->af[2];
    real *qx = s->q[0], *qy = s->q[1], *qz = s->q[2];
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                long x0 = lidx(n, n, i, j, k), x1 = lidx(n, n, i + 1, j, k);
                long y0 = lidx(n + 1, n, i, j, k), y1 = lidx(n + 1, n, i, j + 1, k);
                long z0 = lidx(n, n + 1, i, j, k), z1 = lidx(n, n + 1, i, j, k + 1);
                real div = ax[x1] * u[x1] + qx[x1] - ax[x0] * u[x0] - qx[x0] + ay[y1] * v[y1] + qy[y1] -
                           ay[y0] * v[y0] - qy[y0] + az[z1] * w[z1] + qz[z1] - az[z0] * w[z0] - qz[z0];
                div = div / h;
                rhs[((long)i * n + j) * n + k] = -scale * div;
            }
}

static void st_divergence(orc_scene* s, double dtd) { divergence(s, (real)(s->prm.rho / dtd), s->rhs); }

/* grid.py:238-241 */
static void st_mean(orc_scene* s) {
#ifdef ORC_F64
    double acc = 0.0;
    for (long t = 0; t < s->N; ++t) acc += s->rhs[t];
    real mean = acc / (double)s->N;
#else
    real mean = orc_numpy_mean_f32(s->rhs, s->N);
#endif
    s->st.rhs_mean = (double)mean;
    if (mean != R(0.0))
        for (long t = 0; t < s->N; ++t) s->rhs[t] = s->rhs[t] - mean;
}

static void st_jacobi(orc_scene* s) { jacobi_diag(s->n, s->h, s->af, s->pdiag); }

/* grid.py:243-249 */
static void st_pressure_cg(orc_scene* s) {
    opctx op = {0, s->n, s->h, R(0.0), s->af, NULL, NULL, NULL};
    cgres r;
    pcg(&op, s->pdiag, s->rhs, s->p, s->N, s->prm.cg_tol, s->prm.cg_max_iter, s->prm.cg_check_every, s->cr, s->cz,
        s->cp, s->cAp, &r);
    s->st.p_iters_total += r.exact;
    s->st.p_iters = r.reported;
    s->st.p_iters_exact = r.exact;
    s->st.p_converged = r.converged;
    s->st.p_rr = r.rr;
    s->st.p_atol = r.atol;
}

/* grid.py:140-153, dt/rho formed on the host in double (grid.py:251) */
static void st_project(orc_scene* s, double dtd) {
    int n = s->n;
    real h = s->h, dor = (real)(dtd / s->prm.rho);
    for (int axis = 0; axis < 3; ++axis) {
        int d[3];
        face_dims(axis, n, d);
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < d[0]; ++i)
            for (int j = 0; j < d[1]; ++j)
                for (int k = 0; k < d[2]; ++k) {
                    long f = lidx(d[1], d[2], i, j, k);
                    int idx = axis == 0 ? i : (axis == 1 ? j : k);
                    if (idx == 0 || idx == n) {
                        s->u[axis][f] = R(0.0);
                        continue;
                    }
                    long tc = ((long)i * n + j) * n + k;
                    long tl = axis == 0 ? tc - (long)n * n : (axis == 1 ? tc - n : tc - 1);
                    real gp = s->p[tc] - s->p[tl];
                    s->u[axis][f] = s->u[axis][f] - dor * gp / h;
                }
    }
    for (int a = 0; a < 3; ++a) store_round_array(s->u[a], s->F[a], s->round_u);
}

ORC_API int orc_stage(orc_scene* s, int stage, double dt) {
    switch (stage) {
        case ORC_ST_CLEAR: st_clear(s); break;
        case ORC_ST_P2G: st_p2g(s); break;
        case ORC_ST_ADD_STATIC: st_add_static(s); break;
        case ORC_ST_BLEND: st_blend(s); break;
        case ORC_ST_FIXED_TW: apply_fixed(s, s->Tw); break;
        case ORC_ST_SOURCE: st_source(s, dt); break;
        case ORC_ST_ADVECT_T: st_advect_T(s, dt); break;
        case ORC_ST_CONDUCTIVITY: st_conductivity(s); break;
        case ORC_ST_DIFFUSE: st_diffuse(s, dt); break;
        case ORC_ST_PARTICLE_UPDATE: st_particle_update(s); break;
        case ORC_ST_SWAP_T: st_swap_T(s); break;
        case ORC_ST_FACE_FRACTION: st_face_fraction(s); break;
        case ORC_ST_ADVECT_FACES: st_advect_faces(s, dt); break;
        case ORC_ST_BUOYANCY: st_buoyancy(s, dt); break;
        case ORC_ST_RELAX: st_relax(s, dt); break;
        case ORC_ST_DIVERGENCE: st_divergence(s, dt); break;
        case ORC_ST_MEAN: st_mean(s); break;
        case ORC_ST_JACOBI: st_jacobi(s); break;
        case ORC_ST_PRESSURE_CG: st_pressure_cg(s); break;
        case ORC_ST_PROJECT: st_project(s, dt); break;
        default: return -1;
    }
    return 0;
}

/* solver.py:63-86: the whole substep without Newton's b.step */
ORC_API void orc_substep(orc_scene* s, double dt) {
    for (int st = 0; st < ORC_ST_COUNT; ++st) {
        if (st == ORC_ST_PARTICLE_UPDATE && s->P == 0) continue;
        if (st == ORC_ST_P2G && s->P == 0) continue;
        orc_stage(s, st, dt);
    }
    s->st.substeps += 1;
}

ORC_API void orc_set_storage_rounding(orc_scene* s, int round_u, int round_t, int bt_bf16) {
    s->round_u = round_u;
    s->round_t = round_t;
    s->bt_bf16 = bt_bf16;
}

ORC_API void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

ORC_API int orc_max_threads(void) { return omp_get_max_threads(); }

ORC_API void orc_run(orc_scene* s, double dt, int K) {
    for (int i = 0; i < K; ++i) orc_substep(s, dt);
}

/* grid.py:156-162 */
ORC_API void orc_cell_velocity(orc_scene* s, double* out) {
    int n = s->n;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                long t = ((long)i * n + j) * n + k;
                out[3 * t + 0] = R(0.5) * (s->u[0][lidx(n, n, i, j, k)] + s->u[0][lidx(n, n, i + 1, j, k)]);
                out[3 * t + 1] = R(0.5) * (s->u[1][lidx(n + 1, n, i, j, k)] + s->u[1][lidx(n + 1, n, i, j + 1, k)]);
                out[3 * t + 2] = R(0.5) * (s->u[2][lidx(n, n + 1, i, j, k)] + s->u[2][lidx(n, n + 1, i, j, k + 1)]);
            }
}

/* grid.py:257-263 */
ORC_API void orc_diagnostics(orc_scene* s, double* div_max, double* speed_max) {
    real* tmp = zalloc(s->N);
    divergence(s, R(1.0), tmp);
    double m = 0.0;
    for (long t = 0; t < s->N; ++t) {
        double a = fabs((double)tmp[t]);
        if (a > m) m = a;
    }
    *div_max = (double)(real)m * (double)s->prm.h;
    free(tmp);
    double sm = 0.0;
    for (int a = 0; a < 3; ++a)
        for (long t = 0; t < s->F[a]; ++t) {
            double v = fabs((double)s->u[a][t]);
            if (v > sm) sm = v;
        }
    *speed_max = sm;
}

/* ------------------------------------------------------------------------------------------
 * Unit entry points on explicit arrays (double in/out, REAL inside), for closed-form tests
 * ------------------------------------------------------------------------------------------ */

static real* to_real(const double* a, long n) {
    real* r = zalloc(n);
    for (long t = 0; t < n; ++t) r[t] = (real)a[t];
    return r;
}

static void from_real(const real* r, double* a, long n) {
    for (long t = 0; t < n; ++t) a[t] = (double)r[t];
}

ORC_API void orc_base_index(const double pos[3], const double off[3], double h, int32_t bi[3], double f[3]) {
    real p[3] = {(real)pos[0], (real)pos[1], (real)pos[2]};
    real o[3] = {(real)off[0], (real)off[1], (real)off[2]};
    int b[3];
    real fr[3];
    base_index(p, o, (real)h, b, fr);
    for (int c = 0; c < 3; ++c) {
        bi[c] = b[c];
        f[c] = fr[c];
    }
}

ORC_API double orc_sample(const double* a, int nx, int ny, int nz, const double pos[3], const double off[3],
                          double h) {
    real* ar = to_real(a, (long)nx * ny * nz);
    real p[3] = {(real)pos[0], (real)pos[1], (real)pos[2]};
    real o[3] = {(real)off[0], (real)off[1], (real)off[2]};
    double r = (double)sample(ar, p, o, (real)h, nx, ny, nz);
    free(ar);
    return r;
}

ORC_API void orc_scatter(double* a, int nx, int ny, int nz, const double pos[3], const double off[3], double h,
                         double val) {
    long N = (long)nx * ny * nz;
    real* ar = to_real(a, N);
    real p[3] = {(real)pos[0], (real)pos[1], (real)pos[2]};
    real o[3] = {(real)off[0], (real)off[1], (real)off[2]};
    scatter(ar, p, o, (real)h, nx, ny, nz, (real)val);
    from_real(ar, a, N);
    free(ar);
}

ORC_API void orc_backtrace(const double* u, const double* v, const double* w, int n, double h, double dt,
                           const double x[3], double out[3]) {
    long F = (long)(n + 1) * n * n;
    real *ur = to_real(u, F), *vr = to_real(v, F), *wr = to_real(w, F);
    real xr[3] = {(real)x[0], (real)x[1], (real)x[2]}, o[3];
    backtrace(ur, vr, wr, xr, (real)h, n, (real)dt, o);
    for (int c = 0; c < 3; ++c) out[c] = o[c];
    free(ur); free(vr); free(wr);
}

ORC_API void orc_laplacian_apply(int n, double h, const double* ax, const double* ay, const double* az,
                                 const double* x, const double* y, double* out, double alpha, double beta) {
    long N = (long)n * n * n, F = (long)(n + 1) * n * n;
    real* af[3] = {to_real(ax, F), to_real(ay, F), to_real(az, F)};
    real *xr = to_real(x, N), *yr = to_real(y, N), *o = zalloc(N);
    laplacian_apply(n, (real)h, af, xr, yr, o, (real)alpha, (real)beta);
    from_real(o, out, N);
    free(af[0]); free(af[1]); free(af[2]); free(xr); free(yr); free(o);
}

ORC_API void orc_jacobi_diag(int n, double h, const double* ax, const double* ay, const double* az, double* d) {
    long N = (long)n * n * n, F = (long)(n + 1) * n * n;
    real* af[3] = {to_real(ax, F), to_real(ay, F), to_real(az, F)};
    real* o = zalloc(N);
    jacobi_diag(n, (real)h, af, o);
    from_real(o, d, N);
    free(af[0]); free(af[1]); free(af[2]); free(o);
}

ORC_API void orc_heat_apply(int n, double h, double dt, const double* kap, const int32_t* fixed,
                            const double* tfix, const double* x, const double* y, double* out, double alpha,
                            double beta) {
    long N = (long)n * n * n;
    real *kr = to_real(kap, N), *tr = to_real(tfix, N), *xr = to_real(x, N), *yr = to_real(y, N), *o = zalloc(N);
    heat_apply(n, (real)h, (real)dt, kr, fixed, tr, xr, yr, o, (real)alpha, (real)beta);
    from_real(o, out, N);
    free(kr); free(tr); free(xr); free(yr); free(o);
}

ORC_API void orc_heat_rhs_diag(int n, double h, double dt, const double* T, const double* kap, const int32_t* fixed,
                               const double* tfix, double* rhs, double* diag) {
    long N = (long)n * n * n;
    real *Tr = to_real(T, N), *kr = to_real(kap, N), *tr = to_real(tfix, N), *r = zalloc(N), *d = zalloc(N);
    heat_rhs(n, (real)h, (real)dt, Tr, kr, fixed, tr, r);
    heat_diag(n, (real)h, (real)dt, kr, fixed, d);
    from_real(r, rhs, N);
    from_real(d, diag, N);
    free(Tr); free(kr); free(tr); free(r); free(d);
}

/* kind 0: pressure operator on (ax, ay, az); kind 1: heat operator on (kap, fixed, tfix, dt).
 * out_iters = {exact, reported, converged}; out_rr = {final ||r||^2, atol}. */
ORC_API void orc_pcg(int kind, int n, double h, double dt, const double* ax, const double* ay, const double* az,
                     const double* kap, const int32_t* fixed, const double* tfix, const double* diag,
                     const double* b, double* x, double tol, int maxiter, int check_every, int32_t out_iters[3],
                     double out_rr[2]) {
    long N = (long)n * n * n, F = (long)(n + 1) * n * n;
    real* af[3] = {NULL, NULL, NULL};
    real *kr = NULL, *tr = NULL;
    if (kind == 0) {
        af[0] = to_real(ax, F);
        af[1] = to_real(ay, F);
        af[2] = to_real(az, F);
    } else {
        kr = to_real(kap, N);
        tr = to_real(tfix, N);
    }
    real *dr = to_real(diag, N), *br = to_real(b, N), *xr = to_real(x, N);
    real *r = zalloc(N), *z = zalloc(N), *p = zalloc(N), *Ap = zalloc(N);
    opctx op = {kind, n, (real)h, (real)dt, af, kr, tr, fixed};
    cgres res;
    pcg(&op, dr, br, xr, N, tol, maxiter, check_every, r, z, p, Ap, &res);
    from_real(xr, x, N);
    out_iters[0] = res.exact;
    out_iters[1] = res.reported;
    out_iters[2] = res.converged;
    out_rr[0] = res.rr;
    out_rr[1] = res.atol;
    free(af[0]); free(af[1]); free(af[2]); free(kr); free(tr);
    free(dr); free(br); free(xr); free(r); free(z); free(p); free(Ap);
}

ORC_API int orc_warp_reported_count(int exact, int converged, int maxiter, int check_every) {
    return warp_reported_count(exact, converged, maxiter, check_every);
}

ORC_API int orc_real_bytes(void) { return (int)sizeof(real); }
