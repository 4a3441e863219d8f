// hfm_pcg.hip — Jacobi-preconditioned CG for the pressure projection and the backward-Euler
// heat solve (gfx950).
//
// Semantics are those of warp.optim.linear.cg as the reference calls it (grid.py:247-249,
// heat.py:206-210): atol = max(tol * ||b||, tol); r = b - A x; z = D r; p = z; then per
// iteration Ap = A p, alpha = (rr > atol^2) ? rz/pAp : 0, x += alpha p, r -= alpha Ap,
// z = D r, beta = (rr > atol^2) ? rz'/rz : 0, p = z + beta p.  Once ||r||^2 <= atol^2 alpha and
// beta are zero, so x and r stop changing: the kernels here stop at that iteration (ctl.done)
// and every later launch returns at once.
//
// One iteration is two fused passes over the cells:
//   pass A (dir): p = z + beta p_old at the cell and its six neighbours, Ap = A p, sum p.Ap
//   pass B (upd): x += alpha p, r -= alpha Ap, z = D r, sums r.r and r.z
// The scalars are formed by the workgroup that finishes last (fixed summation tree), which
// also publishes the progress to pinned host memory so the host can stop enqueueing.
// The search direction is double-buffered (pass A reads neighbours of the old direction while
// other workgroups write the new one).
//
// Operators:
//   pressure (grid.py:94-115): sum over in-range faces a_f (p_c - p_nb) / h^2, a_f recomputed
//                              from the solid fraction (grid.py:46-59)
//   heat (heat.py:55-105):     fixed cells identity; free cells x + dt/h^2 sum kappa_f terms
#include "hfm_device.h"
#include "hfm_scene.h"

namespace hfm {

namespace {

struct PcgArgs {
    int n;
    long N;
    float h, hh, dt_hh, amin;
    const float* alpha_s;  // pressure
    const float* kap;      // heat
    const uint8_t* flags;  // heat
    float* x;
    float* b;
    float* r;
    float* z;
    const float* d_old;
    float* d_new;
    float* Ap;
    const float* D;
    CgCtl* ctl;
    CgMirror* mirror;
    double* partials;
    int remove_mean;       // pressure init: b -= mean when mean != 0 (grid.py:240-241)
    int report;            // upd: last iteration of an enqueued batch, post the progress
};

template <int OP>
struct Operator;

// grid.py:94-115 with alpha/beta applied by the caller
template <>
struct Operator<1> {
    template <class Dir>
    __device__ static float apply(const PcgArgs& a, long t, int i, int j, int k, const Dir& dir) {
        const int n = a.n;
        const long nn = (long)n * n;
        const float* as = a.alpha_s;
        float ac = as[t];
        float pc = dir(t);
        float acc = 0.0f;
        if (i > 0) acc += face_fraction(as[t - nn], ac, a.amin) * (pc - dir(t - nn));
        if (i < n - 1) acc += face_fraction(ac, as[t + nn], a.amin) * (pc - dir(t + nn));
        if (j > 0) acc += face_fraction(as[t - n], ac, a.amin) * (pc - dir(t - n));
        if (j < n - 1) acc += face_fraction(ac, as[t + n], a.amin) * (pc - dir(t + n));
        if (k > 0) acc += face_fraction(as[t - 1], ac, a.amin) * (pc - dir(t - 1));
        if (k < n - 1) acc += face_fraction(ac, as[t + 1], a.amin) * (pc - dir(t + 1));
        return acc;
    }
    // out = alpha * acc / h^2 + beta * y
    __device__ static float combine(const PcgArgs& a, float acc, float xc, bool fixed, float alpha) {
        (void)xc;
        (void)fixed;
        return alpha * acc / a.hh;
    }
};

// heat.py:55-105
template <>
struct Operator<0> {
    template <class Dir>
    __device__ static float apply(const PcgArgs& a, long t, int i, int j, int k, const Dir& dir) {
        const int n = a.n;
        const long nn = (long)n * n;
        float kc = a.kap[t];
        float xc = dir(t);
        float acc = 0.0f;
        long nb[6] = {t - nn, t + nn, t - n, t + n, t - 1, t + 1};
        bool in[6] = {i > 0, i < n - 1, j > 0, j < n - 1, k > 0, k < n - 1};
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            if (!in[s]) continue;
            float kf = face_kappa(kc, a.kap[nb[s]]);
            if (a.flags[nb[s]] & FLAG_FIXED)
                acc += kf * xc;
            else
                acc += kf * (xc - dir(nb[s]));
        }
        return acc;
    }
    // fixed: alpha * x ; free: alpha * (x + dt/h^2 * acc)
    __device__ static float combine(const PcgArgs& a, float acc, float xc, bool fixed, float alpha) {
        if (fixed) return alpha * xc;
        return alpha * (xc + a.dt_hh * acc);
    }
};

__device__ __forceinline__ void cell_ijk(long t, int n, int& i, int& j, int& k) {
    const long nn = (long)n * n;
    i = (int)(t / nn);
    j = (int)((t / n) % n);
    k = (int)(t % n);
}

// Publish progress to pinned host memory.  Relaxed system-scope stores go straight to host
// memory.  While the solve runs only the iteration count is posted, by the last iteration of
// each batch the host enqueued (`report`), without waiting; at the end the results are stored,
// their completion waited for, and only then `done` — so the host never sees `done` before the
// results, and no system-scope release fence (a whole-L2 write-back) is used.
__device__ __forceinline__ void mirror_progress(CgMirror* m, const CgCtl* c, int report) {
    if (!c->done) {
        if (report) __hip_atomic_store(&m->it, c->it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return;
    }
    __hip_atomic_store(&m->rr, (double)c->rr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&m->atol, c->atol, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&m->converged, c->converged, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&m->it, c->it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    wait_stores();
    __hip_atomic_store(&m->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Cells of logical workgroup L: CPT consecutive tiles of kBlock cells.
#define FOR_CELLS(CPT, nb)                                                   \
    const int L_ = xcd_tile(blockIdx.x, nb);                                 \
    _Pragma("unroll") for (int u_ = 0; u_ < CPT; ++u_)                       \
        if (long t = ((long)L_ * CPT + u_) * kBlock + threadIdx.x; t < a.N)

// r = b - A x ; z = D r ; sums b.b, r.r, r.z ; then atol and the convergence test
template <int OP, int CPT>
__global__ __launch_bounds__(kBlock) void k_pcg_init(PcgArgs a) {
    __shared__ double lds[3][kWaves];
    const int nb = gridDim.x;
    double v[3] = {0.0, 0.0, 0.0};
    FOR_CELLS(CPT, nb) {
        int i, j, k;
        cell_ijk(t, a.n, i, j, k);
        float bc = a.b[t];
        if (OP == 1 && a.remove_mean) {
            float mean = a.ctl->mean;
            if (mean != 0.0f) {
                bc = bc - mean;
                a.b[t] = bc;
            }
        }
        const float* xs = a.x;
        auto dir = [xs](long q) { return xs[q]; };
        float acc = Operator<OP>::apply(a, t, i, j, k, dir);
        bool fixed = OP == 0 && (a.flags[t] & FLAG_FIXED);
        float rc = Operator<OP>::combine(a, acc, xs[t], fixed, -1.0f) + 1.0f * bc;
        float zc = a.D[t] * rc;
        a.r[t] = rc;
        a.z[t] = zc;
        v[0] += (double)(bc * bc);
        v[1] += (double)(rc * rc);
        v[2] += (double)(rc * zc);
    }
    double tot[3];
    if (!grid_sum_last<3>(v, a.partials, a.ctl->tickets, nb, lds, tot)) return;
    if (threadIdx.x != 0) return;
    CgCtl* c = a.ctl;
    float bb = (float)tot[0], rr = (float)tot[1], rz = (float)tot[2];
    double atol = c->tol * sqrt((double)bb);
    if (atol < c->tol) atol = c->tol;
    c->atol = atol;
    c->atol_sq = (float)(atol * atol);
    c->rr = rr;
    c->rz = rz;
    c->alpha = 0.0f;
    c->beta = 0.0f;
    c->it = 0;
    c->first = 1;
    c->converged = rr <= c->atol_sq;
    c->done = c->converged;
    mirror_progress(a.mirror, c, 1);
}

// pass A: direction update at the cell and its neighbours, SpMV, p.Ap ; then alpha
template <int OP, int CPT>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(PcgArgs a) {
    __shared__ double lds[1][kWaves];
    if (a.ctl->done) return;
    const int nb = gridDim.x;
    const float beta = a.ctl->beta;
    const int first = a.ctl->first;
    double v[1] = {0.0};
    FOR_CELLS(CPT, nb) {
        int i, j, k;
        cell_ijk(t, a.n, i, j, k);
        const float* zs = a.z;
        const float* ds = a.d_old;
        auto dir = [zs, ds, beta, first](long q) { return first ? zs[q] : zs[q] + beta * ds[q]; };
        float pc = dir(t);
        float acc = Operator<OP>::apply(a, t, i, j, k, dir);
        bool fixed = OP == 0 && (a.flags[t] & FLAG_FIXED);
        float apc = Operator<OP>::combine(a, acc, pc, fixed, 1.0f);
        a.d_new[t] = pc;
        a.Ap[t] = apc;
        v[0] += (double)(pc * apc);
    }
    double tot[1];
    if (!grid_sum_last<1>(v, a.partials, a.ctl->tickets, nb, lds, tot)) return;
    if (threadIdx.x != 0) return;
    CgCtl* c = a.ctl;
    float pAp = (float)tot[0];
    c->alpha = (c->rr > c->atol_sq) ? c->rz / pAp : 0.0f;
}

// pass B: x, r update, z = D r, sums r.r and r.z ; then beta and the convergence test
template <int OP, int CPT>
__global__ __launch_bounds__(kBlock) void k_pcg_upd(PcgArgs a) {
    __shared__ double lds[2][kWaves];
    if (a.ctl->done) return;
    const int nb = gridDim.x;
    const float alpha = a.ctl->alpha;
    double v[2] = {0.0, 0.0};
    FOR_CELLS(CPT, nb) {
        float xc = a.x[t] + alpha * a.d_new[t];
        float rc = a.r[t] - alpha * a.Ap[t];
        float zc = a.D[t] * rc;
        a.x[t] = xc;
        a.r[t] = rc;
        a.z[t] = zc;
        v[0] += (double)(rc * rc);
        v[1] += (double)(rc * zc);
    }
    double tot[2];
    if (!grid_sum_last<2>(v, a.partials, a.ctl->tickets, nb, lds, tot)) return;
    if (threadIdx.x != 0) return;
    CgCtl* c = a.ctl;
    float rr = (float)tot[0], rz = (float)tot[1];
    c->beta = (rr > c->atol_sq) ? rz / c->rz : 0.0f;
    c->rz = rz;
    c->rr = rr;
    c->first = 0;
    c->it = c->it + 1;
    if (rr <= c->atol_sq) {
        c->converged = 1;
        c->done = 1;
    } else if (c->it >= c->maxeff) {
        c->done = 1;
    }
    mirror_progress(a.mirror, c, a.report);
}

PcgArgs make_args(Scene* s, int op, float dt) {
    PcgArgs a;
    a.n = s->g.n;
    a.N = s->g.N;
    a.h = s->g.h;
    a.hh = a.h * a.h;
    a.dt_hh = dt / (a.h * a.h);
    a.amin = (float)s->prm.alpha_f_min;
    a.alpha_s = s->alpha_s;
    a.kap = s->kap;
    a.flags = s->flags;
    if (op == 1) {
        a.x = s->p;
        a.b = s->rhs;
        a.D = s->pdiag;
    } else {
        a.x = s->T2;
        a.b = s->hrhs;
        a.D = s->hdiag;
    }
    a.r = s->cg_r;
    a.z = s->cg_z;
    a.d_old = nullptr;
    a.d_new = nullptr;
    a.Ap = s->cg_Ap;
    a.ctl = &s->ctl[op];
    a.mirror = &s->mirror_dev[op];
    a.partials = s->partials;
    a.remove_mean = op == 1;
    a.report = 0;
    return a;
}

}  // namespace

int pcg_blocks(const Scene* s) {
    long tiles = (s->g.N + kBlock - 1) / kBlock;
    int cpt = s->pcg_cpt > 0 ? s->pcg_cpt : 1;
    return (int)((tiles + cpt - 1) / cpt);
}

// dispatch over (operator, cells per thread)
#define PCG_LAUNCH(KERNEL, s, op, a)                                                                         \
    do {                                                                                                    \
        dim3 grid_(pcg_blocks(s)), block_(kBlock);                                                          \
        int cpt_ = (s)->pcg_cpt;                                                                            \
        if (op == 1) {                                                                                      \
            if (cpt_ == 4) hipLaunchKernelGGL((KERNEL<1, 4>), grid_, block_, 0, (s)->stream, a);            \
            else if (cpt_ == 2) hipLaunchKernelGGL((KERNEL<1, 2>), grid_, block_, 0, (s)->stream, a);       \
            else hipLaunchKernelGGL((KERNEL<1, 1>), grid_, block_, 0, (s)->stream, a);                      \
        } else {                                                                                            \
            if (cpt_ == 4) hipLaunchKernelGGL((KERNEL<0, 4>), grid_, block_, 0, (s)->stream, a);            \
            else if (cpt_ == 2) hipLaunchKernelGGL((KERNEL<0, 2>), grid_, block_, 0, (s)->stream, a);       \
            else hipLaunchKernelGGL((KERNEL<0, 1>), grid_, block_, 0, (s)->stream, a);                      \
        }                                                                                                   \
    } while (0)

void launch_pcg_init(Scene* s, int op, float dt) {
    PcgArgs a = make_args(s, op, dt);
    PCG_LAUNCH(k_pcg_init, s, op, a);
}

void launch_pcg_dir(Scene* s, int op, float dt, int parity) {
    PcgArgs a = make_args(s, op, dt);
    a.d_old = s->cg_d + (long)(parity ^ 1) * s->g.N;
    a.d_new = s->cg_d + (long)parity * s->g.N;
    PCG_LAUNCH(k_pcg_dir, s, op, a);
}

void launch_pcg_upd(Scene* s, int op, int parity, int report) {
    PcgArgs a = make_args(s, op, 0.0f);
    a.report = report;
    a.d_new = s->cg_d + (long)parity * s->g.N;
    PCG_LAUNCH(k_pcg_upd, s, op, a);
}

}  // namespace hfm
