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
// Reductions use no atomics: every workgroup stores its partial sums with plain stores, and
// every workgroup of the next kernel sums all of them itself in one fixed order, so all of them
// hold the same bit-identical scalars (the kernel boundary makes the partials visible).  Pass A
// of iteration k+1 therefore also finishes iteration k: it forms r.r and r.z, tests convergence
// and derives beta; pass B forms p.Ap and alpha.  Scalars that live across kernels (r.z of the
// last two iterations, atol) are written by workgroup 0 into parity slots, never read and
// written by the same kernel.  The workgroup count is capped (kMaxPcgBlocks) so the redundant
// prologue sums stay small; each workgroup then covers several consecutive tiles.
// The search direction is double-buffered (pass A reads neighbours of the old direction while
// other workgroups write the new one).
//
// Operators:
//   pressure (grid.py:94-115): sum over in-range faces a_f (p_c - p_nb) / h^2, a_f recomputed
//                              from the solid fraction (grid.py:46-59)
//   heat (heat.py:55-105):     fixed cells identity; free cells x + dt/h^2 sum kappa_f terms
#include <hip/hip_ext.h>

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
    double* part_dir;      // [nb]       p.Ap partials of pass A
    double* part_upd;      // [3][nb]    r.r, r.z (and b.b from init) partials of pass B / init
    int cpt;               // tiles per workgroup
    int nb_dir;            // workgroups of the pass A launch (rows of part_dir)
    int nb_upd;            // workgroups of the init / pass B launches (rows of part_upd)
    // LDS-tiled pressure pass A: 16 x 16 (j, k) tiles, ichunk planes per workgroup (0: flat)
    int tiles_j, tiles_k, ichunk;
    int k;                 // completed iterations before this launch
    int remove_mean;       // pressure init: b -= mean when mean != 0 (grid.py:240-241)
    int report;            // dir: first launch of an enqueued batch, post the progress
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
// memory.  While the solve runs only the iteration count is posted (by the first pass of each
// batch the host enqueued), without waiting; at the end the results are stored, their
// completion waited for, and only then `done` — so the host never sees `done` before the
// results, and no system-scope release fence (a whole-L2 write-back) is used.
__device__ __forceinline__ void mirror_final(CgMirror* m, int it, float rr, double atol, int converged) {
    __hip_atomic_store(&m->rr, (double)rr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&m->atol, atol, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&m->converged, converged, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&m->it, it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    wait_stores();
    __hip_atomic_store(&m->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Whether the solve has stopped, read once per workgroup (thread 0, broadcast through LDS) so
// that every wave of the workgroup takes the same branch: workgroup 0 of a pass A may set the
// flag while other workgroups of the same launch start.
__device__ __forceinline__ bool solve_stopped(const CgCtl* c) {
    __shared__ int stopped;
    if (threadIdx.x == 0) stopped = c->done;
    __syncthreads();
    return stopped != 0;
}

// Sum NV rows of per-workgroup partials (row q at P + q*nb) in a fixed order; every thread of the
// BS-thread workgroup receives the totals.
template <int NV, int BS = kBlock>
__device__ __forceinline__ void sum_partials(const double* __restrict__ P, int nb, double (*lds)[BS / 64],
                                             double (&out)[NV]) {
    __shared__ double bcast[3];
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (int b = threadIdx.x; b < nb; b += BS)
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] += P[q * nb + b];
    block_sum<NV, BS>(acc, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) bcast[q] = acc[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = bcast[q];
}

// Store this workgroup's partial sums (thread 0 after the block sum).
template <int NV, int BS = kBlock>
__device__ __forceinline__ void store_partials(double (&v)[NV], double* __restrict__ P, int nb,
                                               double (*lds)[BS / 64]) {
    block_sum<NV, BS>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) P[q * nb + blockIdx.x] = v[q];
}

// Cells of logical workgroup L: cpt consecutive tiles of kBlock cells.
#define FOR_CELLS(nb)                                                        \
    const int L_ = xcd_tile(blockIdx.x, nb);                                 \
    for (int u_ = 0; u_ < a.cpt; ++u_)                                       \
        if (long t = ((long)L_ * a.cpt + u_) * kBlock + threadIdx.x; t < a.N)

// Prologue of every pass A: finish iteration k from pass B's partial sums (r.r, r.z; b.b too
// at k = 0, where atol is formed), test convergence and the iteration cap, and derive beta.
// Workgroup 0 records the state.  Returns true when the solve stops (in every thread).
template <int BS = kBlock>
__device__ bool finish_iteration(const PcgArgs& a, double (*lds)[BS / 64], float& beta) {
    CgCtl* c = a.ctl;
    const int k = a.k;
    float rr, rz, atol_sq;
    double atol;
    if (k == 0) {
        double tot[3];
        sum_partials<3, BS>(a.part_upd, a.nb_upd, lds, tot);
        rr = (float)tot[0];
        rz = (float)tot[1];
        float bb = (float)tot[2];
        // warp: atol = max(tol * ||b||, tol)
        atol = c->tol * sqrt((double)bb);
        if (atol < c->tol) atol = c->tol;
        atol_sq = (float)(atol * atol);
    } else {
        double tot[2];
        sum_partials<2, BS>(a.part_upd, a.nb_upd, lds, tot);
        rr = (float)tot[0];
        rz = (float)tot[1];
        atol = c->atol;
        atol_sq = c->atol_sq;
    }
    const bool conv = rr <= atol_sq;
    const bool stop = conv || k >= c->maxeff;
    // beta of iteration k (warp: rz_new / rz_old while rr > atol^2)
    beta = (k == 0) ? 0.0f : ((rr > atol_sq) ? rz / c->slot_rz[(k + 1) & 1] : 0.0f);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->slot_rr[k & 1] = rr;
        c->slot_rz[k & 1] = rz;
        if (k == 0) {
            c->atol = atol;
            c->atol_sq = atol_sq;
        }
        if (stop) {
            c->done = 1;
            c->converged = conv;
            c->it = k;
            mirror_final(a.mirror, k, rr, atol, conv);
        } else if (a.report) {
            __hip_atomic_store(&a.mirror->it, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    return stop;
}

// r = b - A x ; z = D r ; partial sums r.r, r.z, b.b
template <int OP>
__global__ __launch_bounds__(kBlock) void k_pcg_init(PcgArgs a) {
    __shared__ double lds[3][kWaves];
    const int nb = gridDim.x;
    double v[3] = {0.0, 0.0, 0.0};
    FOR_CELLS(nb) {
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
        v[0] += (double)(rc * rc);
        v[1] += (double)(rc * zc);
        v[2] += (double)(bc * bc);
    }
    store_partials<3>(v, a.part_upd, nb, lds);
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // a new solve: clear the previous one's stop
        a.ctl->done = 0;
        a.ctl->converged = 0;
        a.ctl->it = 0;
    }
}

// pass A of iteration k+1: finish iteration k (r.r, r.z, convergence, beta), then the direction
// update at the cell and its neighbours, SpMV and partial p.Ap
template <int OP>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(PcgArgs a) {
    __shared__ double lds[3][kWaves];
    CgCtl* c = a.ctl;
    if (solve_stopped(c)) return;
    const int nb = gridDim.x;
    const int k = a.k;
    float beta;
    if (finish_iteration(a, lds, beta)) return;
    double v[1] = {0.0};
    const bool first = k == 0;
    FOR_CELLS(nb) {
        int i, j, kk;
        cell_ijk(t, a.n, i, j, kk);
        const float* zs = a.z;
        const float* ds = a.d_old;
        auto dir = [zs, ds, beta, first](long q) { return first ? zs[q] : zs[q] + beta * ds[q]; };
        float pc = dir(t);
        float acc = Operator<OP>::apply(a, t, i, j, kk, dir);
        bool fixed = OP == 0 && (a.flags[t] & FLAG_FIXED);
        float apc = Operator<OP>::combine(a, acc, pc, fixed, 1.0f);
        a.d_new[t] = pc;
        a.Ap[t] = apc;
        v[0] += (double)(pc * apc);
    }
    store_partials<1>(v, a.part_dir, nb, lds);
}

// Pass A for the pressure operator, LDS-tiled.  A workgroup of 1024 threads owns a 16 x 16 x D
// block of cells.  It first forms the new direction p = z + beta p_old and loads the solid
// fraction on the block framed by a one-cell halo ((D+2) x 18 x 18 cells, all loads in flight at
// once) into LDS, then every thread applies the stencil to D/4 cells from LDS.  Per cell the
// arithmetic is exactly that of k_pcg_dir<1>.
constexpr int kT3 = 16, kF3 = kT3 + 2, kBlock3 = 1024;

template <int D>
__global__ __launch_bounds__(kBlock3) void k_pcg_dir_tile3(PcgArgs a) {
    __shared__ float P[D + 2][kF3][kF3];
    __shared__ float S[D + 2][kF3][kF3];
    __shared__ double lds[3][kBlock3 / 64];
    {
        __shared__ int stopped;
        if (threadIdx.x == 0) stopped = a.ctl->done;
        __syncthreads();
        if (stopped) return;
    }
    const int nb = gridDim.x;
    const int n = a.n;
    const int L = xcd_tile(blockIdx.x, nb);
    const int tjk = a.tiles_j * a.tiles_k;
    const int ti = L / tjk, rem = L % tjk;
    const int i0 = ti * D, j0 = (rem / a.tiles_k) * kT3, k0 = (rem % a.tiles_k) * kT3;
    const bool first = a.k == 0;
    const float* __restrict__ zs = a.z;
    const float* __restrict__ ds = a.d_old;
    const float* __restrict__ as = a.alpha_s;
    // frame operands are loaded before the prologue (they do not depend on beta), so the
    // latency of the partial-sum prologue overlaps theirs
    constexpr int kFrame = (D + 2) * kF3 * kF3;
    constexpr int kPer = (kFrame + kBlock3 - 1) / kBlock3;
    float fz[kPer], fd[kPer], fa[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int f = threadIdx.x + u * kBlock3;
        fz[u] = 0.0f;
        fd[u] = 0.0f;
        fa[u] = 0.0f;
        if (f < kFrame) {
            const int fi = f / (kF3 * kF3), fr = f % (kF3 * kF3);
            const int fj = fr / kF3, fk = fr % kF3;
            const int gi = i0 + fi - 1, gj = j0 + fj - 1, gk = k0 + fk - 1;
            if (gi >= 0 && gi < n && gj >= 0 && gj < n && gk >= 0 && gk < n) {
                const long t = ((long)gi * n + gj) * n + gk;
                fz[u] = zs[t];
                if (!first) fd[u] = ds[t];
                fa[u] = as[t];
            }
        }
    }
    float beta;
    if (finish_iteration<kBlock3>(a, lds, beta)) return;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int f = threadIdx.x + u * kBlock3;
        if (f < kFrame) {
            const int fi = f / (kF3 * kF3), fr = f % (kF3 * kF3);
            const int fj = fr / kF3, fk = fr % kF3;
            P[fi][fj][fk] = first ? fz[u] : fz[u] + beta * fd[u];
            S[fi][fj][fk] = fa[u];
        }
    }
    __syncthreads();
    const float amin = a.amin;
    const int lk = threadIdx.x % kT3, lj = (threadIdx.x / kT3) % kT3, lq = threadIdx.x / (kT3 * kT3);
    const int j = j0 + lj, k = k0 + lk;
    double v[1] = {0.0};
    if (j < n && k < n) {
        const int fj = lj + 1, fk = lk + 1;
#pragma unroll
        for (int m = 0; m < D / 4; ++m) {
            const int li = lq + 4 * m, i = i0 + li;
            if (i >= n) break;
            const int fi = li + 1;
            const float pc = P[fi][fj][fk], ac = S[fi][fj][fk];
            float acc = 0.0f;
            if (i > 0) acc += face_fraction(S[fi - 1][fj][fk], ac, amin) * (pc - P[fi - 1][fj][fk]);
            if (i < n - 1) acc += face_fraction(ac, S[fi + 1][fj][fk], amin) * (pc - P[fi + 1][fj][fk]);
            if (j > 0) acc += face_fraction(S[fi][fj - 1][fk], ac, amin) * (pc - P[fi][fj - 1][fk]);
            if (j < n - 1) acc += face_fraction(ac, S[fi][fj + 1][fk], amin) * (pc - P[fi][fj + 1][fk]);
            if (k > 0) acc += face_fraction(S[fi][fj][fk - 1], ac, amin) * (pc - P[fi][fj][fk - 1]);
            if (k < n - 1) acc += face_fraction(ac, S[fi][fj][fk + 1], amin) * (pc - P[fi][fj][fk + 1]);
            const float apc = 1.0f * acc / a.hh;
            const long t = ((long)i * n + j) * n + k;
            a.d_new[t] = pc;
            a.Ap[t] = apc;
            v[0] += (double)(pc * apc);
        }
    }
    store_partials<1, kBlock3>(v, a.part_dir, nb, lds);
}

// pass B of iteration k+1: alpha from p.Ap, x and r update, z = D r, partial r.r and r.z
template <int OP>
__global__ __launch_bounds__(kBlock) void k_pcg_upd(PcgArgs a) {
    __shared__ double lds[2][kWaves];
    CgCtl* c = a.ctl;
    if (solve_stopped(c)) return;
    const int nb = gridDim.x;
    const int L = xcd_tile(blockIdx.x, nb);
    // the operands of the first kPre tiles are loaded before the p.Ap prologue, so its latency
    // overlaps theirs
    constexpr int kPre = 4;
    float px[kPre], pd[kPre], pr[kPre], pa[kPre], pD[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        long t = ((long)L * a.cpt + u) * kBlock + threadIdx.x;
        if (u < a.cpt && t < a.N) {
            px[u] = a.x[t];
            pd[u] = a.d_new[t];
            pr[u] = a.r[t];
            pa[u] = a.Ap[t];
            pD[u] = a.D[t];
        }
    }
    double tot[1];
    sum_partials<1>(a.part_dir, a.nb_dir, lds, tot);
    const float pAp = (float)tot[0];
    const float rr = c->slot_rr[a.k & 1], rz = c->slot_rz[a.k & 1];
    const float alpha = (rr > c->atol_sq) ? rz / pAp : 0.0f;
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        long t = ((long)L * a.cpt + u) * kBlock + threadIdx.x;
        if (u < a.cpt && t < a.N) {
            float xc = px[u] + alpha * pd[u];
            float rc = pr[u] - alpha * pa[u];
            float zc = pD[u] * rc;
            a.x[t] = xc;
            a.r[t] = rc;
            a.z[t] = zc;
            v[0] += (double)(rc * rc);
            v[1] += (double)(rc * zc);
        }
    }
    for (int u = kPre; u < a.cpt; ++u) {
        long t = ((long)L * a.cpt + u) * kBlock + threadIdx.x;
        if (t >= a.N) break;
        float xc = a.x[t] + alpha * a.d_new[t];
        float rc = a.r[t] - alpha * a.Ap[t];
        float zc = a.D[t] * rc;
        a.x[t] = xc;
        a.r[t] = rc;
        a.z[t] = zc;
        v[0] += (double)(rc * rc);
        v[1] += (double)(rc * zc);
    }
    store_partials<2>(v, a.part_upd, nb, lds);
}

}  // namespace
int pcg_cpt(const Scene* s);
int pcg_blocks(const Scene* s);
void pcg_dir_geometry(const Scene* s, int op, int* nb, int* tiles_j, int* tiles_k, int* ichunk);
namespace {

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
    a.part_dir = s->partials;
    a.part_upd = s->partials + kMaxDirBlocks;
    a.cpt = pcg_cpt(s);
    a.nb_upd = pcg_blocks(s);
    pcg_dir_geometry(s, op, &a.nb_dir, &a.tiles_j, &a.tiles_k, &a.ichunk);
    a.k = 0;
    a.remove_mean = op == 1;
    a.report = 0;
    return a;
}

}  // namespace


// tiles of kBlock cells per PCG workgroup: enough that at most kMaxPcgBlocks workgroups run
// (the scene's pcg_cpt raises it further)
int pcg_cpt(const Scene* s) {
    long tiles = (s->g.N + kBlock - 1) / kBlock;
    long need = (tiles + kMaxPcgBlocks - 1) / kMaxPcgBlocks;
    int cpt = s->pcg_cpt > 0 ? s->pcg_cpt : 1;
    return (int)(need > cpt ? need : cpt);
}

int pcg_blocks(const Scene* s) {
    long tiles = (s->g.N + kBlock - 1) / kBlock;
    int cpt = pcg_cpt(s);
    return (int)((tiles + cpt - 1) / cpt);
}

#define PCG_LAUNCH(KERNEL, s, op, a)                                                        \
    do {                                                                                   \
        dim3 grid_(pcg_blocks(s)), block_(kBlock);                                         \
        if (op == 1)                                                                       \
            hipLaunchKernelGGL((KERNEL<1>), grid_, block_, 0, (s)->stream, a);             \
        else                                                                               \
            hipLaunchKernelGGL((KERNEL<0>), grid_, block_, 0, (s)->stream, a);             \
    } while (0)

void launch_pcg_init(Scene* s, int op, float dt) {
    PcgArgs a = make_args(s, op, dt);
    PCG_LAUNCH(k_pcg_init, s, op, a);
}

// Pass A geometry: the LDS-tiled kernel for the pressure operator (unless HFM_PCG_FLAT_DIR) with
// planes D = 8, else 16, else 4, the first that keeps at most kMaxDirBlocks workgroups
// (HFM_DIR_ICHUNK forces D), else the flat kernel with the pass B geometry.
void pcg_dir_geometry(const Scene* s, int op, int* nb, int* tiles_j, int* tiles_k, int* ichunk) {
    const int n = s->g.n;
    const int tj = (n + kT3 - 1) / kT3, tk = tj;
    int D = 0;
    if (op == 1 && !s->flat_dir) {
        const int cand[3] = {8, 16, 4};
        for (int c = 0; c < 3 && !D; ++c) {
            int d = cand[c];
            if (s->dir_ichunk > 0 && d != s->dir_ichunk) continue;
            long blocks = (long)tj * tk * ((n + d - 1) / d);
            if (blocks <= kMaxDirBlocks) D = d;
        }
    }
    if (!D) {
        *nb = pcg_blocks(s);
        *tiles_j = *tiles_k = *ichunk = 0;
        return;
    }
    *nb = tj * tk * ((n + D - 1) / D);
    *tiles_j = tj;
    *tiles_k = tk;
    *ichunk = D;
}

// With ev0/ev1 the launch carries its own start/stop events (hipExtLaunchKernelGGL): the events
// time the dispatch itself, with no separate event packets around it.
template <class K>
static void launch_timed(K kernel, dim3 grid, dim3 block, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1,
                         const PcgArgs& a) {
    if (ev0)
        hipExtLaunchKernelGGL(kernel, grid, block, 0, st, ev0, ev1, 0, a);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, st, a);
}

void launch_pcg_dir(Scene* s, int op, float dt, int k, int report, hipEvent_t ev0, hipEvent_t ev1) {
    PcgArgs a = make_args(s, op, dt);
    a.k = k;
    a.report = report;
    a.d_old = s->cg_d + (long)((k & 1) ^ 1) * s->g.N;
    a.d_new = s->cg_d + (long)(k & 1) * s->g.N;
    const dim3 blk3(kBlock3), blk(kBlock);
    if (a.ichunk == 4)
        launch_timed(k_pcg_dir_tile3<4>, dim3(a.nb_dir), blk3, s->stream, ev0, ev1, a);
    else if (a.ichunk == 8)
        launch_timed(k_pcg_dir_tile3<8>, dim3(a.nb_dir), blk3, s->stream, ev0, ev1, a);
    else if (a.ichunk == 16)
        launch_timed(k_pcg_dir_tile3<16>, dim3(a.nb_dir), blk3, s->stream, ev0, ev1, a);
    else if (op == 1)
        launch_timed(k_pcg_dir<1>, dim3(a.nb_dir), blk, s->stream, ev0, ev1, a);
    else
        launch_timed(k_pcg_dir<0>, dim3(a.nb_dir), blk, s->stream, ev0, ev1, a);
}

void launch_pcg_upd(Scene* s, int op, int k, hipEvent_t ev0, hipEvent_t ev1) {
    PcgArgs a = make_args(s, op, 0.0f);
    a.k = k;
    a.d_new = s->cg_d + (long)(k & 1) * s->g.N;
    if (op == 1)
        launch_timed(k_pcg_upd<1>, dim3(a.nb_upd), dim3(kBlock), s->stream, ev0, ev1, a);
    else
        launch_timed(k_pcg_upd<0>, dim3(a.nb_upd), dim3(kBlock), s->stream, ev0, ev1, a);
}

}  // namespace hfm
