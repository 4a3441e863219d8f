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
// Launches are batched over scenes (a block map names each workgroup's scene and its index
// inside the scene): every scene runs its own solve with its own scalars, and a scene whose
// solve has stopped skips the rest of the launches.  The host learns the progress of the batch
// from workgroup 0 of the first pass A of each group of iterations it enqueued.
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

// One scene's view of one solve (op 0 heat, 1 pressure) at iteration k.
struct PcgView {
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
    double* part_dir;      // [nb_dir]     p.Ap partials of pass A
    double* part_upd;      // [3][nb_upd]  r.r, r.z (and b.b from init) partials of pass B / init
    int cpt;               // tiles per workgroup (init / pass B / flat pass A)
    int nb_dir, nb_upd;
    int tiles_j, tiles_k;
    int k;                 // completed iterations before this launch
    int lb;                // workgroup index inside the scene
};

template <int OP>
__device__ __forceinline__ PcgView make_view(const SceneDev& a, int lb, int k, float dt) {
    PcgView v;
    v.n = a.n;
    v.N = a.N;
    v.h = a.h;
    v.hh = a.hh;
    v.dt_hh = dt / (a.h * a.h);
    v.amin = a.amin;
    v.alpha_s = a.alpha_s;
    v.kap = a.kap;
    v.flags = a.flags;
    if (OP == 1) {
        v.x = a.p;
        v.b = a.rhs;
        v.D = a.pdiag;
    } else {
        v.x = a.T;
        v.b = a.hrhs;
        v.D = a.hdiag;
    }
    v.r = a.cg_r;
    v.z = a.cg_z;
    v.Ap = a.cg_Ap;
    v.d_new = a.cg_d + (long)(k & 1) * a.N;
    v.d_old = a.cg_d + (long)((k & 1) ^ 1) * a.N;
    v.ctl = a.ctl + OP;
    v.mirror = a.mirror + OP;
    v.part_dir = a.partials;
    v.part_upd = a.partials + kMaxDirBlocks;
    v.cpt = a.cpt;
    v.nb_upd = a.nb_upd;
    v.nb_dir = OP == 1 ? a.nb_dir : a.nb_upd;
    v.tiles_j = a.tiles_j;
    v.tiles_k = a.tiles_k;
    v.k = k;
    v.lb = lb;
    return v;
}

template <int OP>
struct Operator;

// grid.py:94-115 with alpha/beta applied by the caller
template <>
struct Operator<1> {
    template <class Dir>
    __device__ static float apply(const PcgView& a, long t, int i, int j, int k, const Dir& dir) {
        const int n = a.n;
        const long nn = (long)n * n;
        const float* as = a.alpha_s;
        const float ac = as[t];
        const float pc = dir(t);
        float acc = 0.0f;
        if (i > 0) acc += face_fraction(as[t - nn], ac, a.amin) * (pc - dir(t - nn));
        if (i < n - 1) acc += face_fraction(ac, as[t + nn], a.amin) * (pc - dir(t + nn));
        if (j > 0) acc += face_fraction(as[t - n], ac, a.amin) * (pc - dir(t - n));
        if (j < n - 1) acc += face_fraction(ac, as[t + n], a.amin) * (pc - dir(t + n));
        if (k > 0) acc += face_fraction(as[t - 1], ac, a.amin) * (pc - dir(t - 1));
        if (k < n - 1) acc += face_fraction(ac, as[t + 1], a.amin) * (pc - dir(t + 1));
        return acc;
    }
    // alpha * acc / h^2
    __device__ static float combine(const PcgView& a, float acc, float xc, bool fixed, float alpha) {
        (void)xc;
        (void)fixed;
        return alpha * acc / a.hh;
    }
};

// heat.py:55-105
template <>
struct Operator<0> {
    template <class Dir>
    __device__ static float apply(const PcgView& a, long t, int i, int j, int k, const Dir& dir) {
        const int n = a.n;
        const long nn = (long)n * n;
        const float kc = a.kap[t];
        const float xc = dir(t);
        float acc = 0.0f;
        const long nb[6] = {t - nn, t + nn, t - n, t + n, t - 1, t + 1};
        const bool in[6] = {i > 0, i < n - 1, j > 0, j < n - 1, k > 0, k < n - 1};
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            if (!in[s]) continue;
            const float kf = face_kappa(kc, a.kap[nb[s]]);
            if (a.flags[nb[s]] & FLAG_FIXED)
                acc += kf * xc;
            else
                acc += kf * (xc - dir(nb[s]));
        }
        return acc;
    }
    // fixed: alpha * x ; free: alpha * (x + dt/h^2 * acc)
    __device__ static float combine(const PcgView& a, float acc, float xc, bool fixed, float alpha) {
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

// Publish the result of a solve to pinned host memory.  Relaxed system-scope stores go straight
// to host memory; the results are stored, their completion waited for, and only then `done` —
// so the host never sees `done` before the results, and no system-scope release fence (a
// whole-L2 write-back) is used.
__device__ __forceinline__ void mirror_final(CgMirror* m, int it, float rr, double atol, int converged) {
    __hip_atomic_store(&m->rr, (double)rr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&m->atol, atol, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&m->converged, converged, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&m->it, it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    wait_stores();
    __hip_atomic_store(&m->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Progress of a batch: workgroup 0 of a reporting launch posts (tag << 32) | k as it starts.
// The stream runs launches in order, so k iterations of every scene have completed by then.
// The tag names the solve, so a late launch of an earlier solve is not taken for progress.
__device__ __forceinline__ void post_pace(unsigned long long* pace, unsigned tag, int k) {
    if (pace && blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(pace, ((unsigned long long)tag << 32) | (unsigned)k, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
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

// Store this workgroup's partial sums (thread 0 after the block sum) in row slot lb.
template <int NV, int BS = kBlock>
__device__ __forceinline__ void store_partials(double (&v)[NV], double* __restrict__ P, int nb, int lb,
                                               double (*lds)[BS / 64]) {
    block_sum<NV, BS>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) P[q * nb + lb] = v[q];
}

// Cells of logical workgroup L: cpt consecutive tiles of kBlock cells.
#define FOR_CELLS(v, nb)                                                     \
    const int L_ = xcd_tile((v).lb, nb);                                     \
    for (int u_ = 0; u_ < (v).cpt; ++u_)                                     \
        if (long t = ((long)L_ * (v).cpt + u_) * kBlock + threadIdx.x; t < (v).N)

// Prologue of every pass A: finish iteration k from pass B's partial sums (r.r, r.z; b.b too
// at k = 0, where atol is formed), test convergence and the iteration cap, and derive beta.
// Workgroup 0 of the scene records the state.  Returns true when the solve stops (in every
// thread).
template <int BS = kBlock>
__device__ bool finish_iteration(const PcgView& a, double (*lds)[BS / 64], float& beta) {
    CgCtl* c = a.ctl;
    const int k = a.k;
    float rr, rz, atol_sq;
    double atol;
    if (k == 0) {
        double tot[3];
        sum_partials<3, BS>(a.part_upd, a.nb_upd, lds, tot);
        rr = (float)tot[0];
        rz = (float)tot[1];
        const float bb = (float)tot[2];
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
    if (a.lb == 0 && threadIdx.x == 0) {
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
        }
    }
    return stop;
}

// r = b - A x ; z = D r ; partial sums r.r, r.z, b.b
template <int OP>
__global__ __launch_bounds__(kBlock) void k_pcg_init(const SceneDev* __restrict__ tab,
                                                     const MapEntry* __restrict__ map, float dt) {
    HFM_BATCH_ENTRY(tab, map);
    __shared__ double lds[3][kWaves];
    const PcgView v = make_view<OP>(a, lb, 0, dt);
    double s[3] = {0.0, 0.0, 0.0};
    FOR_CELLS(v, v.nb_upd) {
        int i, j, k;
        cell_ijk(t, v.n, i, j, k);
        float bc = v.b[t];
        if (OP == 1) {
            const float mean = v.ctl->mean;  // grid.py:240-241
            if (mean != 0.0f) {
                bc = bc - mean;
                v.b[t] = bc;
            }
        }
        const float* xs = v.x;
        auto dir = [xs](long q) { return xs[q]; };
        const float acc = Operator<OP>::apply(v, t, i, j, k, dir);
        const bool fixed = OP == 0 && (v.flags[t] & FLAG_FIXED);
        const float rc = Operator<OP>::combine(v, acc, xs[t], fixed, -1.0f) + 1.0f * bc;
        const float zc = v.D[t] * rc;
        v.r[t] = rc;
        v.z[t] = zc;
        s[0] += (double)(rc * rc);
        s[1] += (double)(rc * zc);
        s[2] += (double)(bc * bc);
    }
    store_partials<3>(s, v.part_upd, v.nb_upd, lb, lds);
    if (lb == 0 && threadIdx.x == 0) {  // a new solve: clear the previous one's stop
        v.ctl->done = 0;
        v.ctl->converged = 0;
        v.ctl->it = 0;
    }
}

// pass A of iteration k+1: finish iteration k (r.r, r.z, convergence, beta), then the direction
// update at the cell and its neighbours, SpMV and partial p.Ap
template <int OP>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(const SceneDev* __restrict__ tab,
                                                    const MapEntry* __restrict__ map, float dt, int k,
                                                    unsigned long long* pace, unsigned tag) {
    post_pace(pace, tag, k);
    HFM_BATCH_ENTRY(tab, map);
    __shared__ double lds[3][kWaves];
    const PcgView v = make_view<OP>(a, lb, k, dt);
    if (solve_stopped(v.ctl)) return;
    float beta;
    if (finish_iteration(v, lds, beta)) return;
    double s[1] = {0.0};
    const bool first = k == 0;
    FOR_CELLS(v, v.nb_dir) {
        int i, j, kk;
        cell_ijk(t, v.n, i, j, kk);
        const float* zs = v.z;
        const float* ds = v.d_old;
        auto dir = [zs, ds, beta, first](long q) { return first ? zs[q] : zs[q] + beta * ds[q]; };
        const float pc = dir(t);
        const float acc = Operator<OP>::apply(v, t, i, j, kk, dir);
        const bool fixed = OP == 0 && (v.flags[t] & FLAG_FIXED);
        const float apc = Operator<OP>::combine(v, acc, pc, fixed, 1.0f);
        v.d_new[t] = pc;
        v.Ap[t] = apc;
        s[0] += (double)(pc * apc);
    }
    store_partials<1>(s, v.part_dir, v.nb_dir, lb, lds);
}

// Pass A for the pressure operator, LDS-tiled.  A workgroup of 1024 threads owns a 16 x 16 x D
// block of cells.  It first forms the new direction p = z + beta p_old and loads the solid
// fraction on the block framed by a one-cell halo ((D+2) x 18 x 18 cells, all loads in flight at
// once) into LDS, then every thread applies the stencil to D/4 cells from LDS.  Per cell the
// arithmetic is exactly that of k_pcg_dir<1>.
constexpr int kT3 = 16, kF3 = kT3 + 2, kBlock3 = 1024;

template <int D>
__global__ __launch_bounds__(kBlock3) void k_pcg_dir_tile3(const SceneDev* __restrict__ tab,
                                                           const MapEntry* __restrict__ map, int kit,
                                                           unsigned long long* pace, unsigned tag) {
    post_pace(pace, tag, kit);
    HFM_BATCH_ENTRY(tab, map);
    __shared__ float P[D + 2][kF3][kF3];
    __shared__ float S[D + 2][kF3][kF3];
    __shared__ double lds[3][kBlock3 / 64];
    const PcgView v = make_view<1>(a, lb, kit, 0.0f);
    {
        __shared__ int stopped;
        if (threadIdx.x == 0) stopped = v.ctl->done;
        __syncthreads();
        if (stopped) return;
    }
    const int nb = v.nb_dir;
    const int n = v.n;
    const int L = xcd_tile(lb, nb);
    const int tjk = v.tiles_j * v.tiles_k;
    const int ti = L / tjk, rem = L % tjk;
    const int i0 = ti * D, j0 = (rem / v.tiles_k) * kT3, k0 = (rem % v.tiles_k) * kT3;
    const bool first = kit == 0;
    const float* __restrict__ zs = v.z;
    const float* __restrict__ ds = v.d_old;
    const float* __restrict__ as = v.alpha_s;
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
    if (finish_iteration<kBlock3>(v, lds, beta)) return;
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
    const float amin = v.amin;
    const int lk = threadIdx.x % kT3, lj = (threadIdx.x / kT3) % kT3, lq = threadIdx.x / (kT3 * kT3);
    const int j = j0 + lj, k = k0 + lk;
    double s[1] = {0.0};
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
            const float apc = 1.0f * acc / v.hh;
            const long t = ((long)i * n + j) * n + k;
            v.d_new[t] = pc;
            v.Ap[t] = apc;
            s[0] += (double)(pc * apc);
        }
    }
    store_partials<1, kBlock3>(s, v.part_dir, nb, lb, lds);
}

// pass B of iteration k+1: alpha from p.Ap, x and r update, z = D r, partial r.r and r.z
template <int OP>
__global__ __launch_bounds__(kBlock) void k_pcg_upd(const SceneDev* __restrict__ tab,
                                                    const MapEntry* __restrict__ map, int kit) {
    HFM_BATCH_ENTRY(tab, map);
    __shared__ double lds[2][kWaves];
    const PcgView v = make_view<OP>(a, lb, kit, 0.0f);
    CgCtl* c = v.ctl;
    if (solve_stopped(c)) return;
    const int L = xcd_tile(lb, v.nb_upd);
    // the operands of the first kPre tiles are loaded before the p.Ap prologue, so its latency
    // overlaps theirs
    constexpr int kPre = 4;
    float px[kPre], pd[kPre], pr[kPre], pa[kPre], pD[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        const long t = ((long)L * v.cpt + u) * kBlock + threadIdx.x;
        if (u < v.cpt && t < v.N) {
            px[u] = v.x[t];
            pd[u] = v.d_new[t];
            pr[u] = v.r[t];
            pa[u] = v.Ap[t];
            pD[u] = v.D[t];
        }
    }
    double tot[1];
    sum_partials<1>(v.part_dir, v.nb_dir, lds, tot);
    const float pAp = (float)tot[0];
    const float rr = c->slot_rr[kit & 1], rz = c->slot_rz[kit & 1];
    const float alpha = (rr > c->atol_sq) ? rz / pAp : 0.0f;
    double s[2] = {0.0, 0.0};
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        const long t = ((long)L * v.cpt + u) * kBlock + threadIdx.x;
        if (u < v.cpt && t < v.N) {
            const float xc = px[u] + alpha * pd[u];
            const float rc = pr[u] - alpha * pa[u];
            const float zc = pD[u] * rc;
            v.x[t] = xc;
            v.r[t] = rc;
            v.z[t] = zc;
            s[0] += (double)(rc * rc);
            s[1] += (double)(rc * zc);
        }
    }
    for (int u = kPre; u < v.cpt; ++u) {
        const long t = ((long)L * v.cpt + u) * kBlock + threadIdx.x;
        if (t >= v.N) break;
        const float xc = v.x[t] + alpha * v.d_new[t];
        const float rc = v.r[t] - alpha * v.Ap[t];
        const float zc = v.D[t] * rc;
        v.x[t] = xc;
        v.r[t] = rc;
        v.z[t] = zc;
        s[0] += (double)(rc * rc);
        s[1] += (double)(rc * zc);
    }
    store_partials<2>(s, v.part_upd, v.nb_upd, lb, lds);
}

// With ev0/ev1 the launch carries its own start/stop events (hipExtLaunchKernelGGL): the events
// time the dispatch itself, with no separate event packets around it.
template <class K, class... Args>
void launch_timed(K kernel, int blocks, int threads, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1,
                  Args... args) {
    if (blocks <= 0) return;
    if (ev0)
        hipExtLaunchKernelGGL(kernel, dim3(blocks), dim3(threads), 0, st, ev0, ev1, 0, args...);
    else
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(threads), 0, st, args...);
}

}  // namespace

// PCG geometry of a scene.  Init, pass B and the flat pass A cover cpt consecutive tiles of
// kBlock cells per workgroup: at least the scene's pcg_cpt (HFM_PCG_CPT), and enough that at most
// kMaxPcgBlocks workgroups run.  The pressure pass A is LDS-tiled (unless HFM_PCG_FLAT_DIR) with
// planes D = 8, else 16, else 4, the first that keeps at most kMaxDirBlocks workgroups
// (HFM_DIR_ICHUNK forces D); otherwise it is the flat kernel with the pass B geometry.
void pcg_geometry(const Scene* s, int* cpt, int* nb_upd, int* nb_dir, int* tiles_j, int* tiles_k, int* dtile) {
    const long tiles = (s->N + kBlock - 1) / kBlock;
    const long need = (tiles + kMaxPcgBlocks - 1) / kMaxPcgBlocks;
    const int c0 = s->pcg_cpt > 0 ? s->pcg_cpt : 1;
    const int c = (int)(need > c0 ? need : c0);
    *cpt = c;
    *nb_upd = (int)((tiles + c - 1) / c);
    const int n = s->n;
    const int tj = (n + kT3 - 1) / kT3, tk = tj;
    int D = 0;
    if (!s->flat_dir) {
        const int cand[3] = {8, 16, 4};
        for (int q = 0; q < 3 && !D; ++q) {
            const int d = cand[q];
            if (s->dir_ichunk > 0 && d != s->dir_ichunk) continue;
            const long blocks = (long)tj * tk * ((n + d - 1) / d);
            if (blocks <= kMaxDirBlocks) D = d;
        }
    }
    if (!D) {
        *nb_dir = *nb_upd;
        *tiles_j = *tiles_k = *dtile = 0;
        return;
    }
    *nb_dir = tj * tk * ((n + D - 1) / D);
    *tiles_j = tj;
    *tiles_k = tk;
    *dtile = D;
}

void launch_pcg_init(const Launch& L, int op, float dt) {
    if (op == 1)
        launch_timed(k_pcg_init<1>, L.blocks, kBlock, L.stream, nullptr, nullptr, L.tab, L.map, dt);
    else
        launch_timed(k_pcg_init<0>, L.blocks, kBlock, L.stream, nullptr, nullptr, L.tab, L.map, dt);
}

void launch_pcg_dir(const Launch& L, Family fam, int op, float dt, int k, unsigned long long* pace, unsigned tag,
                    hipEvent_t ev0, hipEvent_t ev1) {
    switch (fam) {
        case FAM_DIR4:
            launch_timed(k_pcg_dir_tile3<4>, L.blocks, kBlock3, L.stream, ev0, ev1, L.tab, L.map, k, pace, tag);
            break;
        case FAM_DIR8:
            launch_timed(k_pcg_dir_tile3<8>, L.blocks, kBlock3, L.stream, ev0, ev1, L.tab, L.map, k, pace, tag);
            break;
        case FAM_DIR16:
            launch_timed(k_pcg_dir_tile3<16>, L.blocks, kBlock3, L.stream, ev0, ev1, L.tab, L.map, k, pace, tag);
            break;
        default:
            if (op == 1)
                launch_timed(k_pcg_dir<1>, L.blocks, kBlock, L.stream, ev0, ev1, L.tab, L.map, dt, k, pace, tag);
            else
                launch_timed(k_pcg_dir<0>, L.blocks, kBlock, L.stream, ev0, ev1, L.tab, L.map, dt, k, pace, tag);
    }
}

void launch_pcg_upd(const Launch& L, int op, int k, hipEvent_t ev0, hipEvent_t ev1) {
    if (op == 1)
        launch_timed(k_pcg_upd<1>, L.blocks, kBlock, L.stream, ev0, ev1, L.tab, L.map, k);
    else
        launch_timed(k_pcg_upd<0>, L.blocks, kBlock, L.stream, ev0, ev1, L.tab, L.map, k);
}

}  // namespace hfm
