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
    __shared__ double bcast[NV];
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
    if (OP == 1 && a.cg_fused) return;  // the fused form has no pass B
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

// ---------------------------------------------------------------------------------------------
// Fused pressure PCG: one launch per iteration (HFM_FLAG_FUSED_CG).
//
// Launch k starts from x_k, r_k, p_{k-1} in memory and the partial sums of launch k-1:
// r_k.r_k, r_k.z_k, z_k.A z_k and z_k.A p_{k-1} (z = D r).  Its prologue forms
//   beta = r_k.z_k / r_{k-1}.z_{k-1}   (0 at k = 0, or once converged, as warp)
//   p_k.A p_k = z.Az + 2 beta z.Ap_{k-1} + beta^2 p_{k-1}.A p_{k-1}   (exact algebra, A symmetric)
//   alpha = r_k.z_k / p_k.A p_k        (0 once converged, as warp)
// and then, on a 16 x 16 x D block with a two-cell halo staged in LDS:
//   p_k = z_k + beta p_{k-1}           on block + 2 halo
//   A p_k, r_{k+1} = r_k - alpha A p_k, z_{k+1} = D r_{k+1}   on block + 1 halo
//   x_{k+1} = x_k + alpha p_k, A z_{k+1} and the four partial sums   on the block
// Halo cells are recomputed by every block that needs them with the same operations on the same
// inputs, so all copies agree bit for bit.  Memory per iteration: reads x, r, p, D, alpha_s,
// writes x, r, p (8 words per cell) in one launch, against 12 words in two launches for the
// two-pass form.  r and p are double-buffered by iteration parity (neighbours read the old ones
// while owners write the new), as are the partial sums.
// ---------------------------------------------------------------------------------------------
constexpr int kF2 = kT3 + 4;

template <int D>
__global__ __launch_bounds__(kBlock3) void k_pcg_fused(const SceneDev* __restrict__ tab,
                                                       const MapEntry* __restrict__ map, int kit,
                                                       unsigned long long* pace, unsigned tag) {
    post_pace(pace, tag, kit);
    HFM_BATCH_ENTRY(tab, map);
    __shared__ float P2[D + 4][kF2][kF2];
    __shared__ float S2[D + 4][kF2][kF2];
    __shared__ float Z1[D + 2][kF3][kF3];
    __shared__ double lds[kFusedRows][kBlock3 / 64];
    CgCtl* c = a.ctl + 1;
    {
        __shared__ int stopped;
        if (threadIdx.x == 0) stopped = c->done;
        __syncthreads();
        if (stopped) return;
    }
    const int n = a.n;
    const long N = a.N;
    const int nb = a.nb_dir;
    const int L = xcd_tile(lb, nb);
    const int tjk = a.tiles_j * a.tiles_k;
    const int ti = L / tjk, rem = L % tjk;
    const int i0 = ti * D, j0 = (rem / a.tiles_k) * kT3, k0 = (rem % a.tiles_k) * kT3;
    const bool first = kit == 0;
    const float* __restrict__ rk = a.cg_r + (long)(kit & 1) * N;
    float* __restrict__ rn = a.cg_r + (long)((kit + 1) & 1) * N;
    const float* __restrict__ pold = a.cg_d + (long)((kit + 1) & 1) * N;
    float* __restrict__ pk = a.cg_d + (long)(kit & 1) * N;
    const float* __restrict__ Dg = a.pdiag;
    const float* __restrict__ as = a.alpha_s;
    float* __restrict__ x = a.p;
    const double* part_in = a.partials + (long)(kit & 1) * kFusedRows * kMaxDirBlocks;
    double* part_out = a.partials + (long)((kit + 1) & 1) * kFusedRows * kMaxDirBlocks;
    // ---- operand loads, issued before the prologue (they do not depend on its scalars)
    constexpr int kFr2 = (D + 4) * kF2 * kF2;
    constexpr int kP2 = (kFr2 + kBlock3 - 1) / kBlock3;
    float f_r[kP2], f_p[kP2], f_d[kP2], f_a[kP2];
#pragma unroll
    for (int u = 0; u < kP2; ++u) {
        const int f = threadIdx.x + u * kBlock3;
        f_r[u] = 0.0f;
        f_p[u] = 0.0f;
        f_d[u] = 0.0f;
        f_a[u] = 0.0f;
        if (f < kFr2) {
            const int fi = f / (kF2 * kF2), fr = f % (kF2 * kF2);
            const int gi = i0 + fi - 2, gj = j0 + fr / kF2 - 2, gk = k0 + fr % kF2 - 2;
            if (gi >= 0 && gi < n && gj >= 0 && gj < n && gk >= 0 && gk < n) {
                const long t = ((long)gi * n + gj) * n + gk;
                f_r[u] = rk[t];
                if (!first) f_p[u] = pold[t];
                f_d[u] = Dg[t];
                f_a[u] = as[t];
            }
        }
    }
    constexpr int kFr1 = (D + 2) * kF3 * kF3;
    constexpr int kP1 = (kFr1 + kBlock3 - 1) / kBlock3;
    float g_r[kP1], g_d[kP1], g_x[kP1];
#pragma unroll
    for (int u = 0; u < kP1; ++u) {
        const int f = threadIdx.x + u * kBlock3;
        g_r[u] = 0.0f;
        g_d[u] = 0.0f;
        g_x[u] = 0.0f;
        if (f < kFr1) {
            const int fi = f / (kF3 * kF3), fr = f % (kF3 * kF3);
            const int fj = fr / kF3, fk = fr % kF3;
            const int gi = i0 + fi - 1, gj = j0 + fj - 1, gk = k0 + fk - 1;
            if (gi >= 0 && gi < n && gj >= 0 && gj < n && gk >= 0 && gk < n) {
                const long t = ((long)gi * n + gj) * n + gk;
                g_r[u] = rk[t];
                g_d[u] = Dg[t];
                if (fi >= 1 && fi <= D && fj >= 1 && fj <= kT3 && fk >= 1 && fk <= kT3) g_x[u] = x[t];
            }
        }
    }
    // ---- prologue: scalars of iteration k from launch k-1's partial sums
    float rr, rz, atol_sq;
    double zaz, zap, atol;
    if (first) {
        double tot[5];
        sum_partials<5, kBlock3>(part_in, a.nb_upd, lds, tot);  // rows written by the init launch
        rr = (float)tot[0];
        rz = (float)tot[1];
        zaz = tot[2];
        zap = tot[3];
        const float bb = (float)tot[4];
        atol = c->tol * sqrt((double)bb);  // warp: atol = max(tol * ||b||, tol)
        if (atol < c->tol) atol = c->tol;
        atol_sq = (float)(atol * atol);
    } else {
        double tot[4];
        sum_partials<4, kBlock3>(part_in, nb, lds, tot);
        rr = (float)tot[0];
        rz = (float)tot[1];
        zaz = tot[2];
        zap = tot[3];
        atol = c->atol;
        atol_sq = c->atol_sq;
    }
    const bool conv = rr <= atol_sq;
    const bool stop = conv || kit >= c->maxeff;
    const float beta = first ? 0.0f : (rr > atol_sq ? rz / c->slot_rz[(kit + 1) & 1] : 0.0f);
    const double pap = first ? zaz
                             : zaz + 2.0 * (double)beta * zap + (double)beta * (double)beta * c->slot_pap[(kit + 1) & 1];
    const float alpha = rr > atol_sq ? rz / (float)pap : 0.0f;
    if (lb == 0 && threadIdx.x == 0) {
        c->slot_rr[kit & 1] = rr;
        c->slot_rz[kit & 1] = rz;
        c->slot_pap[kit & 1] = pap;
        if (first) {
            c->atol = atol;
            c->atol_sq = atol_sq;
        }
        if (stop) {
            c->done = 1;
            c->converged = conv;
            c->it = kit;
            mirror_final(a.mirror + 1, kit, rr, atol, conv);
        }
    }
    if (stop) return;
    // ---- stage 1: p_k and alpha_s on block + 2 halo
#pragma unroll
    for (int u = 0; u < kP2; ++u) {
        const int f = threadIdx.x + u * kBlock3;
        if (f < kFr2) {
            const int fi = f / (kF2 * kF2), fr = f % (kF2 * kF2);
            const float z = f_d[u] * f_r[u];
            P2[fi][fr / kF2][fr % kF2] = first ? z : z + beta * f_p[u];
            S2[fi][fr / kF2][fr % kF2] = f_a[u];
        }
    }
    __syncthreads();
    const float amin = a.amin;
    const float hh = a.hh;
    // ---- stage 2: A p_k, r_{k+1}, z_{k+1} on block + 1 halo
    float k_ap[kP1], k_rn[kP1], k_zn[kP1];
#pragma unroll
    for (int u = 0; u < kP1; ++u) {
        const int f = threadIdx.x + u * kBlock3;
        k_ap[u] = 0.0f;
        k_rn[u] = 0.0f;
        k_zn[u] = 0.0f;
        if (f < kFr1) {
            const int fi = f / (kF3 * kF3), fr = f % (kF3 * kF3);
            const int fj = fr / kF3, fk = fr % kF3;
            const int i = i0 + fi - 1, j = j0 + fj - 1, k = k0 + fk - 1;
            float zn = 0.0f;
            if (i >= 0 && i < n && j >= 0 && j < n && k >= 0 && k < n) {
                const int ci = fi + 1, cj = fj + 1, ck = fk + 1;
                const float pc = P2[ci][cj][ck], ac = S2[ci][cj][ck];
                float acc = 0.0f;
                if (i > 0) acc += face_fraction(S2[ci - 1][cj][ck], ac, amin) * (pc - P2[ci - 1][cj][ck]);
                if (i < n - 1) acc += face_fraction(ac, S2[ci + 1][cj][ck], amin) * (pc - P2[ci + 1][cj][ck]);
                if (j > 0) acc += face_fraction(S2[ci][cj - 1][ck], ac, amin) * (pc - P2[ci][cj - 1][ck]);
                if (j < n - 1) acc += face_fraction(ac, S2[ci][cj + 1][ck], amin) * (pc - P2[ci][cj + 1][ck]);
                if (k > 0) acc += face_fraction(S2[ci][cj][ck - 1], ac, amin) * (pc - P2[ci][cj][ck - 1]);
                if (k < n - 1) acc += face_fraction(ac, S2[ci][cj][ck + 1], amin) * (pc - P2[ci][cj][ck + 1]);
                const float ap = 1.0f * acc / hh;
                const float rnew = g_r[u] - alpha * ap;
                zn = g_d[u] * rnew;
                k_ap[u] = ap;
                k_rn[u] = rnew;
                k_zn[u] = zn;
            }
            Z1[fi][fj][fk] = zn;
        }
    }
    __syncthreads();
    // ---- stage 3: x_{k+1}, A z_{k+1} and the partial sums on the block
    double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < kP1; ++u) {
        const int f = threadIdx.x + u * kBlock3;
        if (f < kFr1) {
            const int fi = f / (kF3 * kF3), fr = f % (kF3 * kF3);
            const int fj = fr / kF3, fk = fr % kF3;
            const int i = i0 + fi - 1, j = j0 + fj - 1, k = k0 + fk - 1;
            const bool inner = fi >= 1 && fi <= D && fj >= 1 && fj <= kT3 && fk >= 1 && fk <= kT3;
            if (inner && i < n && j < n && k < n) {
                const int ci = fi + 1, cj = fj + 1, ck = fk + 1;
                const float zc = Z1[fi][fj][fk], ac = S2[ci][cj][ck];
                float acc = 0.0f;
                if (i > 0) acc += face_fraction(S2[ci - 1][cj][ck], ac, amin) * (zc - Z1[fi - 1][fj][fk]);
                if (i < n - 1) acc += face_fraction(ac, S2[ci + 1][cj][ck], amin) * (zc - Z1[fi + 1][fj][fk]);
                if (j > 0) acc += face_fraction(S2[ci][cj - 1][ck], ac, amin) * (zc - Z1[fi][fj - 1][fk]);
                if (j < n - 1) acc += face_fraction(ac, S2[ci][cj + 1][ck], amin) * (zc - Z1[fi][fj + 1][fk]);
                if (k > 0) acc += face_fraction(S2[ci][cj][ck - 1], ac, amin) * (zc - Z1[fi][fj][fk - 1]);
                if (k < n - 1) acc += face_fraction(ac, S2[ci][cj][ck + 1], amin) * (zc - Z1[fi][fj][fk + 1]);
                const float az = 1.0f * acc / hh;
                const float pc = P2[ci][cj][ck];
                const long t = ((long)i * n + j) * n + k;
                x[t] = g_x[u] + alpha * pc;
                pk[t] = pc;
                rn[t] = k_rn[u];
                v[0] += (double)(k_rn[u] * k_rn[u]);
                v[1] += (double)(k_rn[u] * k_zn[u]);
                v[2] += (double)(k_zn[u] * az);
                v[3] += (double)(k_zn[u] * k_ap[u]);
            }
        }
    }
    store_partials<4, kBlock3>(v, part_out, nb, lb, lds);
}

// Scalars of the first fused iteration from r_0 (written by k_pcg_init<1>): r.r, r.z, z.Az,
// z.Ap_{-1} = 0 and b.b, into the partial-sum buffer launch 0 reads.
__global__ __launch_bounds__(kBlock) void k_pcg_init_fused(const SceneDev* __restrict__ tab,
                                                           const MapEntry* __restrict__ map) {
    HFM_BATCH_ENTRY(tab, map);
    if (!a.cg_fused) return;
    __shared__ double lds[kFusedRows][kWaves];
    const PcgView v = make_view<1>(a, lb, 0, 0.0f);
    const float* r = a.cg_r;
    const float* Dg = a.pdiag;
    double s[kFusedRows] = {0.0, 0.0, 0.0, 0.0, 0.0};
    FOR_CELLS(v, v.nb_upd) {
        int i, j, k;
        cell_ijk(t, v.n, i, j, k);
        auto zf = [r, Dg](long q) { return Dg[q] * r[q]; };
        const float acc = Operator<1>::apply(v, t, i, j, k, zf);
        const float az = Operator<1>::combine(v, acc, 0.0f, false, 1.0f);
        const float rc = r[t], zc = zf(t), bc = v.b[t];
        s[0] += (double)(rc * rc);
        s[1] += (double)(rc * zc);
        s[2] += (double)(zc * az);
        s[4] += (double)(bc * bc);
    }
    store_partials<kFusedRows>(s, a.partials, v.nb_upd, lb, lds);
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

void launch_pcg_init_fused(const Launch& L) {
    launch_timed(k_pcg_init_fused, L.blocks, kBlock, L.stream, nullptr, nullptr, L.tab, L.map);
}

void launch_pcg_fused(const Launch& L, Family fam, int k, unsigned long long* pace, unsigned tag, hipEvent_t ev0,
                      hipEvent_t ev1) {
    switch (fam) {
        case FAM_FUSED4:
            launch_timed(k_pcg_fused<4>, L.blocks, kBlock3, L.stream, ev0, ev1, L.tab, L.map, k, pace, tag);
            break;
        case FAM_FUSED8:
            launch_timed(k_pcg_fused<8>, L.blocks, kBlock3, L.stream, ev0, ev1, L.tab, L.map, k, pace, tag);
            break;
        case FAM_FUSED16:
            launch_timed(k_pcg_fused<16>, L.blocks, kBlock3, L.stream, ev0, ev1, L.tab, L.map, k, pace, tag);
            break;
        default:
            break;
    }
}

void launch_pcg_upd(const Launch& L, int op, int k, hipEvent_t ev0, hipEvent_t ev1) {
    if (op == 1)
        launch_timed(k_pcg_upd<1>, L.blocks, kBlock, L.stream, ev0, ev1, L.tab, L.map, k);
    else
        launch_timed(k_pcg_upd<0>, L.blocks, kBlock, L.stream, ev0, ev1, L.tab, L.map, k);
}

}  // namespace hfm
