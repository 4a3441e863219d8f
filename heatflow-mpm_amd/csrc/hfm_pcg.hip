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
// Breakdown stop (no counterpart in warp, which runs to its cap whatever the scalars hold): the
// fp32 solve also stops, with converged = -1, where an iteration would take a step that is not a
// CG step: p.Ap not positive or alpha not finite (x and r then keep the last iterate), r.z of the
// previous iteration not positive or beta not finite, or a non-finite ||r||^2.  A converging solve
// never meets these conditions, so its arithmetic and bits are unchanged; a solve stalled below its
// floating-point floor can (tiny p.Ap of either sign), and a step with such an alpha would fill x
// with huge or non-finite values.
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
//
// Mixed precision (HFM_PREC_MIXED) runs each solve as iterative refinement.  A refinement step
// forms the true residual r = b - A x (fp32 arithmetic, ||r||^2 and ||b||^2 in fp64) and stops
// the solve when ||r|| <= atol = max(tol ||b||, tol); otherwise it scales r to unit RMS
// (sigma = ||r|| / sqrt(N)) and an inner PCG solves A^ d = r / sigma with A^ = c A (c = h^2 for
// the pressure operator, whose A^ is then sum a_f (p_c - p_nb) with entries of order one; c = 1
// for the heat operator) and Jacobi D^ = D / c.  The inner solve keeps r, z, p and A p in FP16
// (all arithmetic fp32, dot products fp64), runs the same two passes as the fp32 solve, stops
// when its ||r||^2 falls below max(eta^2 ||r_0||^2, min((atol / sigma)^2, ||r_0||^2 / 4))
// (eta = 1e-2), and adds its
// update to x directly: x += sigma c alpha p.  The FP16 vectors hold values of order one: |r^|
// <= sqrt(N), |z^| <= |r^| / sum a_f with sum a_f >= 3 amin.
#include <cstdlib>

#include <hip/hip_ext.h>

#include "hfm_device.h"
#include "hfm_scene.h"


namespace hfm {

namespace {

// One scene's view of one solve (op 0 heat, 1 pressure) at iteration k; V is the storage type of
// the iteration vectors (float, or _Float16 in the inner solve of a refinement).
template <class V>
struct PcgViewT {
    int n;
    long N;
    float h, hh, dt_hh, amin;
    float rhh;             // RN(1 / h^2)
    int fastdiv;           // division by h^2 through div_const's corrected product
    const float* alpha_s;  // pressure
    const float* kap;      // heat
    const uint8_t* flags;  // heat
    float* x;
    float* b;
    V* r;
    V* z;
    const V* d_old;
    V* d_new;
    V* Ap;
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
    int floor_stop;        // the pressure solve of a multigrid scene: stop at the fp32 floor
};
using PcgView = PcgViewT<float>;

template <bool IR>
struct VecOf {
    using type = float;
};
template <>
struct VecOf<true> {
    using type = _Float16;
};

constexpr float kIrEta2 = 1.0e-4f;  // inner solve: reduce ||r||^2 by eta^2, eta = 1e-2
constexpr int kIrInnerMax = 256;    // iterations of one inner solve at most (then a refinement step)
constexpr int kIrPart = kMaxDirBlocks + 3 * kMaxPcgBlocks;  // residual partials of a refinement step
// The floor stop of a multigrid-preconditioned fp32 solve (no counterpart in warp, whose Jacobi PCG
// stalls at the floor of its fp32 residual): with the V-cycle as preconditioner the iteration diverges
// once it runs below that floor (the recursively updated residual grows without bound), so the solve
// stops when ||r||^2 has set no new minimum for kFloorStall iterations or has grown kFloorGrowth times
// above its minimum.  A converging solve lowers ||r||^2 at (nearly) every iteration and never meets it.
constexpr int kFloorStall = 8;
constexpr float kFloorGrowth = 100.0f;

template <int OP, bool IR = false>
__device__ __forceinline__ PcgViewT<typename VecOf<IR>::type> make_view(const SceneG& a, int lb, int k, float dt) {
    PcgViewT<typename VecOf<IR>::type> v;
    v.n = a.n;
    v.N = a.N;
    v.h = a.h;
    v.hh = a.hh;
    v.rhh = a.rhh;
    v.fastdiv = a.fastdiv;
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
    if constexpr (IR) {
        v.r = a.ir_r;
        v.z = a.ir_z;
        v.Ap = a.ir_q;
        v.d_new = a.ir_d + (long)(k & 1) * a.N;
        v.d_old = a.ir_d + (long)((k & 1) ^ 1) * a.N;
    } else {
        v.r = a.cg_r;
        v.z = a.cg_z;
        v.Ap = a.cg_Ap;
        v.d_new = a.cg_d + (long)(k & 1) * a.N;
        v.d_old = a.cg_d + (long)((k & 1) ^ 1) * a.N;
    }
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
    v.floor_stop = OP == 1 && !IR && a.cg_mg;
    return v;
}

template <int OP>
struct Operator;

// grid.py:94-115 with alpha/beta applied by the caller.  Branch-free as Operator<0>: a neighbour
// outside the grid is the cell itself, whose term a_f (p_c - p_c) is +0 (a_f in [amin, 1]).
template <>
struct Operator<1> {
    template <class View, class Dir>
    __device__ static float apply(const View& a, long t, int i, int j, int k, const Dir& dir) {
        const int n = a.n;
        const long nn = (long)n * n;
        const float* as = a.alpha_s;
        const float ac = as[t];
        const float pc = dir(t);
        const long nb[6] = {i > 0 ? t - nn : t, i < n - 1 ? t + nn : t, j > 0 ? t - n : t,
                            j < n - 1 ? t + n : t,  k > 0 ? t - 1 : t,   k < n - 1 ? t + 1 : t};
        float an[6], pn[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            an[s] = as[nb[s]];
            pn[s] = dir(nb[s]);
        }
        float acc = 0.0f;
#pragma unroll
        for (int s = 0; s < 6; ++s)   // the lower cell of a face first, as face_fraction takes them
            acc += ((s & 1) ? face_fraction(ac, an[s], a.amin) : face_fraction(an[s], ac, a.amin)) * (pc - pn[s]);
        return acc;
    }
    // alpha * acc / h^2; the inner solve of a refinement applies A^ = h^2 A: alpha * acc
    template <bool IR = false, class View>
    __device__ static float combine(const View& a, float acc, float xc, bool fixed, float alpha) {
        (void)xc;
        (void)fixed;
        if constexpr (IR) return alpha * acc;
        return div_const(alpha * acc, a.hh, a.rhh, a.fastdiv);
    }
    // Jacobi of A^ = h^2 A: D / h^2
    __device__ static float inner_diag_scale(float rhh) { return rhh; }
};

// heat.py:55-105.  Branch-free: a neighbour outside the grid is replaced by the cell itself, whose
// term kf (x_c - x_c) is +0 (the combine ignores acc for a fixed cell, so the cell is free there);
// acc starts at +0 and a sum that starts at +0 never becomes -0, so adding +0 leaves it as
// skipping the term would, and every load of the stencil issues at once.
template <>
struct Operator<0> {
    template <class View, class Dir>
    __device__ static float apply(const View& a, long t, int i, int j, int k, const Dir& dir) {
        const int n = a.n;
        const long nn = (long)n * n;
        const float kc = a.kap[t];
        const float xc = dir(t);
        const long nb[6] = {i > 0 ? t - nn : t, i < n - 1 ? t + nn : t, j > 0 ? t - n : t,
                            j < n - 1 ? t + n : t,  k > 0 ? t - 1 : t,   k < n - 1 ? t + 1 : t};
        float kn[6], xn[6];
        bool fixed[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            kn[s] = a.kap[nb[s]];
            fixed[s] = (a.flags[nb[s]] & FLAG_FIXED) != 0;
            xn[s] = dir(nb[s]);
        }
        float acc = 0.0f;
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const float kf = face_kappa(kc, kn[s]);
            acc += fixed[s] ? kf * xc : kf * (xc - xn[s]);
        }
        return acc;
    }
    // fixed: alpha * x ; free: alpha * (x + dt/h^2 * acc) (also the inner operator: c = 1)
    template <bool IR = false, class View>
    __device__ static float combine(const View& a, float acc, float xc, bool fixed, float alpha) {
        if (fixed) return alpha * xc;
        return alpha * (xc + a.dt_hh * acc);
    }
    __device__ static float inner_diag_scale(float) { return 1.0f; }
};

__device__ __forceinline__ void cell_ijk(long t, int n, int& i, int& j, int& k) {
    split3((unsigned)t, (unsigned)n, (unsigned)n, i, j, k);
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

// Partial sums of the previous launch: every thread adds entries threadIdx.x, + blockDim.x, ...
// of each row, then reduce_bcast reduces them over the workgroup in a fixed order (wave trees,
// then waves in order) into every thread's out[0..nv), with one barrier that also broadcasts
// thread 0's `flag`; it returns the flag in every thread.  (Letting wave 0 alone load and reduce
// the rows measured slower: the other waves then idle at the barrier.)
// The first entry of each row is taken as the thread's value (no add), so nothing waits for these
// loads before the reduction uses them: loads issued after them (a kernel's data) stay in flight
// meanwhile.  Further entries (rows longer than the workgroup) are added in order, four per round
// with their loads issued together.  AGENT: the rows are read with agent-scope loads (rows that
// other workgroups of the same launch wrote); the sums are the same.
template <int NV, bool AGENT = false>
__device__ __forceinline__ void load_rows(const double* P, int nb, int nv, double (&acc)[NV]) {
    auto rd = [P](int i) -> double { return AGENT ? load_agent(P + i) : P[i]; };
    const int b0 = threadIdx.x, bd = blockDim.x;
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = (q < nv && b0 < nb) ? rd(q * nb + b0) : 0.0;
    int b = b0 + bd;
#pragma unroll 1
    for (; b + 3 * bd < nb; b += 4 * bd) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            if (q < nv) {
                const double v0 = rd(q * nb + b), v1 = rd(q * nb + b + bd), v2 = rd(q * nb + b + 2 * bd),
                             v3 = rd(q * nb + b + 3 * bd);
                acc[q] = acc[q] + v0;
                acc[q] = acc[q] + v1;
                acc[q] = acc[q] + v2;
                acc[q] = acc[q] + v3;
            }
        }
    }
#pragma unroll 1
    for (; b < nb; b += bd)
#pragma unroll
        for (int q = 0; q < NV; ++q)
            if (q < nv) acc[q] += rd(q * nb + b);
}

// nb: the number of row entries; a wave whose threads all lie past it holds zeros and skips its tree
// (the wave's total is then +0, which leaves every sum it joins unchanged).
template <int NV, int BS>
__device__ __forceinline__ int reduce_bcast(const double (&acc)[NV], int nv, int flag, double (&out)[NV], int nb) {
    __shared__ double wsum[NV][BS / 64];
    __shared__ int flag_s;
    const bool any = (int)(threadIdx.x & ~63u) < nb;   // uniform over the wave
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        if (q < nv) {
            const double w = any ? wave_sum(acc[q]) : 0.0;
            if ((threadIdx.x & 63) == 0) wsum[q][threadIdx.x >> 6] = w;
        }
    }
    if (threadIdx.x == 0) flag_s = flag;
    __syncthreads();
    const int f = flag_s;
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = q < nv ? sum_wave_values<BS / 64>(wsum[q]) : 0.0;
    return f;
}

// Operands of a pass A prologue, loaded together with the block's data: this workgroup's share
// of the partial sums of pass B (r.r, r.z; b.b too from the init launch at k = 0) and the control
// scalars written by earlier launches.
struct PassAIn {
    double acc[3];
    int done, brk, maxeff, total;
    float rz_prev, atol_sq, gl;
    double atol, tol;
    float rr_min;   // floor stop: smallest ||r||^2 before iteration k, and its iteration
    int k_min;
};

// IR: the inner solve of a refinement step, whose first launch takes r.r and r.z from the scale
// kernel (no b.b)
template <int BS, bool IR = false, class View>
__device__ __forceinline__ void pass_a_load(const View& a, PassAIn& p) {
    const CgCtl* c = a.ctl;
    const bool first = a.k == 0 && !IR;
    load_rows<3>(a.part_upd, a.nb_upd, first ? 3 : 2, p.acc);
    p.done = c->done;
    p.brk = c->brk;
    p.maxeff = c->maxeff;
    p.rz_prev = c->slot_rz[(a.k + 1) & 1];
    p.atol_sq = c->atol_sq;
    p.atol = c->atol;
    p.tol = c->tol;
    p.total = c->ir_total;
    p.gl = c->ir_gl;
    p.rr_min = c->slot_rrmin[(a.k + 1) & 1];
    p.k_min = c->slot_kmin[(a.k + 1) & 1];
}

// Prologue of every pass A: finish iteration k from pass B's partial sums (r.r, r.z; b.b too at
// k = 0, where atol is formed), test convergence and the iteration cap, and derive beta.
// Workgroup 0 of the scene records the state.  Returns true (in every thread) when the
// workgroup has nothing to do: the solve stopped in an earlier launch, or stops now.
// In the inner solve of a refinement (IR) the threshold is set at k = 0 from the scale kernel's
// r.r, and a stop ends only the inner solve: the host then starts the next refinement step.
template <int BS = kBlock, bool IR = false, class View>
__device__ bool finish_iteration(const View& a, const PassAIn& p, float& beta) {
    CgCtl* c = a.ctl;
    const int k = a.k;
    double tot[3];
    if (reduce_bcast<3, BS>(p.acc, k == 0 && !IR ? 3 : 2, p.done, tot, a.nb_upd)) return true;
    const float rr = (float)tot[0];
    const float rz = (float)tot[1];
    float atol_sq = p.atol_sq;
    double atol = p.atol;
    if (IR && k == 0) {
        // max(eta^2 ||r_0||^2, (atol / sigma)^2), the latter capped at ||r_0||^2 / 4 so that every
        // inner solve reduces the residual (FP16 rounding of r_0 could otherwise leave it above a
        // threshold that the true residual has nearly reached)
        const float rel = kIrEta2 * rr, gl = fminf(p.gl, 0.25f * rr);
        atol_sq = rel > gl ? rel : gl;
    } else if (k == 0) {
        const float bb = (float)tot[2];
        // warp: atol = max(tol * ||b||, tol)
        atol = p.tol * sqrt((double)bb);
        if (atol < p.tol) atol = p.tol;
        atol_sq = (float)(atol * atol);
    }
    const bool conv = rr <= atol_sq;
    // beta of iteration k (warp: rz_new / rz_old while rr > atol^2); the inner FP16 solve of a
    // refinement takes beta = 0 on a breakdown (r.z of the previous iteration not positive), the
    // fp32 solve stops there (and on pass B's alpha breakdown, or a non-finite ||r||^2)
    bool brk = false;
    if (k == 0 || !(rr > atol_sq)) {
        beta = 0.0f;
    } else if (IR) {
        beta = p.rz_prev > 0.0f ? rz / p.rz_prev : 0.0f;
    } else {
        beta = rz / p.rz_prev;
        brk = !(p.rz_prev > 0.0f) || !(fabsf(beta) <= 3.4028235e38f);
    }
    if (!IR) brk = brk || p.brk || !(rr <= 3.4028235e38f);
    // the floor stop of a multigrid scene (see kFloorStall); iteration 0 sets the first minimum
    float rr_min = rr;
    int k_min = k;
    bool floor = false;
    if (a.floor_stop && k > 0) {
        const bool lower = rr < p.rr_min;
        rr_min = lower ? rr : p.rr_min;
        k_min = lower ? k : p.k_min;
        floor = !conv && (k - k_min >= kFloorStall || rr > kFloorGrowth * p.rr_min);
    }
    const bool stop = conv || k >= p.maxeff || brk || floor;
    if (a.lb == 0 && threadIdx.x == 0) {
        c->slot_rr[k & 1] = rr;
        c->slot_rz[k & 1] = rz;
        if (a.floor_stop) {
            c->slot_rrmin[k & 1] = rr_min;
            c->slot_kmin[k & 1] = k_min;
        }
        if (k == 0) {
            if (!IR) c->atol = atol;
            c->atol_sq = atol_sq;
        }
        if (stop) {
            const int cv = conv ? 1 : (brk ? -1 : (floor ? 2 : 0));
            c->done = 1;
            c->converged = cv;
            c->it = k;
            if (IR) {
                c->ir_total = p.total + k;
                __hip_atomic_store(&a.mirror->it, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                wait_stores();
                __hip_atomic_store(&a.mirror->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            } else {
                mirror_final(a.mirror, k, rr, atol, cv);
            }
        }
    }
    return stop;
}

// The heater energy balance of the substep: workgroup 0 of the heat solve's first launch sums the
// heat prep's per-workgroup source rises in a fixed order (fp64) into ctl[0].src_dt and the host
// mirror.
template <class A>
__device__ __forceinline__ void sum_source_rises(const A& a, int lb) {
    if (lb != 0 || !a.rate_on) return;
    double acc[1], tot[1];
    load_rows<1>(a.src_part, nblocks(a.N), 1, acc);
    reduce_bcast<1, kBlock>(acc, 1, 0, tot, nblocks(a.N));
    if (threadIdx.x == 0) {
        a.ctl[0].src_dt = tot[0];
        __hip_atomic_store(&a.mirror[0].src_dt, tot[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// r = b - A x ; z = D r ; partial sums r.r, r.z, b.b
template <int OP, class Src>
__global__ __launch_bounds__(kBlock) void k_pcg_init(Src src, float dt) {
    HFM_ENTRY(src);
    if (OP == 0) sum_source_rises(a, lb);
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
        v.ctl->brk = 0;
        v.ctl->converged = 0;
        v.ctl->it = 0;
    }
}

// pass A of iteration k+1: finish iteration k (r.r, r.z, convergence, beta), then the direction
// update at the cell and its neighbours, SpMV and partial p.Ap
template <int OP, bool IR, class Src>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(Src src, float dt, int k,
                                                    unsigned long long* pace, unsigned tag) {
    using V = typename VecOf<IR>::type;
    post_pace(pace, tag, k);
    HFM_ENTRY(src);
    if (batch_solve_done(src, a.ctl + OP)) return;
    if (IR && a.ctl[OP].done) return;   // as k_pcg_dir_tile3: launches queued past an inner solve's stop
    __shared__ double lds[3][kWaves];
    const auto v = make_view<OP, IR>(a, lb, k, dt);
    PassAIn pre;
    pass_a_load<kBlock, IR>(v, pre);
    float beta;
    if (finish_iteration<kBlock, IR>(v, pre, beta)) return;
    double s[1] = {0.0};
    const bool first = k == 0;
    FOR_CELLS(v, v.nb_dir) {
        int i, j, kk;
        cell_ijk(t, v.n, i, j, kk);
        const V* zs = v.z;
        const V* ds = v.d_old;
        auto dir = [zs, ds, beta, first](long q) {
            return first ? (float)zs[q] : (float)zs[q] + beta * (float)ds[q];
        };
        const float pc = dir(t);
        const float acc = Operator<OP>::apply(v, t, i, j, kk, dir);
        const bool fixed = OP == 0 && (v.flags[t] & FLAG_FIXED);
        const float apc = Operator<OP>::template combine<IR>(v, acc, pc, fixed, 1.0f);
        const V pv = (V)pc, av = (V)apc;
        v.d_new[t] = pv;
        v.Ap[t] = av;
        s[0] += (double)((float)pv * (float)av);
    }
    store_partials<1>(s, v.part_dir, v.nb_dir, lb, lds);
}

// Pass A for the pressure operator, LDS-tiled.  A workgroup of 1024 threads owns a 16 x 16 x D
// block of cells and stages it framed by a one-cell halo ((D+2) x 18 x 18 cells) in LDS: the
// solid fraction, and once the prologue has formed beta, the new direction p = z + beta p_old.
// For the staging, 972 threads each own one (j, k) position of the 18 x 18 frame plane and step
// through every third plane, so a load costs an address increment; all loads are in flight
// before the prologue, whose latency they overlap.  A frame cell outside the box takes the
// values of the clamped in-box cell; for every frame cell a stencil reaches, that is the very cell
// whose stencil reaches it, so the missing neighbour's term is a * (p - p) = +0 and the sum is
// exactly the one k_pcg_dir<1> forms by skipping the term, with no branch.  Every thread then
// applies the stencil to D/4 cells from LDS.
constexpr int kT3 = 16, kF3 = kT3 + 2, kBlock3 = 1024;
constexpr int kPlane3 = kF3 * kF3;            // 324 positions of a frame plane
constexpr int kGroups3 = kBlock3 / kPlane3;   // 3 plane groups

template <int D, bool IR, class Src>
__global__ __launch_bounds__(kBlock3) void k_pcg_dir_tile3(Src src, int kit, unsigned long long* pace,
                                                           unsigned tag) {
    using V = typename VecOf<IR>::type;
    post_pace(pace, tag, kit);
    HFM_ENTRY(src);
    if (batch_solve_done(src, a.ctl + 1)) return;
    // an inner solve of the refinement takes a few iterations, so launches queued past its stop are
    // common: they skip the frame's loads (a scene alone otherwise issues them with the prologue's)
    if (IR && a.ctl[1].done) return;
    __shared__ float P[D + 2][kF3][kF3];
    __shared__ float S[D + 2][kF3][kF3];
    __shared__ double lds[1][kBlock3 / 64];
    const auto v = make_view<1, IR>(a, lb, kit, 0.0f);
    const int nb = v.nb_dir;
    const int n = v.n;
    const int L = xcd_tile(lb, nb);
    const int tjk = v.tiles_j * v.tiles_k;
    const int ti = L / tjk, rem = L % tjk;
    const int i0 = ti * D, j0 = (rem / v.tiles_k) * kT3, k0 = (rem % v.tiles_k) * kT3;
    const bool first = kit == 0;
    const V* __restrict__ zs = v.z;
    const V* __restrict__ ds = v.d_old;
    const float* __restrict__ as = v.alpha_s;
    const unsigned nn = (unsigned)n * n;
    // staging: thread -> (plane group, frame position); out-of-box frame cells load from a clamped
    // in-box address and hold 0
    constexpr int kM = (D + 2 + kGroups3 - 1) / kGroups3;
    const int grp = threadIdx.x / kPlane3, pos = threadIdx.x % kPlane3;
    const int fj = pos / kF3, fk = pos % kF3;
    const int gj = j0 + fj - 1, gk = k0 + fk - 1;
    const unsigned t_jk = (unsigned)clampi(gj, 0, n - 1) * n + clampi(gk, 0, n - 1);
    // the prologue's operands first (vector loads complete in issue order, so the prologue then waits
    // for these only), then the frame, in flight during the prologue
    PassAIn pre;
    pass_a_load<kBlock3, IR>(v, pre);
    float fz[kM], fd[kM], fa[kM];
#pragma unroll
    for (int m = 0; m < kM; ++m) {
        const int fi = grp + kGroups3 * m, gi = i0 + fi - 1;
        const unsigned t = (unsigned)clampi(gi, 0, n - 1) * nn + t_jk;
        fz[m] = (float)ld32(zs, t);
        fd[m] = first ? 0.0f : (float)ld32(ds, t);
        fa[m] = ld32(as, t);
    }
    float beta;
    if (finish_iteration<kBlock3, IR>(v, pre, beta)) return;
    if (grp < kGroups3) {
#pragma unroll
        for (int m = 0; m < kM; ++m) {
            const int fi = grp + kGroups3 * m;
            if (fi < D + 2) {
                P[fi][fj][fk] = first ? fz[m] : fz[m] + beta * fd[m];
                S[fi][fj][fk] = fa[m];
            }
        }
    }
    __syncthreads();
    const float amin = v.amin, hh = v.hh, rhh = a.rhh;
    const int fastdiv = a.fastdiv;
    const int lk = threadIdx.x % kT3, lj = (threadIdx.x / kT3) % kT3, lq = threadIdx.x / (kT3 * kT3);
    const int j = j0 + lj, k = k0 + lk;
    double sum[1] = {0.0};
    if (j < n && k < n) {
        const int sj = lj + 1, sk = lk + 1;
        const unsigned tb = ((unsigned)i0 * n + j) * n + k;
#pragma unroll
        for (int m = 0; m < D / 4; ++m) {
            const int li = lq + 4 * m, i = i0 + li;
            if (i >= n) break;
            const int si = li + 1;
            const float pc = P[si][sj][sk], ac = S[si][sj][sk];
            float acc = 0.0f;
            acc += face_fraction(S[si - 1][sj][sk], ac, amin) * (pc - P[si - 1][sj][sk]);
            acc += face_fraction(ac, S[si + 1][sj][sk], amin) * (pc - P[si + 1][sj][sk]);
            acc += face_fraction(S[si][sj - 1][sk], ac, amin) * (pc - P[si][sj - 1][sk]);
            acc += face_fraction(ac, S[si][sj + 1][sk], amin) * (pc - P[si][sj + 1][sk]);
            acc += face_fraction(S[si][sj][sk - 1], ac, amin) * (pc - P[si][sj][sk - 1]);
            acc += face_fraction(ac, S[si][sj][sk + 1], amin) * (pc - P[si][sj][sk + 1]);
            float apc;
            if constexpr (IR)
                apc = acc;  // A^ = h^2 A
            else
                apc = div_const(1.0f * acc, hh, rhh, fastdiv);
            const unsigned t = tb + (unsigned)li * nn;
            const V pv = (V)pc, av = (V)apc;
            st32(v.d_new, t, pv);
            st32(v.Ap, t, av);
            sum[0] += (double)((float)pv * (float)av);
        }
    }
    store_partials<1, kBlock3>(sum, v.part_dir, nb, lb, lds);
}

// pass B of iteration k+1: alpha from p.Ap, x and r update, z = D r, partial r.r and r.z
// In the inner solve of a refinement (IR) x is the fp32 solution itself: x += (sigma c alpha) p,
// r and z are FP16 and z = (D / c) r.
template <int OP, bool IR, class Src>
__global__ __launch_bounds__(kBlock) void k_pcg_upd(Src src, int kit) {
    using V = typename VecOf<IR>::type;
    HFM_ENTRY(src);
    if (OP == 1 && a.cg_fused) return;  // the fused form has no pass B
    if (batch_solve_done(src, a.ctl + OP)) return;
    __shared__ double lds[2][kWaves];
    const auto v = make_view<OP, IR>(a, lb, kit, 0.0f);
    const CgCtl* c = v.ctl;
    const int L = xcd_tile(lb, v.nb_upd);
    const float dsc = IR ? Operator<OP>::inner_diag_scale(v.rhh) : 1.0f;
    // a multigrid-preconditioned pressure solve replaces z and the r.z sums in the V-cycle that
    // follows: this pass neither loads D nor stores z (6 words per cell instead of 8)
    const bool mgz = OP == 1 && a.cg_mg;
    // the partial sums of pass A and the scalars first (vector loads complete in issue order, so the
    // p.Ap reduction then waits for these only), then the operands of the first kPre tiles, in flight
    // during the reduction
    double acc[1];
    load_rows<1>(v.part_dir, v.nb_dir, 1, acc);
    const int done = c->done;
    const float rr = c->slot_rr[kit & 1], rz = c->slot_rz[kit & 1], atol_sq = c->atol_sq;
    const float step = IR ? c->ir_step : 1.0f;
    constexpr int kPre = 4;
    float px[kPre], pd[kPre], pr[kPre], pa[kPre], pD[kPre];
    const unsigned N32 = (unsigned)v.N;
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        const unsigned t = ((unsigned)L * v.cpt + u) * kBlock + threadIdx.x;
        if (u < v.cpt && t < N32) {
            px[u] = ld32(v.x, t);
            pd[u] = (float)ld32(v.d_new, t);
            pr[u] = (float)ld32(v.r, t);
            pa[u] = (float)ld32(v.Ap, t);
            pD[u] = mgz ? 0.0f : ld32(v.D, t);
        }
    }
    double tot[1];
    if (reduce_bcast<1, kBlock>(acc, 1, done, tot, v.nb_dir)) return;
    const float pAp = (float)tot[0];
    // alpha = 0 on a breakdown (p.Ap not positive or not a number; fp32 also alpha not finite): x and
    // r stay.  The inner FP16 solve of a refinement then runs to its cap and the refinement step
    // takes over; the fp32 solve is stopped by the next pass A (brk, written by workgroup 0 and read
    // by later launches only).
    float alpha = 0.0f;
    if (rr > atol_sq) {
        if (IR) {
            alpha = pAp > 0.0f ? rz / pAp : 0.0f;
        } else {
            alpha = rz / pAp;
            if (!(pAp > 0.0f) || !(fabsf(alpha) <= 3.4028235e38f)) {
                alpha = 0.0f;
                if (lb == 0 && threadIdx.x == 0) v.ctl->brk = 1;
            }
        }
    }
    const float xa = IR ? step * alpha : alpha;
    double s[2] = {0.0, 0.0};
    auto update = [&](unsigned t, float x0, float d0, float r0, float a0, float D0) {
        const float xc = x0 + xa * d0;
        const V rv = (V)(r0 - alpha * a0);
        const float rc = (float)rv;
        const V zv = (V)((IR ? D0 * dsc : D0) * rc);
        const float zc = (float)zv;
        st32(v.x, t, xc);
        st32(v.r, t, rv);
        s[0] += (double)(rc * rc);
        if (!mgz) {
            st32(v.z, t, zv);
            s[1] += (double)(rc * zc);
        }
    };
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        const unsigned t = ((unsigned)L * v.cpt + u) * kBlock + threadIdx.x;
        if (u < v.cpt && t < N32) update(t, px[u], pd[u], pr[u], pa[u], pD[u]);
    }
    for (int u = kPre; u < v.cpt; ++u) {
        const unsigned t = ((unsigned)L * v.cpt + u) * kBlock + threadIdx.x;
        if (t >= N32) break;
        update(t, ld32(v.x, t), (float)ld32(v.d_new, t), (float)ld32(v.r, t), (float)ld32(v.Ap, t),
               mgz ? 0.0f : ld32(v.D, t));
    }
    store_partials<2>(s, v.part_upd, v.nb_upd, lb, lds);
}

// ---------------------------------------------------------------------------------------------
// Mixed precision: the refinement step (HFM_PREC_MIXED).
// ---------------------------------------------------------------------------------------------

// True residual of x: r = b - A x (fp32, as the fp32 solve's first residual), ||r||^2 in fp64
// (products and sums in double); the first step of a solve also forms ||b||^2 and removes the
// pressure rhs mean (grid.py:240-241), and clears the solve's state.
template <int OP, class Src>
__global__ __launch_bounds__(kBlock) void k_ir_residual(Src src, float dt, int first) {
    HFM_ENTRY(src);
    CgCtl* c = a.ctl + OP;
    if (!first && c->ir_end) return;  // set by an earlier launch: this scene's solve has ended
    if (OP == 0 && first) sum_source_rises(a, lb);
    __shared__ double lds[2][kWaves];
    const PcgView v = make_view<OP>(a, lb, 0, dt);
    double s[2] = {0.0, 0.0};
    FOR_CELLS(v, v.nb_upd) {
        int i, j, k;
        cell_ijk(t, v.n, i, j, k);
        float bc = v.b[t];
        if (OP == 1 && first) {
            const float mean = c->mean;
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
        v.r[t] = rc;
        s[0] += (double)rc * (double)rc;
        s[1] += (double)bc * (double)bc;
    }
    store_partials<2>(s, a.partials + kIrPart, v.nb_upd, lb, lds);
    if (lb == 0 && threadIdx.x == 0) {
        if (first) {
            c->done = 0;
            c->brk = 0;
            c->converged = 0;
            c->it = 0;
            c->ir_total = 0;
            c->ir_outer = 0;
            c->ir_end = 0;
            c->ir_inner_conv = 0;
        } else {
            c->ir_inner_conv = c->converged == 1;   // the inner solve that just ended reached its tolerance
        }
    }
}

// Decide the refinement step from the residual's partial sums: stop the solve when
// ||r|| <= atol (atol = max(tol ||b||, tol), formed at the first step), when the last step cut
// ||r|| by less than sqrt(2), or, in a multigrid scene, by less than 4 after an inner solve that
// reached its tolerance (its 1-3 multigrid iterations cut the true residual about 1/eta = 100 times
// away from the floor; Jacobi inner solves run hundreds of FP16 iterations and can cut it less, so
// they keep the first rule) -- either way the true residual has reached the floor of its fp32
// evaluation, which can lie above atol: the fp32 solve, which tests its recursively updated residual,
// stops there too -- at the iteration cap, or on a non-finite residual; otherwise write r^ = r / sigma
// and z^ = (D / c) r^ in FP16 with their partial sums, which the inner solve's first pass A
// reads, and set the inner solve's state.
// step: the refinement step (0 the first), which the host counts; ||r||^2 of each step is kept in
// slot step & 1, so this launch reads the previous step's slot and workgroup 0 writes its own (every
// workgroup must reach the same decision, so no workgroup may read a value that workgroup 0 of the
// same launch writes).
template <int OP, class Src>
__global__ __launch_bounds__(kBlock) void k_ir_scale(Src src, int step) {
    HFM_ENTRY(src);
    CgCtl* c = a.ctl + OP;
    if (c->ir_end) return;
    const bool first = step == 0;
    __shared__ double lds[2][kWaves];
    const auto v = make_view<OP, true>(a, lb, 0, 0.0f);
    const double* part = a.partials + kIrPart;
    double acc[2];
    load_rows<2>(part, v.nb_upd, first ? 2 : 1, acc);
    const double tol = c->tol;
    double atol = c->atol;
    const int total = c->ir_total, outer = step, maxiter = c->ir_maxiter;
    const double rr_prev = c->ir_rr2[(step + 1) & 1];
    double tot[2];
    reduce_bcast<2, kBlock>(acc, first ? 2 : 1, 0, tot, v.nb_upd);
    const double rr = tot[0];
    if (first) {
        atol = tol * sqrt(tot[1]);
        if (atol < tol) atol = tol;
    }
    const bool finite = rr <= 1.0e300;   // false for NaN and infinity
    const bool conv = finite && rr <= atol * atol;
    const bool stalled = !first && finite && !conv &&
                         (rr > 0.5 * rr_prev || (a.cg_mg && c->ir_inner_conv && rr > 0.0625 * rr_prev));
    if (conv || stalled || !finite || total >= maxiter) {
        if (lb == 0 && threadIdx.x == 0) {
            c->ir_end = 1;
            c->done = 1;
            c->converged = conv;
            c->it = total;
            if (first) c->atol = atol;
            CgMirror* m = v.mirror;
            __hip_atomic_store(&m->rr, rr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&m->atol, atol, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&m->converged, conv ? 1 : (stalled ? 2 : (finite ? 0 : -1)), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&m->it, total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&m->outer, outer, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&m->end, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            wait_stores();
            __hip_atomic_store(&m->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        return;
    }
    const double sigma = sqrt(rr / (double)v.N);
    const float inv = (float)(1.0 / sigma);
    const float dsc = Operator<OP>::inner_diag_scale(v.rhh);
    const float* __restrict__ r32 = a.cg_r;  // the residual kernel's fp32 r
    double s[2] = {0.0, 0.0};
    FOR_CELLS(v, v.nb_upd) {
        const _Float16 rv = (_Float16)(r32[t] * inv);
        const float rc = (float)rv;
        const _Float16 zv = (_Float16)((v.D[t] * dsc) * rc);
        const float zc = (float)zv;
        v.r[t] = rv;
        v.z[t] = zv;
        s[0] += (double)(rc * rc);
        s[1] += (double)(rc * zc);
    }
    store_partials<2>(s, v.part_upd, v.nb_upd, lb, lds);
    if (lb == 0 && threadIdx.x == 0) {
        const double cop = OP == 1 ? (double)a.hh : 1.0;
        c->ir_outer = outer + 1;
        c->ir_gl = (float)(atol * atol / rr * (double)v.N);   // (atol / sigma)^2
        c->ir_step = (float)(sigma * cop);
        c->ir_rr2[step & 1] = rr;
        if (first) c->atol = atol;
        c->done = 0;
        c->converged = 0;
        c->it = 0;
        c->maxeff = maxiter - total < kIrInnerMax ? maxiter - total : kIrInnerMax;
    }
}

// ---------------------------------------------------------------------------------------------
// One-launch pressure PCG (HFM_FLAG_FUSED_CG): warp's Jacobi PCG in the single-reduction form of
// Chronopoulos and Gear (J. Comput. Appl. Math. 25, 1989).  With u = D r (never stored), w = A u and
// s = A p carried by the recurrence s = w + beta s, one iteration is
//   beta_k = g_k / g_{k-1},  p.Ap_k = d_k - beta_k^2 p.Ap_{k-1}  (g = r.u, d = w.u; k = 0: beta 0, p.Ap d_0)
//   alpha_k = g_k / p.Ap_k
//   p_k = u_k + beta_k p_{k-1},  s_k = w_k + beta_k s_{k-1},  x_{k+1} = x_k + alpha_k p_k,
//   r_{k+1} = r_k - alpha_k s_k,  u_{k+1} = D r_{k+1},  w_{k+1} = A u_{k+1}
// and the only global reduction, g, d and r.r of step k + 1, closes it.  In exact arithmetic these are
// warp's iterates (p_k.A p_{k-1} = 0 turns z.Az + 2 beta z.Ap_{k-1} + beta^2 p.Ap_{k-1} into d - beta^2
// p.Ap_{k-1}, and s_k = A p_k); the stopping rule and its count are warp's, on r.r of the recursive r.
// In fp32 the iteration counts equal the two-pass form's and the solutions agree to ~1e-7 on the
// pressure systems of the coupled and plume scenes.
//
// Launch k works on a 16 x 16 x D block framed by a one-cell halo (as pass A).  Every frame cell gets
// s_k, r_{k+1} and u_{k+1} from r_k, w_k, s_{k-1} and D (a frame cell outside the box takes the clamped
// in-box cell's values, so its stencil term is +0); u_{k+1} and the solid fraction go to LDS, and then
// w_{k+1} = A u_{k+1} is formed from LDS.  The halo cells (the frame minus the block: the two end planes
// and the 68-cell rings of the inner planes) are handled in a compact numbering, at most two per
// thread; the block's own cells in the stencil's layout, D / 4 per thread, which also writes their
// r_{k+1}, s_k, p_k, x_{k+1} and then w_{k+1}, and sums r.u and r.r (then w.u) over its cells in that
// order.  Neighbouring blocks read r, w and s at halo cells while the owners write them, so those three
// are double-buffered by iteration parity (r_k, w_k in buffer k & 1, s_k in buffer k & 1); p and x are
// read and written by their owners only.  Memory per iteration: reads r, w, s, D, alpha_s, x, p; writes
// r, w, s, p, x (12 words per cell), one launch and one reduction instead of two of each.
// ---------------------------------------------------------------------------------------------
template <int D>
constexpr int kHalo = 2 * kPlane3 + D * (kPlane3 - kT3 * kT3);

// Frame coordinates of halo cell h: the end planes fi = 0 and D + 1 (18 x 18 each), then for each inner
// plane its ring: the rows fj = 0 and 17, then the columns fk = 0 and 17 of rows 1 .. 16.
template <int D>
__device__ __forceinline__ void halo_cell(int h, int& fi, int& fj, int& fk) {
    constexpr int kRing = kPlane3 - kT3 * kT3;   // 68
    if (h < 2 * kPlane3) {
        const bool lo = h < kPlane3;
        const int q = lo ? h : h - kPlane3;
        fi = lo ? 0 : D + 1;
        fj = q / kF3;
        fk = q % kF3;
    } else {
        const int q = h - 2 * kPlane3;
        fi = 1 + q / kRing;
        const int r = q % kRing;
        if (r < 2 * kF3) {
            fj = r < kF3 ? 0 : kF3 - 1;
            fk = r < kF3 ? r : r - kF3;
        } else {
            const int col = r - 2 * kF3;
            fj = 1 + (col >> 1);
            fk = (col & 1) ? kF3 - 1 : 0;
        }
    }
}

// Memory index of frame cell (fi, fj, fk) of the block at (i0, j0, k0), clamped into the box.
__device__ __forceinline__ unsigned frame_index(int i0, int j0, int k0, int fi, int fj, int fk, int n) {
    const unsigned ci = (unsigned)clampi(i0 + fi - 1, 0, n - 1), cj = (unsigned)clampi(j0 + fj - 1, 0, n - 1),
                   ck = (unsigned)clampi(k0 + fk - 1, 0, n - 1);
    return (ci * (unsigned)n + cj) * (unsigned)n + ck;
}

// Scalars of iteration k from the reduced rows r.r, g, d (and b.b, row 4, at k = 0), with the breakdown
// stop of the two-pass form: r.z of the previous iteration or p.Ap not positive, beta or alpha not
// finite, or a non-finite r.r.
struct CgScalars {
    float rr, g, alpha, beta, pap;
    bool conv, brk, stop;
};
__device__ __forceinline__ CgScalars cg_scalars(const double* tot, bool first, float g_prev, float pap_prev,
                                                double tol, double& atol, float& atol_sq, int kit, int maxeff) {
    CgScalars q;
    q.rr = (float)tot[0];
    q.g = (float)tot[1];
    const float dd = (float)tot[2];
    if (first) {
        const float bb = (float)tot[4];
        atol = tol * sqrt((double)bb);   // warp: atol = max(tol * ||b||, tol)
        if (atol < tol) atol = tol;
        atol_sq = (float)(atol * atol);
    }
    q.conv = q.rr <= atol_sq;
    q.beta = 0.0f;
    q.pap = dd;
    if (!first && q.rr > atol_sq) {
        q.beta = q.g / g_prev;
        q.pap = dd - (q.beta * q.beta) * pap_prev;
    }
    q.alpha = q.rr > atol_sq ? q.g / q.pap : 0.0f;
    q.brk = !(q.rr <= 3.4028235e38f) ||
            (q.rr > atol_sq && ((!first && !(g_prev > 0.0f)) || !(fabsf(q.beta) <= 3.4028235e38f) ||
                                !(q.pap > 0.0f) || !(fabsf(q.alpha) <= 3.4028235e38f)));
    q.stop = q.conv || kit >= maxeff || q.brk;
    return q;
}

// Workgroup 0's record of iteration k's scalars, and of the solve's end.
__device__ __forceinline__ void cg_record(CgCtl* c, CgMirror* mirror, const CgScalars& q, bool first, int kit,
                                          double atol, float atol_sq) {
    c->slot_rr[kit & 1] = q.rr;
    c->slot_rz[kit & 1] = q.g;
    c->slot_pap[kit & 1] = (double)q.pap;
    if (first) {
        c->atol = atol;
        c->atol_sq = atol_sq;
    }
    if (q.stop) {
        const int cv = q.conv ? 1 : (q.brk ? -1 : 0);
        c->done = 1;
        c->converged = cv;
        c->it = kit;
        mirror_final(mirror, kit, q.rr, atol, cv);
    }
}

// s_k, r_{k+1} and (returned) u_{k+1} of a frame cell
__device__ __forceinline__ float cg_frame(float r, float w, float s_old, float d, bool first, float alpha, float beta,
                                          float& s, float& r1) {
    s = first ? w : w + beta * s_old;
    r1 = r - alpha * s;
    return d * r1;
}

// w = A u at block cell (li, lj, lk), from u (returned in uc) and alpha_s on the frame in LDS
template <int D>
__device__ __forceinline__ float cg_stencil(const float (*U)[kF3][kF3], const float (*S)[kF3][kF3], int li, int lj,
                                            int lk, float amin, float hh, float rhh, int fastdiv, float& uc) {
    const int si = li + 1, sj = lj + 1, sk = lk + 1;
    uc = U[si][sj][sk];
    const float ac = S[si][sj][sk];
    float acc6 = 0.0f;
    acc6 += face_fraction(S[si - 1][sj][sk], ac, amin) * (uc - U[si - 1][sj][sk]);
    acc6 += face_fraction(ac, S[si + 1][sj][sk], amin) * (uc - U[si + 1][sj][sk]);
    acc6 += face_fraction(S[si][sj - 1][sk], ac, amin) * (uc - U[si][sj - 1][sk]);
    acc6 += face_fraction(ac, S[si][sj + 1][sk], amin) * (uc - U[si][sj + 1][sk]);
    acc6 += face_fraction(S[si][sj][sk - 1], ac, amin) * (uc - U[si][sj][sk - 1]);
    acc6 += face_fraction(ac, S[si][sj][sk + 1], amin) * (uc - U[si][sj][sk + 1]);
    return div_const(1.0f * acc6, hh, rhh, fastdiv);
}

// cg_stencil with the block's face fractions formed ahead (FX below plane li, FY below row lj, FZ below
// column lk): the same six terms in the same order, so the same bits
template <int D>
__device__ __forceinline__ float cg_stencil_ff(const float (*U)[kF3][kF3], const float (*FX)[kT3][kT3],
                                               const float (*FY)[kT3 + 1][kT3], const float (*FZ)[kT3][kT3 + 1],
                                               int li, int lj, int lk, float hh, float rhh, int fastdiv, float& uc) {
    const int si = li + 1, sj = lj + 1, sk = lk + 1;
    uc = U[si][sj][sk];
    float acc6 = 0.0f;
    acc6 += FX[li][lj][lk] * (uc - U[si - 1][sj][sk]);
    acc6 += FX[li + 1][lj][lk] * (uc - U[si + 1][sj][sk]);
    acc6 += FY[li][lj][lk] * (uc - U[si][sj - 1][sk]);
    acc6 += FY[li][lj + 1][lk] * (uc - U[si][sj + 1][sk]);
    acc6 += FZ[li][lj][lk] * (uc - U[si][sj][sk - 1]);
    acc6 += FZ[li][lj][lk + 1] * (uc - U[si][sj][sk + 1]);
    return div_const(1.0f * acc6, hh, rhh, fastdiv);
}

// The persistent form's frame of u in LDS, laid out so that the stencil's reads of a wave (4 rows of 16
// k-consecutive cells) hit 64 distinct banks: rows of the 16 inner k positions, [D + 2][18][16], then the
// two k-halo positions of each row, [D + 2][18][2] (the same 18 x 18 x (D + 2) words as the plain frame).
template <int D>
__device__ __forceinline__ unsigned urow_offset(int fi, int fj, int fk) {
    return fk >= 1 && fk <= kT3 ? (unsigned)((fi * kF3 + fj) * kT3 + fk - 1)
                                : (unsigned)((D + 2) * kF3 * kT3 + (fi * kF3 + fj) * 2 + (fk == 0 ? 0 : 1));
}

// The six-term sum of cg_stencil_ff on the row layout (w before its division by h^2): u at k - 1 and
// k + 1 comes from the neighbouring lanes of the 16-lane row (DPP row_shr:1 / row_shl:1), the row's
// first and last lanes take the k-halo pair; the same six terms in the same order, so the same bits.
// Every lane of the wave must run it (the moves read the neighbours' registers).
template <int D>
__device__ __forceinline__ float cg_stencil_rows_sum(const float (*UI)[kF3][kT3], const float2 (*UK)[kF3],
                                                     const float (*FX)[kT3][kT3], const float (*FY)[kT3 + 1][kT3],
                                                     const float (*FZ)[kT3][kT3 + 1], int li, int lj, int lk,
                                                     float& uc) {
    const int si = li + 1, sj = lj + 1;
    uc = UI[si][sj][lk];
    const float2 kh = UK[si][sj];
    const float ukm = __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(kh.x), __float_as_int(uc), 0x111, 0xf, 0xf, false));  // row_shr:1
    const float ukp = __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(kh.y), __float_as_int(uc), 0x101, 0xf, 0xf, false));  // row_shl:1
    float acc6 = 0.0f;
    acc6 += FX[li][lj][lk] * (uc - UI[si - 1][sj][lk]);
    acc6 += FX[li + 1][lj][lk] * (uc - UI[si + 1][sj][lk]);
    acc6 += FY[li][lj][lk] * (uc - UI[si][sj - 1][lk]);
    acc6 += FY[li][lj + 1][lk] * (uc - UI[si][sj + 1][lk]);
    acc6 += FZ[li][lj][lk] * (uc - ukm);
    acc6 += FZ[li][lj][lk + 1] * (uc - ukp);
    return acc6;
}

// cg_stencil_rows_sum for a cell whose six face fractions are exactly 1 (no solid in or beside it): the
// same terms without reading the fractions (1 x d is d exactly), so the same bits.
template <int D>
__device__ __forceinline__ float cg_stencil_rows_open(const float (*UI)[kF3][kT3], const float2 (*UK)[kF3], int li,
                                                      int lj, int lk, float& uc) {
    const int si = li + 1, sj = lj + 1;
    uc = UI[si][sj][lk];
    const float2 kh = UK[si][sj];
    const float ukm = __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(kh.x), __float_as_int(uc), 0x111, 0xf, 0xf, false));  // row_shr:1
    const float ukp = __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(kh.y), __float_as_int(uc), 0x101, 0xf, 0xf, false));  // row_shl:1
    float acc6 = 0.0f;
    acc6 += uc - UI[si - 1][sj][lk];
    acc6 += uc - UI[si + 1][sj][lk];
    acc6 += uc - UI[si][sj - 1][lk];
    acc6 += uc - UI[si][sj + 1][lk];
    acc6 += uc - ukm;
    acc6 += uc - ukp;
    return acc6;
}

template <int D, class Src>
__global__ __launch_bounds__(kBlock3, D <= 8 ? 8 : 4) void k_pcg_fused(Src src, int kit, unsigned long long* pace,
                                                                       unsigned tag) {
    post_pace(pace, tag, kit);
    HFM_ENTRY(src);
    if (batch_solve_done(src, a.ctl + 1)) return;
    __shared__ float U[D + 2][kF3][kF3];   // u_{k+1}
    __shared__ float S[D + 2][kF3][kF3];   // alpha_s
    __shared__ double lds[3][kBlock3 / 64];
    CgCtl* c = a.ctl + 1;
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
    const float* __restrict__ wk = a.cg_Ap + (long)(kit & 1) * N;
    float* __restrict__ wn = a.cg_Ap + (long)((kit + 1) & 1) * N;
    const float* __restrict__ sprev = a.cg_d + (long)((kit + 1) & 1) * N;
    float* __restrict__ sk = a.cg_d + (long)(kit & 1) * N;
    float* __restrict__ pv = a.cg_z;
    float* __restrict__ x = a.p;
    const float* __restrict__ Dg = a.pdiag;
    const float* __restrict__ as = a.alpha_s;
    const double* part_in = a.partials + (long)(kit & 1) * kFusedRows * kMaxDirBlocks;
    double* part_out = a.partials + (long)((kit + 1) & 1) * kFusedRows * kMaxDirBlocks;
    // ---- the prologue's operands first (vector loads complete in issue order, so the prologue then
    // waits for these only): the partial sums of launch k - 1 (the init launch wrote rows of its own
    // count: r.r, g, d, -, b.b; later launches r.r, g, d) and the control scalars
    double acc[kFusedRows];
    load_rows<kFusedRows>(part_in, first ? a.nb_upd : nb, first ? kFusedRows : 3, acc);
    const int done = c->done, maxeff = c->maxeff;
    const float g_prev = c->slot_rz[(kit + 1) & 1];
    const float pap_prev = (float)c->slot_pap[(kit + 1) & 1];
    const double tol = c->tol;
    double atol = c->atol;
    float atol_sq = c->atol_sq;
    // ---- then the frame, in flight during the prologue: halo cells, then the block's own cells
    constexpr int kH = kHalo<D>, kHM = (kH + kBlock3 - 1) / kBlock3, kO = D / 4;
    // (threads past the last halo cell load the last one's values and do not use them: no branch, so
    // every load issues at once)
    float hr[kHM], hw[kHM], hs[kHM], hd[kHM], ha[kHM];
#pragma unroll
    for (int q = 0; q < kHM; ++q) {
        const int h = (int)threadIdx.x + q * kBlock3;
        int fi, fj, fk;
        halo_cell<D>(h < kH ? h : kH - 1, fi, fj, fk);
        const unsigned t = frame_index(i0, j0, k0, fi, fj, fk, n);
        hr[q] = ld32(rk, t);
        hw[q] = ld32(wk, t);
        hs[q] = first ? 0.0f : ld32(sprev, t);
        hd[q] = ld32(Dg, t);
        ha[q] = ld32(as, t);
    }
    const int lk = threadIdx.x % kT3, lj = (threadIdx.x / kT3) % kT3, lq = threadIdx.x / (kT3 * kT3);
    const int j = j0 + lj, k = k0 + lk;
    float orr[kO], ow[kO], os[kO], od[kO], oa[kO];
#pragma unroll
    for (int m = 0; m < kO; ++m) {
        const unsigned t = frame_index(i0, j0, k0, lq + 4 * m + 1, lj + 1, lk + 1, n);
        orr[m] = ld32(rk, t);
        ow[m] = ld32(wk, t);
        os[m] = first ? 0.0f : ld32(sprev, t);
        od[m] = ld32(Dg, t);
        oa[m] = ld32(as, t);
    }
    // ---- prologue: scalars of iteration k
    double tot[kFusedRows];
    if (reduce_bcast<kFusedRows, kBlock3>(acc, first ? kFusedRows : 3, done, tot, first ? a.nb_upd : nb)) return;
    const CgScalars q = cg_scalars(tot, first, g_prev, pap_prev, tol, atol, atol_sq, kit, maxeff);
    if (lb == 0 && threadIdx.x == 0) cg_record(c, a.mirror + 1, q, first, kit, atol, atol_sq);
    if (q.stop) return;
    // x and p of the own cells, needed only after the frame pass: loaded now (fewer registers live
    // across the prologue)
    float ox[kO], op[kO];
#pragma unroll
    for (int m = 0; m < kO; ++m) {
        const int i = i0 + lq + 4 * m;
        const bool in_box = i < n && j < n && k < n;
        const unsigned t = frame_index(i0, j0, k0, lq + 4 * m + 1, lj + 1, lk + 1, n);
        ox[m] = in_box ? ld32((const float*)x, t) : 0.0f;
        op[m] = in_box && !first ? ld32((const float*)pv, t) : 0.0f;
    }
    // ---- halo cells: u_{k+1}
#pragma unroll
    for (int qh = 0; qh < kHM; ++qh) {
        const int h = (int)threadIdx.x + qh * kBlock3;
        if (h < kH) {
            int fi, fj, fk;
            halo_cell<D>(h, fi, fj, fk);
            float s, r1;
            U[fi][fj][fk] = cg_frame(hr[qh], hw[qh], hs[qh], hd[qh], first, q.alpha, q.beta, s, r1);
            S[fi][fj][fk] = ha[qh];
        }
    }
    // ---- own cells: s_k, r_{k+1}, u_{k+1}, p_k, x_{k+1}
    double sum_g = 0.0, sum_rr = 0.0, sum_d = 0.0;   // r.u, r.r, w.u of step k + 1
#pragma unroll
    for (int m = 0; m < kO; ++m) {
        const int li = lq + 4 * m, i = i0 + li;
        float s, r1;
        const float u1 = cg_frame(orr[m], ow[m], os[m], od[m], first, q.alpha, q.beta, s, r1);
        U[li + 1][lj + 1][lk + 1] = u1;
        S[li + 1][lj + 1][lk + 1] = oa[m];
        if (i < n && j < n && k < n) {
            const unsigned t = ((unsigned)i * n + j) * n + k;
            const float uk = od[m] * orr[m];
            const float pk = first ? uk : uk + q.beta * op[m];
            st32(rn, t, r1);
            st32(sk, t, s);
            st32(pv, t, pk);
            st32(x, t, ox[m] + q.alpha * pk);
            sum_g += (double)(r1 * u1);
            sum_rr += (double)(r1 * r1);
        }
    }
    __syncthreads();
    // ---- w_{k+1} = A u_{k+1} on the block, from LDS
    const float amin = a.amin, hh = a.hh, rhh = a.rhh;
    const int fastdiv = a.fastdiv;
#pragma unroll
    for (int m = 0; m < kO; ++m) {
        const int li = lq + 4 * m, i = i0 + li;
        if (i < n && j < n && k < n) {
            float uc;
            const float w1 = cg_stencil<D>(U, S, li, lj, lk, amin, hh, rhh, fastdiv, uc);
            st32(wn, ((unsigned)i * n + j) * n + k, w1);
            sum_d += (double)(w1 * uc);
        }
    }
    // rows r.r, g, d in the order the prologue reads them
    double out[3] = {sum_rr, sum_g, sum_d};
    store_partials<3, kBlock3>(out, part_out, nb, lb, lds);
}

// The first iteration's inputs from r_0 (written by k_pcg_init<1>): w_0 = A D r_0 into w's buffer 0,
// and the partial sums r.r, g = r.u, d = w.u and b.b into the partial-sum buffer launch 0 reads.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_pcg_init_fused(Src src) {
    HFM_ENTRY(src);
    if (!a.cg_fused) return;
    __shared__ double lds[kFusedRows][kWaves];
    const PcgView v = make_view<1>(a, lb, 0, 0.0f);
    const float* r = a.cg_r;
    const float* Dg = a.pdiag;
    float* w = a.cg_Ap;
    double s[kFusedRows] = {0.0, 0.0, 0.0, 0.0, 0.0};
    FOR_CELLS(v, v.nb_upd) {
        int i, j, k;
        cell_ijk(t, v.n, i, j, k);
        auto uf = [r, Dg](long q) { return Dg[q] * r[q]; };
        const float acc = Operator<1>::apply(v, t, i, j, k, uf);
        const float wc = Operator<1>::combine(v, acc, 0.0f, false, 1.0f);
        const float rc = r[t], uc = uf(t), bc = v.b[t];
        w[t] = wc;
        s[0] += (double)(rc * rc);
        s[1] += (double)(rc * uc);
        s[2] += (double)(wc * uc);
        s[4] += (double)(bc * bc);
    }
    store_partials<kFusedRows>(s, a.partials, v.nb_upd, lb, lds);
    // the persistent form's partial-sum slots start empty (every copy)
    if (a.cg_persist) {
        double* slots = a.partials + kPersistSlots;
        for (int q = lb * kBlock + (int)threadIdx.x; q < 9 * a.nb_dir; q += v.nb_upd * kBlock)
#pragma unroll
            for (int c = 0; c < kPersistReplicas; ++c)
                slots[c * kPersistReplicaStride + q] = __builtin_bit_cast(double, kEmptySum);
    }
}

// ---------------------------------------------------------------------------------------------
// Persistent pressure PCG (HFM_FLAG_PERSISTENT_CG): every iteration of k_pcg_fused in one
// cooperative launch of one workgroup per tile, all of them resident for the whole solve.  A
// workgroup keeps its own cells' r, w, s and D in registers and their p and x in LDS (the stencil's
// layout, D / 4 cells per thread), D of its halo cells in registers and the frame's solid fraction in LDS, so memory carries
// only what a neighbour needs: the r, s and w of the cells on the tile's faces, and the three partial
// sums.  The launch boundary that separates two iterations of the one-launch form becomes the partial
// sums themselves: iteration k stores its faces, waits for those stores to complete, then stores its
// three sums in slot k % 3 of every copy of the slots (kPersistReplicas, one per XCD, each on pages of
// its own); iteration k + 1 starts when no sum of that slot in the copy of the workgroup's XCD reads as
// empty (a signalling NaN pattern, which no arithmetic produces), so the polls of the whole device
// spread over eight sets of pages.  A workgroup runs at most one iteration ahead of any
// other (it cannot pass an iteration before every workgroup has posted its sums), so the faces keep the
// one-launch form's two parity buffers, and a workgroup empties its entries of slot (k + 1) % 3 (in
// every copy) as iteration k starts: every reader of what they held (the sums of iteration k - 2) has passed
// iteration k - 1, and the next sums they take are those of iteration k + 1, posted after the emptying
// (a copy's entries are emptied and posted by the same thread, whose stores complete in order of its
// waits).  Every cross-workgroup access is a relaxed
// agent-scope atomic, coherent across the XCDs without L2 write-back or invalidation.
//
// Each workgroup computes what its launch of k_pcg_fused computes, from the same values, in the same
// order: the same tiles and layouts, the same per-workgroup partial sums and the same reduction of
// them, so the iterates, the scalars and the iteration count are the one-launch form's bit for bit.
// Every workgroup reaches the stop decision from the same sums at the same iteration; the final x and
// p go to memory where the one-launch form leaves them.  A wait is bounded: a workgroup whose sums do
// not arrive within `spin_limit` polls reports the solve as stopped with converged = -2 (the host turns
// that into an error) and ends, and so, in turn, does every workgroup waiting on it.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float load_agent_f(const float* p, unsigned i) {
    return __hip_atomic_load(const_cast<float*>(p) + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent_f(float* p, unsigned i, float v) {
    __hip_atomic_store(p + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool empty_sum(double v) {
    return __builtin_bit_cast(unsigned long long, v) == kEmptySum;
}

// The workgroup index of tile L (the inverse of xcd_tile).
__device__ __forceinline__ int tile_workgroup(int L, int nb) {
    const int per = nb / kXcds, rem = nb % kXcds, big = rem * (per + 1);
    const int x = L < big ? L / (per + 1) : rem + (L - big) / per;
    const int q = L < big ? L % (per + 1) : (L - big) % per;
    return q * kXcds + x;
}

// A workgroup's late flag (LDS): a thread whose bounded wait ran out sets it, every thread reads it after
// the barrier that follows.  It is never cleared: the workgroup ends once it reads the flag set, so one
// plain store and one barrier do what a workgroup OR (three barriers) would.
__device__ __forceinline__ bool late_any(int late, int* late_s) {
    if (late) *late_s = 1;
    __syncthreads();
    return *(volatile int*)late_s != 0;
}

// Wait until the r.r entry of workgroup `nbr` in slot P reads as posted (threads with nbr < 0 do not
// wait); false in every thread if it did not within `limit` polls (at once for a negative limit), or if
// the late flag was already set.  A workgroup posts its sums only after its faces of that iteration
// are in memory.
__device__ __forceinline__ bool wait_neighbour(const double* P, int nbr, int limit, int* late_s) {
    int late = 0;
    if (nbr >= 0 && limit < 0) {
        late = 1;
    } else if (nbr >= 0) {
        for (int spins = 0;; ++spins) {
            if (!empty_sum(load_agent(P + nbr))) break;
            if (spins >= limit) {
                late = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return !late_any(late, late_s);
}

// Rows r.r, g, d of slot P (row stride nb): thread b < nb polls its entries until none reads as empty and
// takes them as its share of the reduction (load_rows' first entries); returns 1 if they did not arrive
// within `limit` polls, at once for a negative limit (no barrier: the caller's next one combines the flags).
__device__ __forceinline__ int poll_sums(const double* P, int nb, int limit, double (&acc)[kFusedRows]) {
    int late = 0;
#pragma unroll
    for (int q = 0; q < kFusedRows; ++q) acc[q] = 0.0;
    const int b = (int)threadIdx.x;
    if (b < nb && limit < 0) {
        late = 1;
    } else if (b < nb) {
        for (int spins = 0;; ++spins) {
            const double v0 = load_agent(P + b), v1 = load_agent(P + nb + b), v2 = load_agent(P + 2 * nb + b);
            if (!empty_sum(v0) && !empty_sum(v1) && !empty_sum(v2)) {
                acc[0] = v0;
                acc[1] = v1;
                acc[2] = v2;
                break;
            }
            if (spins >= limit) {
                late = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return late;
}

template <int D, class Src>
__global__ __launch_bounds__(kBlock3) void k_pcg_persist(Src src, int spin_limit) {
    HFM_ENTRY(src);
    constexpr int kH = kHalo<D>, kHM = (kH + kBlock3 - 1) / kBlock3, kO = D / 4;
    // u_{k+1} on the frame in the row layout (urow_offset; alpha_s in the plain layout while the face
    // fractions are formed)
    __shared__ __attribute__((aligned(16))) float U[D + 2][kF3][kF3];
    __shared__ float FX[D + 1][kT3][kT3];  // face fractions of the block's cells: x-face below plane li
    __shared__ float FY[D][kT3 + 1][kT3];  // y-face below row lj
    __shared__ float FZ[D][kT3][kT3 + 1];  // z-face below column lk
    __shared__ float Pl[D][kT3][kT3];      // p_{k-1} of the own cells
    __shared__ float Xl[D][kT3][kT3];      // x_k of the own cells
    __shared__ uint2 Htab[kHM * kBlock3];  // halo cell h: memory index (clamped into the box), offset in U
    __shared__ double lds[3][kBlock3 / 64];
    __shared__ double wsum[kFusedRows][kBlock3 / 64];   // per-wave sums of the rows
    __shared__ int late_s;                              // the late flag (late_any)
    __shared__ CgScalars qs;                            // iteration k's scalars, formed by wave 0
    CgCtl* c = a.ctl + 1;
    const int n = a.n;
    const long N = a.N;
    const int nb = a.nb_dir;
    const int L = xcd_tile(lb, nb);
    const int tjk = a.tiles_j * a.tiles_k;
    const int ti = L / tjk, rem = L % tjk;
    const int tj = rem / a.tiles_k, tk = rem % a.tiles_k;
    const int i0 = ti * D, j0 = tj * kT3, k0 = tk * kT3;
    float* __restrict__ pv = a.cg_z;
    float* __restrict__ x = a.p;
    double* slots = a.partials + kPersistSlots;   // kPersistReplicas copies of [3][3 rows][nb]
    const double* rslots = slots + (lb % kXcds) * kPersistReplicaStride;   // the copy this workgroup polls
    // the workgroups of the tiles around this one (thread t < 27, t != 13: offset (t / 9, t / 3 % 3,
    // t % 3) - 1), whose faces the frame reads; -1 for none
    int nbr = -1;
    if (threadIdx.x < 27 && threadIdx.x != 13) {
        const int ni = ti + (int)threadIdx.x / 9 - 1, nj = tj + (int)threadIdx.x / 3 % 3 - 1,
                  nk = tk + (int)threadIdx.x % 3 - 1;
        if (ni >= 0 && ni < (n + D - 1) / D && nj >= 0 && nj < a.tiles_j && nk >= 0 && nk < a.tiles_k)
            nbr = tile_workgroup((ni * a.tiles_j + nj) * a.tiles_k + nk, nb);
    }
    // the layouts of k_pcg_fused: halo cells in the compact numbering, own cells in the stencil's layout
    const int lk = threadIdx.x % kT3, lj = (threadIdx.x / kT3) % kT3, lq = threadIdx.x / (kT3 * kT3);
    const int j = j0 + lj, k = k0 + lk;
    auto in_box = [&](int m) { return i0 + lq + 4 * m < n && j < n && k < n; };
    // an own cell is published (its r, s, w go to memory every iteration) when it lies on a face of the
    // block or of the box: the cells that some frame reads
    const bool edge_jk = lj == 0 || lj == kT3 - 1 || j == n - 1 || lk == 0 || lk == kT3 - 1 || k == n - 1;
    auto pub = [&](int m) {
        const int li = lq + 4 * m;
        return edge_jk || li == 0 || li == D - 1 || i0 + li == n - 1;
    };
    auto own_index = [&](int m) { return frame_index(i0, j0, k0, lq + 4 * m + 1, lj + 1, lk + 1, n); };
    // constants of the solve: D at the halo and own cells, the halo cells' indices (each thread reads
    // only the entries it writes) and, from alpha_s on the frame, the face fractions of the block's cells
    float (*S)[kF3][kF3] = U;
    float hd[kHM], od[kO];
#pragma unroll
    for (int q = 0; q < kHM; ++q) {
        const int h = (int)threadIdx.x + q * kBlock3;
        int fi, fj, fk;
        halo_cell<D>(h < kH ? h : kH - 1, fi, fj, fk);
        const unsigned t = frame_index(i0, j0, k0, fi, fj, fk, n);
        Htab[h] = make_uint2(t, urow_offset<D>(fi, fj, fk));
        hd[q] = ld32((const float*)a.pdiag, t);
        if (h < kH) S[fi][fj][fk] = ld32((const float*)a.alpha_s, t);
    }
    // own cells: r_k, w_k, s_{k-1} in registers, p_{k-1} and x_k in LDS, for the whole solve
    float orr[kO], ow[kO], os[kO];
#pragma unroll
    for (int m = 0; m < kO; ++m) {
        const unsigned t = own_index(m);
        od[m] = ld32((const float*)a.pdiag, t);
        S[lq + 4 * m + 1][lj + 1][lk + 1] = ld32((const float*)a.alpha_s, t);
        Xl[lq + 4 * m][lj][lk] = in_box(m) ? ld32((const float*)x, t) : 0.0f;   // x_0
        orr[m] = ow[m] = os[m] = 0.0f;
    }
    if (threadIdx.x == 0) late_s = 0;
    const float amin = a.amin, hh = a.hh, rhh = a.rhh;
    const int fastdiv = a.fastdiv;
    __syncthreads();
    for (int e = (int)threadIdx.x; e < (D + 1) * kT3 * kT3; e += kBlock3) {
        const int li = e / (kT3 * kT3), lj2 = e / kT3 % kT3, lk2 = e % kT3;
        FX[li][lj2][lk2] = face_fraction(S[li][lj2 + 1][lk2 + 1], S[li + 1][lj2 + 1][lk2 + 1], amin);
    }
    for (int e = (int)threadIdx.x; e < D * (kT3 + 1) * kT3; e += kBlock3) {
        const int li = e / ((kT3 + 1) * kT3), jj = e / kT3 % (kT3 + 1), lk2 = e % kT3;
        FY[li][jj][lk2] = face_fraction(S[li + 1][jj][lk2 + 1], S[li + 1][jj + 1][lk2 + 1], amin);
    }
    for (int e = (int)threadIdx.x; e < D * kT3 * (kT3 + 1); e += kBlock3) {
        const int li = e / (kT3 * (kT3 + 1)), lj2 = e / (kT3 + 1) % kT3, kk = e % (kT3 + 1);
        FZ[li][lj2][kk] = face_fraction(S[li + 1][lj2 + 1][kk], S[li + 1][lj2 + 1][kk + 1], amin);
    }
    __syncthreads();   // U is u from here on, in the row layout
    // bit m: cell m of every lane of this wave has six open faces (fraction exactly 1), so the wave forms
    // its stencil without reading the fractions (uniform over the wave)
    unsigned open_cells = 0;
#pragma unroll
    for (int m = 0; m < kO; ++m) {
        const int li = lq + 4 * m;
        const bool open = FX[li][lj][lk] == 1.0f && FX[li + 1][lj][lk] == 1.0f && FY[li][lj][lk] == 1.0f &&
                          FY[li][lj + 1][lk] == 1.0f && FZ[li][lj][lk] == 1.0f && FZ[li][lj][lk + 1] == 1.0f;
        if (__all(open)) open_cells |= 1u << m;
    }
    float (*UI)[kF3][kT3] = reinterpret_cast<float (*)[kF3][kT3]>(&U[0][0][0]);
    const float2 (*UK)[kF3] = reinterpret_cast<const float2 (*)[kF3]>(&U[0][0][0] + (D + 2) * kF3 * kT3);
    const int maxeff = c->maxeff;
    const double tol = c->tol;
    double atol = c->atol;
    float atol_sq = c->atol_sq;
    float g_prev = 0.0f, pap_prev = 0.0f;
    for (int kit = 0;; ++kit) {
        const bool first = kit == 0;
        const float* __restrict__ rk = a.cg_r + (long)(kit & 1) * N;
        float* __restrict__ rn = a.cg_r + (long)((kit + 1) & 1) * N;
        const float* __restrict__ wk = a.cg_Ap + (long)(kit & 1) * N;
        float* __restrict__ wn = a.cg_Ap + (long)((kit + 1) & 1) * N;
        const float* __restrict__ sprev = a.cg_d + (long)((kit + 1) & 1) * N;
        float* __restrict__ sk = a.cg_d + (long)(kit & 1) * N;
        // ---- the partial sums of iteration k - 1 (the init launch's at k = 0): first the neighbours' (their
        // faces are then in memory, and the halo loads go out while the others arrive), then everyone's
        double acc[kFusedRows];
        const double* P = rslots + ((kit + 2) % 3) * 3 * nb;   // slot (k - 1) % 3
        bool ok = true;
        if (first)
            load_rows<kFusedRows>(a.partials, a.nb_upd, kFusedRows, acc);
        else
            ok = wait_neighbour(P, nbr, spin_limit, &late_s);
        // ---- r_k, w_k, s_{k-1} of the halo cells, and of the own cells at k = 0 (and of positions of the
        // block outside the box: the clamped cell's), from memory
        // (threads past the last halo cell load the last one's values and do not use them)
        float hr[kHM], hw[kHM], hs[kHM];
#pragma unroll
        for (int q = 0; q < kHM; ++q) {
            const unsigned t = Htab[threadIdx.x + q * kBlock3].x;
            hr[q] = load_agent_f(rk, t);
            hw[q] = load_agent_f(wk, t);
            hs[q] = first ? 0.0f : load_agent_f(sprev, t);
        }
#pragma unroll
        for (int m = 0; m < kO; ++m) {
            if (first || !in_box(m)) {
                const unsigned t = own_index(m);
                orr[m] = load_agent_f(rk, t);
                ow[m] = load_agent_f(wk, t);
                os[m] = first ? 0.0f : load_agent_f(sprev, t);
            }
        }
        int late = ok ? 0 : 1;
        if (!first && ok) late = poll_sums(P, nb, spin_limit, acc);
        // ---- scalars of iteration k: the waves holding entries reduce them (reduce_bcast's tree), wave 0
        // adds the waves and forms the scalars once, the workgroup reads them from LDS
        {
            const int count = first ? a.nb_upd : nb;
            const bool any = (int)(threadIdx.x & ~63u) < count;   // uniform over the wave
            if (first) {
                double w[kFusedRows];
#pragma unroll
                for (int r = 0; r < kFusedRows; ++r) w[r] = acc[r];
                if (any) wave_sums<kFusedRows>(w);
#pragma unroll
                for (int r = 0; r < kFusedRows; ++r)
                    if ((threadIdx.x & 63) == 0) wsum[r][threadIdx.x >> 6] = any ? w[r] : 0.0;
            } else {
                double w[3] = {acc[0], acc[1], acc[2]};
                if (any) wave_sums<3>(w);
#pragma unroll
                for (int r = 0; r < 3; ++r)
                    if ((threadIdx.x & 63) == 0) wsum[r][threadIdx.x >> 6] = any ? w[r] : 0.0;
            }
        }
        if (late_any(late, &late_s)) {
            if (threadIdx.x == 0) {
                c->done = 1;
                mirror_final(a.mirror + 1, kit, __builtin_nanf(""), atol, -2);
            }
            return;
        }
        if (threadIdx.x < 64) {
            double tot[kFusedRows];
            if (first) {
                sum_wave_values_n<kBlock3 / 64, kFusedRows>(wsum, tot);
            } else {
                double t3[3];
                sum_wave_values_n<kBlock3 / 64, 3>(wsum, t3);
#pragma unroll
                for (int r = 0; r < kFusedRows; ++r) tot[r] = r < 3 ? t3[r] : 0.0;
            }
            const CgScalars s0 = cg_scalars(tot, first, g_prev, pap_prev, tol, atol, atol_sq, kit, maxeff);
            if (threadIdx.x == 0) {
                qs = s0;
                if (lb == 0) cg_record(c, a.mirror + 1, s0, first, kit, atol, atol_sq);
            }
            if (kit >= 2 && threadIdx.x < kPersistReplicas)   // empty this workgroup's sums of iteration
#pragma unroll                                                 // k - 2, thread c in copy c
                for (int r = 0; r < 3; ++r)
                    store_agent(slots + threadIdx.x * kPersistReplicaStride + ((kit + 1) % 3) * 3 * nb + r * nb + lb,
                                __builtin_bit_cast(double, kEmptySum));
        }
        __syncthreads();
        const CgScalars q = qs;
        if (q.stop) {
            // x_k and p_{k-1} where the one-launch form leaves them
#pragma unroll
            for (int m = 0; m < kO; ++m)
                if (in_box(m)) {
                    st32(x, own_index(m), Xl[lq + 4 * m][lj][lk]);
                    if (!first) st32(pv, own_index(m), Pl[lq + 4 * m][lj][lk]);
                }
            return;
        }
        // ---- halo cells: u_{k+1}
#pragma unroll
        for (int qh = 0; qh < kHM; ++qh) {
            const int h = (int)threadIdx.x + qh * kBlock3;
            if (h < kH) {
                float s, r1;
                (&U[0][0][0])[Htab[h].y] = cg_frame(hr[qh], hw[qh], hs[qh], hd[qh], first, q.alpha, q.beta, s, r1);
            }
        }
        // ---- own cells: s_k, r_{k+1}, u_{k+1} (p_k and x_{k+1}, which no later step of the iteration
        // reads, once the partial sums are posted)
        double sum_g = 0.0, sum_rr = 0.0, sum_d = 0.0;   // r.u, r.r, w.u of step k + 1
        float rk_own[kO];
#pragma unroll
        for (int m = 0; m < kO; ++m) {
            float s, r1;
            const float u1 = cg_frame(orr[m], ow[m], os[m], od[m], first, q.alpha, q.beta, s, r1);
            UI[lq + 4 * m + 1][lj + 1][lk] = u1;
            rk_own[m] = orr[m];
            if (in_box(m)) {
                orr[m] = r1;
                os[m] = s;
                if (pub(m)) {
                    store_agent_f(rn, own_index(m), r1);
                    store_agent_f(sk, own_index(m), s);
                }
                sum_g += (double)(r1 * u1);
                sum_rr += (double)(r1 * r1);
            }
        }
        __syncthreads();
        // ---- w_{k+1} = A u_{k+1} on the block, from LDS
#pragma unroll
        for (int m = 0; m < kO; ++m) {
            float uc;   // every lane forms its cell's w (the k neighbours come from the row's lanes)
            const float acc6 = (open_cells >> m) & 1u
                                   ? cg_stencil_rows_open<D>(UI, UK, lq + 4 * m, lj, lk, uc)
                                   : cg_stencil_rows_sum<D>(UI, UK, FX, FY, FZ, lq + 4 * m, lj, lk, uc);
            const float w1 = div_const(1.0f * acc6, hh, rhh, fastdiv);
            if (in_box(m)) {
                ow[m] = w1;
                if (pub(m)) store_agent_f(wn, own_index(m), w1);
                sum_d += (double)(w1 * uc);
            }
        }
        // ---- partial sums (rows r.r, g, d) in slot k % 3, once every store of this iteration is done
        wait_stores();
        double out[3] = {sum_rr, sum_g, sum_d};
        block_sum<3, kBlock3>(out, lds);
        if (threadIdx.x < kPersistReplicas)   // thread c posts to copy c
#pragma unroll
            for (int r = 0; r < 3; ++r)
                store_agent(slots + threadIdx.x * kPersistReplicaStride + (kit % 3) * 3 * nb + r * nb + lb, out[r]);
        // ---- p_k and x_{k+1} of the own cells, while the other workgroups' sums are on their way
#pragma unroll
        for (int m = 0; m < kO; ++m) {
            if (in_box(m)) {
                const float uk = od[m] * rk_own[m];
                float& pl = Pl[lq + 4 * m][lj][lk];
                float& xl = Xl[lq + 4 * m][lj][lk];
                const float pk = first ? uk : uk + q.beta * pl;
                pl = pk;
                xl = xl + q.alpha * pk;
            }
        }
        g_prev = q.g;
        pap_prev = q.pap;
    }
}

}  // namespace

// PCG geometry of a scene.  Init, pass B and the flat pass A cover cpt consecutive tiles of
// kBlock cells per workgroup: at least the scene's pcg_cpt (HFM_PCG_CPT), and enough that at most
// kMaxPcgBlocks workgroups run.  The pressure pass A is LDS-tiled (unless HFM_PCG_FLAT_DIR) with
// planes D = 16, else 8, else 4, the first that keeps at most kMaxDirBlocks workgroups (deeper blocks
// re-read less halo and launch fewer, longer workgroups, which measured fastest); the one-launch form
// tries 8 first (two of its workgroups then fit a CU), its persistent form 16 (fewest workgroups to
// keep resident).  HFM_DIR_ICHUNK forces D.  Without a tiling pass A is the flat kernel with the pass B
// geometry.
void pcg_geometry(const Scene* s, bool fused, bool persist, int* cpt, int* nb_upd, int* nb_dir, int* tiles_j,
                  int* tiles_k, int* dtile) {
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
        // the one-launch form prefers 8 planes: two of its workgroups fit a CU (at most 64 VGPRs)
        const int cand[3] = {fused && !persist ? 8 : 16, fused && !persist ? 16 : 8, 4};
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
        HFM_LAUNCH_T(k_pcg_init, 1, L, kBlock, nullptr, nullptr, dt);
    else
        HFM_LAUNCH_T(k_pcg_init, 0, L, kBlock, nullptr, nullptr, dt);
}

template <bool IR>
void launch_dir(const Launch& L, Family fam, int op, float dt, int k, unsigned long long* pace, unsigned tag,
                hipEvent_t ev0, hipEvent_t ev1) {
    switch (fam) {
        case FAM_DIR4: HFM_LAUNCH_T2(k_pcg_dir_tile3, 4, IR, L, kBlock3, ev0, ev1, k, pace, tag); break;
        case FAM_DIR8: HFM_LAUNCH_T2(k_pcg_dir_tile3, 8, IR, L, kBlock3, ev0, ev1, k, pace, tag); break;
        case FAM_DIR16: HFM_LAUNCH_T2(k_pcg_dir_tile3, 16, IR, L, kBlock3, ev0, ev1, k, pace, tag); break;
        default:
            if (op == 1)
                HFM_LAUNCH_T2(k_pcg_dir, 1, IR, L, kBlock, ev0, ev1, dt, k, pace, tag);
            else
                HFM_LAUNCH_T2(k_pcg_dir, 0, IR, L, kBlock, ev0, ev1, dt, k, pace, tag);
    }
}

void launch_pcg_dir(const Launch& L, Family fam, int op, float dt, int k, unsigned long long* pace, unsigned tag,
                    hipEvent_t ev0, hipEvent_t ev1, bool ir) {
    if (ir)
        launch_dir<true>(L, fam, op, dt, k, pace, tag, ev0, ev1);
    else
        launch_dir<false>(L, fam, op, dt, k, pace, tag, ev0, ev1);
}

void launch_pcg_init_fused(const Launch& L) { HFM_LAUNCH(k_pcg_init_fused, L, kBlock, nullptr, nullptr); }

void launch_pcg_fused(const Launch& L, Family fam, int k, unsigned long long* pace, unsigned tag, hipEvent_t ev0,
                      hipEvent_t ev1) {
    switch (fam) {
        case FAM_FUSED4: HFM_LAUNCH_T(k_pcg_fused, 4, L, kBlock3, ev0, ev1, k, pace, tag); break;
        case FAM_FUSED8: HFM_LAUNCH_T(k_pcg_fused, 8, L, kBlock3, ev0, ev1, k, pace, tag); break;
        case FAM_FUSED16: HFM_LAUNCH_T(k_pcg_fused, 16, L, kBlock3, ev0, ev1, k, pace, tag); break;
        default: break;
    }
}

template <class Src>
static const void* persist_kernel(int D) {
    switch (D) {
        case 4: return (const void*)k_pcg_persist<4, Src>;
        case 8: return (const void*)k_pcg_persist<8, Src>;
        case 16: return (const void*)k_pcg_persist<16, Src>;
        default: return nullptr;
    }
}

int pcg_persist_capacity(int D, int device, bool batch) {
    const void* fn = batch ? persist_kernel<BatchSrc>(D) : persist_kernel<SoloSrc>(D);
    if (!fn) return 0;
    int per_cu = 0, cus = 0, coop = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock3, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess || !coop) return 0;
    return per_cu * cus;
}

// HFM_PCG_PERSIST_ORDINARY=1 launches the same kernel as an ordinary dispatch, for profilers that cannot
// trace a cooperative launch (rocprofv3 7.2 ends without writing its results).  Co-residency is then not
// guaranteed by the runtime, only likely on an otherwise idle device (one workgroup per CU); the waits
// stay bounded either way.  HFM_PCG_PERSIST_SPINS (read at every launch) replaces the polls a wait may
// take; a negative value gives up at the first wait, which forces the timeout path from the host.
static bool persist_ordinary() {
    static const bool ordinary = [] {
        const char* e = getenv("HFM_PCG_PERSIST_ORDINARY");
        return e && atoi(e) != 0;
    }();
    return ordinary;
}

hipError_t launch_pcg_persist(const Launch& L, int D, hipEvent_t ev0, hipEvent_t ev1) {
    const bool ordinary = persist_ordinary();
    const void* fn = persist_kernel<SoloSrc>(D);
    if (!fn || !L.solo) return hipErrorInvalidValue;
    SoloSrc src{*reinterpret_cast<const SceneG*>(L.solo)};
    int limit = kPersistSpinLimit;
    if (const char* e = getenv("HFM_PCG_PERSIST_SPINS")) limit = atoi(e);
    void* args[] = {(void*)&src, (void*)&limit};
    if (ev0) (void)hipEventRecord(ev0, L.stream);
    const hipError_t e = ordinary ? hipLaunchKernel(fn, L.sgrid, dim3(kBlock3), args, 0, L.stream)
                                  : hipLaunchCooperativeKernel(fn, L.sgrid, dim3(kBlock3), args, 0, L.stream);
    if (ev1) (void)hipEventRecord(ev1, L.stream);
    return e;
}

hipError_t launch_pcg_persist_batch(const SceneDev* tab, const MapEntry* map, int blocks, int D, hipStream_t stream,
                                    hipEvent_t ev0, hipEvent_t ev1) {
    const void* fn = persist_kernel<BatchSrc>(D);
    if (!fn || blocks <= 0) return hipErrorInvalidValue;
    BatchSrc src;
    src.tab = (const HFM_CONST SceneG*)(const void*)tab;
    src.map = (const HFM_CONST MapEntry*)(const void*)map;
    int limit = kPersistSpinLimit;
    if (const char* e = getenv("HFM_PCG_PERSIST_SPINS")) limit = atoi(e);
    void* args[] = {(void*)&src, (void*)&limit};
    if (ev0) (void)hipEventRecord(ev0, stream);
    const hipError_t e = persist_ordinary() ? hipLaunchKernel(fn, dim3(blocks), dim3(kBlock3), args, 0, stream)
                                            : hipLaunchCooperativeKernel(fn, dim3(blocks), dim3(kBlock3), args, 0, stream);
    if (ev1) (void)hipEventRecord(ev1, stream);
    return e;
}

void launch_pcg_upd(const Launch& L, int op, int k, hipEvent_t ev0, hipEvent_t ev1, bool ir) {
    if (op == 1 && ir)
        HFM_LAUNCH_T2(k_pcg_upd, 1, true, L, kBlock, ev0, ev1, k);
    else if (op == 1)
        HFM_LAUNCH_T2(k_pcg_upd, 1, false, L, kBlock, ev0, ev1, k);
    else if (ir)
        HFM_LAUNCH_T2(k_pcg_upd, 0, true, L, kBlock, ev0, ev1, k);
    else
        HFM_LAUNCH_T2(k_pcg_upd, 0, false, L, kBlock, ev0, ev1, k);
}

void launch_ir_residual(const Launch& L, int op, float dt, bool first) {
    if (op == 1)
        HFM_LAUNCH_T(k_ir_residual, 1, L, kBlock, nullptr, nullptr, dt, (int)first);
    else
        HFM_LAUNCH_T(k_ir_residual, 0, L, kBlock, nullptr, nullptr, dt, (int)first);
}

void launch_ir_scale(const Launch& L, int op, int step) {
    if (op == 1)
        HFM_LAUNCH_T(k_ir_scale, 1, L, kBlock, nullptr, nullptr, step);
    else
        HFM_LAUNCH_T(k_ir_scale, 0, L, kBlock, nullptr, nullptr, step);
}

}  // namespace hfm
