// hfm_mg.hip — geometric multigrid V-cycle as the preconditioner of the pressure PCG
// (HFM_FLAG_MG_PCG; no counterpart in the reference, whose PCG is Jacobi-preconditioned).
//
// The PCG runs unchanged (warp's iteration and stopping rule, hfm_pcg.hip); after its init and
// after every pass B, the V-cycle below replaces z = D r by z = M r and the partial sums of r.z
// that the next pass A reads.  Levels: level 0 is the scene's grid, each next level halves the
// cells per side, down to at most 8.  The levels work with the operator A^ = h^2 A = sum over
// in-range faces a_f (e_c - e_nb) and right-hand side s = h^2 r, so a level is its face
// coefficients alone: level 0 takes the air fractions of the solid fraction (grid.py:46-59, the
// pressure operator's own), a coarse face the average of the 4 fine faces it covers.
//   down, level l -> l+1:  e = omega s / d (one weighted-Jacobi sweep from zero, formed where it is
//                          read), t = s - A^ e, s_{l+1} = r * (sum of t over the 8 children)
//   coarsest level:        16 weighted-Jacobi sweeps from zero
//   up, level l+1 -> l:    e = omega s / d + e_{l+1}(parent) (injection), then one sweep
//                          e += omega (s - A^ e) / d; at level 0 that is z, with r.z partials
// Pre- and post-smoothing are the same symmetric sweep and injection is 1 / r times the transpose
// of the restriction, so M is symmetric, as CG needs, and positive definite for omega < 1 (the
// eigenvalues of D^-1 A^ lie in [0, 2] for these diagonally dominant operators, so the sweep
// contracts and the coarsest sweeps' polynomial (1 - (1 - omega lambda)^k) / lambda stays positive).  d = the sum of a cell's face coefficients; at
// level 0, omega s / d = omega D r with the PCG's own Jacobi diagonal D; on the coarse levels
// w = omega_c / d is stored once per solve (omega_c: the coarse levels' own weight).
//
// Kernels.  The level-0 and large-level passes are memory-latency bound stencils: a neighbour
// outside the grid is replaced by the cell itself, whose term c (e_c - e_c) is exactly 0, so
// each stencil is branch-free and issues all its loads at once.  The restriction runs one thread
// per fine cell and sums the 8 children of a coarse cell across 8 lanes (a fixed DPP tree).  The
// levels of at most 12^3 cells (the "tail": one level, or one and its 5^3 / 6^3 coarsest) run in
// one workgroup per scene out of LDS: restriction, the coarsest sweeps and the prolongation back
// to the tail level in one launch.
#include <hip/hip_ext.h>

#include "hfm_device.h"
#include "hfm_scene.h"

namespace hfm {

namespace {

// omega (level 0), omega on the coarse levels and r, measured on MI355X (coupled scene, PCG
// iterations per substep at 96^3 / 192^3): 0.8, 0.8, 0.5: 12.8 / 20.8; 0.9, 0.9, 0.5: 10.9 / 17.6;
// 0.9, 0.9, 0.4: 10.3 / 16.1; 0.9, 0.95, 0.4: 9.8 / 15.6; 0.9, 0.98, 0.4: 9.6 / 15.2.  A single omega
// of 0.88, 0.93 or 0.95, r = 0.35, 0.45 or 0.55, coarse omega 0.85, and 8 or 12 coarsest sweeps took
// more; 24 sweeps, a level-0 omega of 0.92, or a separate r at level 0 (0.35-0.45) changed little.
constexpr float kMgOmega = 0.9f;
constexpr float kMgOmegaCoarse = 0.98f;
constexpr float kMgRestrict = 0.4f;
constexpr int kMgCoarseSweeps = 16;
constexpr int kTailBlock = 1024;
constexpr int kTailFine = kMgTailMaxN * kMgTailMaxN * kMgTailMaxN;   // cells of the tail level
constexpr int kTailCoarse = (kMgTailMaxN / 2) * (kMgTailMaxN / 2) * (kMgTailMaxN / 2);
static_assert(kTailCoarse <= kTailBlock && 512 <= kTailBlock, "a coarsest level is swept one cell per thread");

// cell coordinates of index t of an n^3 level (32-bit: n <= 1024)
__device__ __forceinline__ void cell3(unsigned t, unsigned n, int& i, int& j, int& k) {
    const unsigned nn = n * n;
    const unsigned a = t / nn, r = t - a * nn, b = r / n;
    i = (int)a;
    j = (int)b;
    k = (int)(r - b * n);
}

// A cell and its six neighbours in the order -x, +x, -y, +y, -z, +z, each clamped to the cell
// itself where it would leave the n^3 grid; with parents(), the parents of all seven on the level
// above (m = n / 2 cells per side).
struct Stencil {
    unsigned t;       // the cell's index
    unsigned q[6];    // neighbour indices
    unsigned pc;      // parent of the cell
    unsigned pq[6];   // parents of the neighbours
};

__device__ __forceinline__ Stencil stencil(int i, int j, int k, int n) {
    Stencil s;
    const unsigned nn = (unsigned)n * n;
    s.t = ((unsigned)i * n + j) * n + k;
    s.q[0] = i > 0 ? s.t - nn : s.t;
    s.q[1] = i + 1 < n ? s.t + nn : s.t;
    s.q[2] = j > 0 ? s.t - n : s.t;
    s.q[3] = j + 1 < n ? s.t + n : s.t;
    s.q[4] = k > 0 ? s.t - 1 : s.t;
    s.q[5] = k + 1 < n ? s.t + 1 : s.t;
    return s;
}

// (i - 1) >> 1 is the parent's row less one where i is even; (i + 1) >> 1 one more where i is odd
__device__ __forceinline__ void parents(Stencil& s, int i, int j, int k, int n, int m) {
    const unsigned mm = (unsigned)m * m;
    s.pc = ((unsigned)(i >> 1) * m + (j >> 1)) * m + (k >> 1);
    s.pq[0] = (i > 0 && !(i & 1)) ? s.pc - mm : s.pc;
    s.pq[1] = (i + 1 < n && (i & 1)) ? s.pc + mm : s.pc;
    s.pq[2] = (j > 0 && !(j & 1)) ? s.pc - m : s.pc;
    s.pq[3] = (j + 1 < n && (j & 1)) ? s.pc + m : s.pc;
    s.pq[4] = (k > 0 && !(k & 1)) ? s.pc - 1 : s.pc;
    s.pq[5] = (k + 1 < n && (k & 1)) ? s.pc + 1 : s.pc;
}

// sum over faces of c_f (e_c - e_nb), in the stencil's order
__device__ __forceinline__ float apply6(float ec, const float (&e)[6], const float (&c)[6]) {
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < 6; ++d) acc += c[d] * (ec - e[d]);
    return acc;
}

// Face coefficients of a stencil.  Level 0: air fractions of the two cells of each face, as the
// pressure operator forms them; a clamped face gives some finite value that meets e_c - e_c = 0.
template <class A>
__device__ __forceinline__ void coefs0(const A& a, const Stencil& s, float (&c)[6]) {
    const float* as = a.alpha_s;
    const float ac = as[s.t];
    float an[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) an[d] = as[s.q[d]];
#pragma unroll
    for (int d = 0; d < 6; ++d) c[d] = (d & 1) ? face_fraction(ac, an[d], a.amin) : face_fraction(an[d], ac, a.amin);
}

// Coarse level at `off`: the -x/-y/-z face of the cell and of its + neighbour (clamped: the
// cell's own -face, finite).
__device__ __forceinline__ void coefs_coarse(const float* ax, const float* ay, const float* az, long off,
                                             const Stencil& s, float (&c)[6]) {
    c[0] = ax[off + s.t];
    c[1] = ax[off + s.q[1]];
    c[2] = ay[off + s.t];
    c[3] = ay[off + s.q[3]];
    c[4] = az[off + s.t];
    c[5] = az[off + s.q[5]];
}

// Restriction lane layout: lane = 16 (2 di + dj) + 2 c + dk holds child (di, dj, dk) of the
// wave's coarse cell c (0..7), so that 16 consecutive lanes read one fine row.  The sum over the
// 8 children pairs k, then j, then i neighbours, the same tree in every lane (fp32 addition
// commutes): lane 16 (2 di + dj) + 2 c + dk ends with coarse cell c's sum for every (di, dj, dk).
__device__ __forceinline__ float sum_children(float x) {
    x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xf, 0xf, false));  // lane ^ 1
    x += __shfl_xor(x, 16);
    x += __shfl_xor(x, 32);
    return x;
}

// the same tree over 8 values held by one thread
__device__ __forceinline__ float sum8(const float (&v)[8]) {
    return ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}

// ---------------------------------------------------------------------------------------------
// Coefficients (once per solve): face coefficients and w = omega / d of level c = l + 1.
// ---------------------------------------------------------------------------------------------

// minus-face coefficient of level l at (i, j, k) along axis ax (0 at the grid's low wall)
template <class A>
__device__ __forceinline__ float face_minus(const A& a, int l, int ax, int i, int j, int k) {
    const int n = a.mg_n[l];
    const int x[3] = {i, j, k};
    if (x[ax] <= 0) return 0.0f;
    const long t = ((long)i * n + j) * n + k;
    if (l == 0) {
        const long st = ax == 0 ? (long)n * n : (ax == 1 ? n : 1);
        return face_fraction(a.alpha_s[t - st], a.alpha_s[t], a.amin);
    }
    const float* f = ax == 0 ? a.mg_ax : (ax == 1 ? a.mg_ay : a.mg_az);
    return f[a.mg_off[l] + t];
}

template <class Src>
__global__ __launch_bounds__(kBlock) void k_mg_coef(Src src, int l) {
    HFM_ENTRY(src);
    const int c = l + 1;
    if (!a.cg_mg || c >= a.mg_levels) return;
    const int m = a.mg_n[c];
    const unsigned Nc = (unsigned)m * m * m;
    const unsigned T = (unsigned)xcd_tile(lb, nblocks(Nc)) * kBlock + threadIdx.x;
    if (T >= Nc) return;
    int I, J, K;
    cell3(T, m, I, J, K);
    const int i = 2 * I, j = 2 * J, k = 2 * K;
    // a coarse -x face covers the fine -x faces of its 4 children at i = 2I; the +x face those
    // at i = 2I + 2 (absent at the high wall)
    const float ax = 0.25f * (face_minus(a, l, 0, i, j, k) + face_minus(a, l, 0, i, j + 1, k) +
                              face_minus(a, l, 0, i, j, k + 1) + face_minus(a, l, 0, i, j + 1, k + 1));
    const float ay = 0.25f * (face_minus(a, l, 1, i, j, k) + face_minus(a, l, 1, i + 1, j, k) +
                              face_minus(a, l, 1, i, j, k + 1) + face_minus(a, l, 1, i + 1, j, k + 1));
    const float az = 0.25f * (face_minus(a, l, 2, i, j, k) + face_minus(a, l, 2, i + 1, j, k) +
                              face_minus(a, l, 2, i, j + 1, k) + face_minus(a, l, 2, i + 1, j + 1, k));
    const float axp = I + 1 < m ? 0.25f * (face_minus(a, l, 0, i + 2, j, k) + face_minus(a, l, 0, i + 2, j + 1, k) +
                                           face_minus(a, l, 0, i + 2, j, k + 1) +
                                           face_minus(a, l, 0, i + 2, j + 1, k + 1))
                                : 0.0f;
    const float ayp = J + 1 < m ? 0.25f * (face_minus(a, l, 1, i, j + 2, k) + face_minus(a, l, 1, i + 1, j + 2, k) +
                                           face_minus(a, l, 1, i, j + 2, k + 1) +
                                           face_minus(a, l, 1, i + 1, j + 2, k + 1))
                                : 0.0f;
    const float azp = K + 1 < m ? 0.25f * (face_minus(a, l, 2, i, j, k + 2) + face_minus(a, l, 2, i + 1, j, k + 2) +
                                           face_minus(a, l, 2, i, j + 1, k + 2) +
                                           face_minus(a, l, 2, i + 1, j + 1, k + 2))
                                : 0.0f;
    const long o = a.mg_off[c] + T;
    a.mg_ax[o] = ax;
    a.mg_ay[o] = ay;
    a.mg_az[o] = az;
    const float d = ax + axp + ay + ayp + az + azp;
    a.mg_w[o] = kMgOmegaCoarse / (d > 1.0e-12f ? d : 1.0e-12f);
}

// ---------------------------------------------------------------------------------------------
// Down-sweep l -> l + 1 (l below the tail level): one thread per fine cell (kCells passes of the
// workgroup), 8 threads per coarse cell;
// s_{l+1} = r * sum over the 8 children of (s - A^ (w s)).
// ---------------------------------------------------------------------------------------------
// Cells per thread of the level passes and of the final up-sweep.  One is fastest on MI355X: at
// 96^3 and 192^3, 2 or 4 cells per thread (their loads issued together) lowered the occupancy
// more than they overlapped latency (3 % and 10 % slower substeps).
constexpr int kCells = 1;
constexpr int kUp0Cells = 1;

// s - A^ (w s) at fine cell (i, j, k) of level l.  Level 0 of the fp32 PCG: s = h^2 r,
// w s = omega D r; of the FP16 inner solve of a refinement step (IR), whose operator is already
// A^: s = r (FP16), w s = omega (D / h^2) r.
template <bool IR, class A>
__device__ __forceinline__ float smoothed_residual(const A& a, int l, int i, int j, int k, int n) {
    const Stencil s = stencil(i, j, k, n);
    float c[6], e[6], ec, sc;
    if (l == 0 && IR) {
        coefs0(a, s, c);
        const _Float16* r = a.ir_r;
        const float* D = a.pdiag;
        const float rhh = a.rhh;
#pragma unroll
        for (int d = 0; d < 6; ++d) e[d] = kMgOmega * (float)r[s.q[d]] * (D[s.q[d]] * rhh);
        const float rc = (float)r[s.t];
        ec = kMgOmega * rc * (D[s.t] * rhh);
        sc = rc;
    } else if (l == 0) {
        coefs0(a, s, c);
        const float* r = a.cg_r;
        const float* D = a.pdiag;
#pragma unroll
        for (int d = 0; d < 6; ++d) e[d] = kMgOmega * r[s.q[d]] * D[s.q[d]];
        const float rc = r[s.t];
        ec = kMgOmega * rc * D[s.t];
        sc = a.hh * rc;
    } else {
        const long off = a.mg_off[l];
        coefs_coarse(a.mg_ax, a.mg_ay, a.mg_az, off, s, c);
        const float* S = a.mg_s + off;
        const float* W = a.mg_w + off;
#pragma unroll
        for (int d = 0; d < 6; ++d) e[d] = S[s.q[d]] * W[s.q[d]];
        sc = S[s.t];
        ec = sc * W[s.t];
    }
    return sc - apply6(ec, e, c);
}

template <bool IR, class Src>
__global__ __launch_bounds__(kBlock) void k_mg_restrict(Src src, int l) {
    HFM_ENTRY(src);
    if (!a.cg_mg || l >= a.mg_tail) return;
    if (a.ctl[1].done) return;   // solve ended (launches queued past the stop do nothing)
    const int n = a.mg_n[l], m = a.mg_n[l + 1];
    const unsigned Mc = (unsigned)m * m * m;
    constexpr unsigned kPerPass = kBlock / 8;           // coarse cells per pass of the workgroup
    constexpr unsigned kPerBlock = kPerPass * kCells;
    const int nb = (int)((Mc + kPerBlock - 1) / kPerBlock);
    if (lb >= nb) return;
    const int lane = threadIdx.x & 63;
    const int dk = lane & 1, dj = (lane >> 4) & 1, di = lane >> 5;
    const unsigned base = (unsigned)xcd_tile(lb, nb) * kPerBlock + (threadIdx.x >> 6) * 8 + ((lane & 15) >> 1);
    float tv[kCells];
#pragma unroll
    for (int u = 0; u < kCells; ++u) {   // coarse cells past the end compute the last one, unstored
        const unsigned T = base + u * kPerPass;
        int I, J, K;
        cell3(T < Mc ? T : Mc - 1, m, I, J, K);
        tv[u] = smoothed_residual<IR>(a, l, 2 * I + di, 2 * J + dj, 2 * K + dk, n);
    }
    float* out = a.mg_s + a.mg_off[l + 1];
#pragma unroll
    for (int u = 0; u < kCells; ++u) {
        const unsigned T = base + u * kPerPass;
        const float sum = sum_children(tv[u]);
        if (T < Mc && (lane & 0x31) == 0) out[T] = kMgRestrict * sum;
    }
}

// ---------------------------------------------------------------------------------------------
// The tail: one workgroup per scene, out of LDS.
// ---------------------------------------------------------------------------------------------

// kMgCoarseSweeps weighted-Jacobi sweeps from zero of an m^3 level, one cell per thread
// (threads t >= m^3 only take part in the barriers); E0 / E1 are the LDS ping-pong buffers.
// Returns the thread's final value; the final field is left in E0 (the sweep count is even).
__device__ __forceinline__ float coarse_sweeps(int t, int m, float s, float w, const float (&c)[6],
                                               const Stencil& st, float* E0, float* E1) {
    static_assert(kMgCoarseSweeps % 2 == 0, "the final sweep lands in E0");
    const int nc = m * m * m;
    float e = 0.0f;
    if (t < nc) E0[t] = 0.0f;
    __syncthreads();
#pragma unroll 2
    for (int it = 0; it < kMgCoarseSweeps; ++it) {
        const float* ein = (it & 1) ? E1 : E0;
        float* eout = (it & 1) ? E0 : E1;
        if (t < nc) {
            float en[6];
#pragma unroll
            for (int d = 0; d < 6; ++d) en[d] = ein[st.q[d]];
            e = e + w * (s - apply6(e, en, c));
            eout[t] = e;
        }
        __syncthreads();
    }
    return e;
}

template <class Src>
__global__ __launch_bounds__(kTailBlock) void k_mg_tail(Src src) {
    HFM_ENTRY(src);
    if (!a.cg_mg) return;
    if (a.ctl[1].done) return;
    static_assert(kTailFine <= 2 * kTailBlock, "two tail cells per thread");
    __shared__ float S[kTailFine], W[kTailFine], AX[kTailFine], AY[kTailFine], AZ[kTailFine];
    __shared__ float PRE[kTailFine], RES[kTailFine];
    __shared__ float CE0[kTailCoarse], CE1[kTailCoarse];
    const int tid = threadIdx.x;
    const int T = a.mg_tail, n = a.mg_n[T];
    const int nf = n * n * n;
    const long off = a.mg_off[T];
    if (T + 1 >= a.mg_levels) {
        // the tail level is the coarsest (n <= 8): sweeps alone, PRE / RES as the ping-pong
        const int t = tid < nf ? tid : nf - 1;
        int i, j, k;
        cell3(t, n, i, j, k);
        const Stencil st = stencil(i, j, k, n);
        float c[6];
        coefs_coarse(a.mg_ax, a.mg_ay, a.mg_az, off, st, c);
        const float e = coarse_sweeps(tid, n, a.mg_s[off + t], a.mg_w[off + t], c, st, PRE, RES);
        if (tid < nf) a.mg_e[off + tid] = e;
        return;
    }
    // two levels: the tail level (n <= 12) and its coarsest (m = n / 2 <= 6); every global load
    // of the kernel is issued here, in one round
    const int m = n / 2, L = T + 1;
    const int nc = m * m * m;
    const long coff = a.mg_off[L];
    float fs[2], fw[2], fx[2], fy[2], fz[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int t0 = tid + u * kTailBlock, t = t0 < nf ? t0 : nf - 1;
        fs[u] = a.mg_s[off + t];
        fw[u] = a.mg_w[off + t];
        fx[u] = a.mg_ax[off + t];
        fy[u] = a.mg_ay[off + t];
        fz[u] = a.mg_az[off + t];
    }
    const int tc = tid < nc ? tid : nc - 1;
    int I, J, K;
    cell3(tc, m, I, J, K);
    const Stencil st = stencil(I, J, K, m);
    float c[6];
    coefs_coarse(a.mg_ax, a.mg_ay, a.mg_az, coff, st, c);
    const float w = a.mg_w[coff + tc];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int t = tid + u * kTailBlock;
        if (t < nf) {
            S[t] = fs[u];
            W[t] = fw[u];
            PRE[t] = fs[u] * fw[u];
            AX[t] = fx[u];
            AY[t] = fy[u];
            AZ[t] = fz[u];
        }
    }
    __syncthreads();
    // residual of the pre-smoothed field
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int t = tid + u * kTailBlock;
        if (t < nf) {
            int i, j, k;
            cell3(t, n, i, j, k);
            const Stencil sf = stencil(i, j, k, n);
            float cf[6], e[6];
            coefs_coarse(AX, AY, AZ, 0, sf, cf);
#pragma unroll
            for (int d = 0; d < 6; ++d) e[d] = PRE[sf.q[d]];
            RES[t] = S[t] - apply6(PRE[t], e, cf);
        }
    }
    __syncthreads();
    // restriction and the coarsest level's sweeps, one coarse cell per thread
    float ch[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) ch[q] = RES[((2 * I + (q >> 2)) * n + 2 * J + ((q >> 1) & 1)) * n + 2 * K + (q & 1)];
    const float s = kMgRestrict * sum8(ch);
    coarse_sweeps(tid, m, s, w, c, st, CE0, CE1);
    // prolongation to the tail level and its post-sweep
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int t = tid + u * kTailBlock;
        if (t < nf) {
            int i, j, k;
            cell3(t, n, i, j, k);
            Stencil sf = stencil(i, j, k, n);
            parents(sf, i, j, k, n, m);
            float cf[6], e[6];
            coefs_coarse(AX, AY, AZ, 0, sf, cf);
#pragma unroll
            for (int d = 0; d < 6; ++d) e[d] = PRE[sf.q[d]] + CE0[sf.pq[d]];
            const float ec = PRE[t] + CE0[sf.pc];
            a.mg_e[off + t] = ec + W[t] * (S[t] - apply6(ec, e, cf));
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Up-sweeps.
// ---------------------------------------------------------------------------------------------

// e = w s + e_{l+1}(parent) and one sweep at cell t of level l >= 1
template <class A>
__device__ __forceinline__ float prolong_sweep(const A& a, int l, unsigned t, int n, int m) {
    int i, j, k;
    cell3(t, n, i, j, k);
    Stencil s = stencil(i, j, k, n);
    parents(s, i, j, k, n, m);
    const long off = a.mg_off[l];
    const float* S = a.mg_s + off;
    const float* W = a.mg_w + off;
    const float* P = a.mg_e + a.mg_off[l + 1];
    float c[6], e[6];
    coefs_coarse(a.mg_ax, a.mg_ay, a.mg_az, off, s, c);
#pragma unroll
    for (int d = 0; d < 6; ++d) e[d] = S[s.q[d]] * W[s.q[d]] + P[s.pq[d]];
    const float sc = S[t], wc = W[t];
    const float ec = sc * wc + P[s.pc];
    return ec + wc * (sc - apply6(ec, e, c));
}

// l + 1 -> l for 1 <= l < tail, into mg_e
template <class Src>
__global__ __launch_bounds__(kBlock) void k_mg_up(Src src, int l) {
    HFM_ENTRY(src);
    if (!a.cg_mg || l >= a.mg_tail) return;
    if (a.ctl[1].done) return;
    const int n = a.mg_n[l], m = a.mg_n[l + 1];
    const unsigned N = (unsigned)n * n * n;
    const int nb = (nblocks(N) + kCells - 1) / kCells;
    if (lb >= nb) return;
    const unsigned base = (unsigned)xcd_tile(lb, nb) * (kBlock * kCells) + threadIdx.x;
    float ev[kCells];
#pragma unroll
    for (int u = 0; u < kCells; ++u) {
        const unsigned t = base + u * kBlock;
        ev[u] = prolong_sweep(a, l, t < N ? t : N - 1, n, m);
    }
    float* out = a.mg_e + a.mg_off[l];
#pragma unroll
    for (int u = 0; u < kCells; ++u) {
        const unsigned t = base + u * kBlock;
        if (t < N) out[t] = ev[u];
    }
}

// 1 -> 0: z = e + omega D (r - A e) with e = omega D r + e_1(parent); z and the partial sums of
// r.z (row 1 of the PCG partials, in the PCG geometry, as pass B writes them).  IR: the FP16
// inner solve of a refinement step (r and z FP16, D / h^2 for D, r.z of the rounded values, as
// its pass B forms them), with z stored as (M r) 4 / n^2: M r approximates the inverse of an
// operator whose smallest nonzero eigenvalue is about (pi / n)^2, so it is up to ~(n / pi)^2
// times r (r has RMS 1 in the inner solve) and would overflow FP16 from n ~ 128 on; CG is
// invariant to a constant factor on the preconditioner (alpha grows by it, p shrinks by it).
template <bool IR, class Src>
__global__ __launch_bounds__(kBlock) void k_mg_up0(Src src) {
    HFM_ENTRY(src);
    if (!a.cg_mg) return;
    if (a.ctl[1].done) return;
    __shared__ double lds[1][kWaves];
    const int n = a.n, m = a.mg_n[1];
    const unsigned N = (unsigned)a.N;
    const float* D = a.pdiag;
    const float* P = a.mg_e + a.mg_off[1];
    const float dsc = IR ? a.rhh : 1.0f;
    const float zscale = 4.0f / ((float)n * (float)n);
    double acc[1] = {0.0};
    const int L_ = xcd_tile(lb, a.nb_upd);
    for (int u0 = 0; u0 < a.cpt; u0 += kUp0Cells) {
        float z[kUp0Cells], rz[kUp0Cells];
#pragma unroll
        for (int u = 0; u < kUp0Cells; ++u) {
            const unsigned t0 = ((unsigned)L_ * a.cpt + u0 + u) * kBlock + threadIdx.x;
            const unsigned t = t0 < N ? t0 : N - 1;
            int i, j, k;
            cell3(t, n, i, j, k);
            Stencil s = stencil(i, j, k, n);
            parents(s, i, j, k, n, m);
            float c[6], e[6], rn[6];
            coefs0(a, s, c);
#pragma unroll
            for (int d = 0; d < 6; ++d) rn[d] = IR ? (float)a.ir_r[s.q[d]] : a.cg_r[s.q[d]];
            const float rc = IR ? (float)a.ir_r[t] : a.cg_r[t];
            if (IR) {
#pragma unroll
                for (int d = 0; d < 6; ++d) e[d] = kMgOmega * rn[d] * (D[s.q[d]] * dsc) + P[s.pq[d]];
                const float Dc = D[t] * dsc;
                const float ec = kMgOmega * rc * Dc + P[s.pc];
                const float zf = (ec + kMgOmega * Dc * (rc - apply6(ec, e, c))) * zscale;
                const _Float16 zh = (_Float16)zf;
                z[u] = (float)zh;
            } else {
#pragma unroll
                for (int d = 0; d < 6; ++d) e[d] = kMgOmega * rn[d] * D[s.q[d]] + P[s.pq[d]];
                const float Dc = D[t];
                const float ec = kMgOmega * rc * Dc + P[s.pc];
                z[u] = ec + kMgOmega * Dc * a.rhh * (a.hh * rc - apply6(ec, e, c));
            }
            rz[u] = rc * z[u];
        }
#pragma unroll
        for (int u = 0; u < kUp0Cells; ++u) {
            const unsigned t = ((unsigned)L_ * a.cpt + u0 + u) * kBlock + threadIdx.x;
            if (u0 + u < a.cpt && t < N) {
                if (IR)
                    a.ir_z[t] = (_Float16)z[u];
                else
                    a.cg_z[t] = z[u];
                acc[0] += (double)rz[u];
            }
        }
    }
    double* part = a.partials + kMaxDirBlocks + a.nb_upd;   // row 1 of pass B's partial sums
    block_sum<1, kBlock>(acc, lds);
    if (threadIdx.x == 0) part[lb] = acc[0];
}

}  // namespace

void launch_mg_coefficients(const Batch* b, int levels) {
    for (int l = 0; l + 1 < levels && l + 1 < kMgMaxLevels; ++l)
        HFM_LAUNCH(k_mg_coef, family(b, (Family)(FAM_MG0 + l + 1)), kBlock, nullptr, nullptr, l);
}

static void launch_mg_vcycle_kernels(const Batch* b, int tail, bool ir) {
    for (int l = 0; l < tail; ++l) {
        if (ir && l == 0)
            HFM_LAUNCH_T(k_mg_restrict, true, family(b, FAM_MG0), kBlock, nullptr, nullptr, l);
        else
            HFM_LAUNCH_T(k_mg_restrict, false, family(b, (Family)(FAM_MG0 + l)), kBlock, nullptr, nullptr, l);
    }
    HFM_LAUNCH(k_mg_tail, family(b, FAM_MGC), kTailBlock, nullptr, nullptr);
    for (int l = tail - 1; l >= 1; --l)
        HFM_LAUNCH(k_mg_up, family(b, (Family)(FAM_MG0 + l)), kBlock, nullptr, nullptr, l);
    if (ir)
        HFM_LAUNCH_T(k_mg_up0, true, family(b, FAM_PCG), kBlock, nullptr, nullptr);
    else
        HFM_LAUNCH_T(k_mg_up0, false, family(b, FAM_PCG), kBlock, nullptr, nullptr);
}

// The V-cycle's launches are the same in every PCG iteration of every solve of a batch: they take no
// iteration index, and their block maps and descriptors are the batch's own, fixed until batch_prepare
// rebuilds them.  With HFM_MG_GRAPH=1 (read at every call) the batch captures them once into a HIP graph
// (kernel nodes only) and replays it with one hipGraphLaunch per iteration, one host call instead of
// 2 tail + 1 launches; the graph is captured again when the tail depth changes or the batch's maps were
// rebuilt (its kernel arguments point at the maps, or hold the descriptor of a batch of one).  The replays
// give the same bits, but on MI355X with ROCm 7.2 the replayed chain takes longer on the device than the
// same launches on the stream (0.87 against 0.81 ms per multigrid substep at 96^3, 0.48 against 0.46 at
// 48^3), so the launches one by one are the default.
int launch_mg_vcycle(Batch* b, int tail, bool ir) {
    const char* env = getenv("HFM_MG_GRAPH");
    if (!env || atoi(env) == 0) {
        launch_mg_vcycle_kernels(b, tail, ir);
        return HFM_OK;
    }
    const int g = ir ? 1 : 0;
    if (!b->mg_exec[g] || b->mg_exec_tail[g] != tail || b->mg_exec_epoch[g] != b->maps_epoch) {
        if (b->mg_exec[g]) (void)hipGraphExecDestroy(b->mg_exec[g]);
        b->mg_exec[g] = nullptr;
        hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            set_error("multigrid V-cycle: stream capture could not begin: %s", hipGetErrorString(e));
            return HFM_ERR_HIP;
        }
        launch_mg_vcycle_kernels(b, tail, ir);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(b->stream, &graph);
        if (e == hipSuccess) e = hipGraphInstantiate(&b->mg_exec[g], graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            b->mg_exec[g] = nullptr;
            set_error("multigrid V-cycle: graph capture failed: %s", hipGetErrorString(e));
            return HFM_ERR_HIP;
        }
        b->mg_exec_tail[g] = tail;
        b->mg_exec_epoch[g] = b->maps_epoch;
    }
    const hipError_t e = hipGraphLaunch(b->mg_exec[g], b->stream);
    if (e != hipSuccess) {
        set_error("multigrid V-cycle: graph launch failed: %s", hipGetErrorString(e));
        return HFM_ERR_HIP;
    }
    return HFM_OK;
}

}  // namespace hfm
