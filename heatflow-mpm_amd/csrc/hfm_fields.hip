// hfm_fields.hip — per-cell and per-face passes of the substep (gfx950), batched over scenes.
//
// Each pass fuses a run of the reference's Warp launches that touch the same cells, keeping
// their arithmetic order:
//   heat_prep   k_blend + k_apply_fixed + k_source                     (heat.py:21-28,147-158)
//   advect_T    k_advect + k_conductivity + k_heat_rhs + k_heat_diag   (heat.py:31-47,108-144)
//   air_faces   k_face_fraction + k_advect_face + k_buoyancy + k_relax (grid.py:25-79)
//   divergence  k_divergence_rhs + k_jacobi                            (grid.py:82-137)
//   rhs_mean    numpy float32 mean of the rhs                          (grid.py:238-241)
//   project     k_project                                              (grid.py:140-153)
//   g2p         k_particle_update                                      (heat.py:161-174)
// Buffer roles are fixed, so the per-scene descriptors never change between substeps: heat_prep
// keeps the previous T in T2 (what the reference's swap leaves there, solver.py:85) and
// advection writes the new T in place; the face pass writes the advected velocity to us and the
// projection writes the projected velocity back to u.
#include "hfm_device.h"
#include "hfm_scene.h"

namespace hfm {

namespace {

__device__ __forceinline__ void face_ijk(int axis, long f, int n, int& i, int& j, int& k) {
    const unsigned u = (unsigned)n, v = u + 1;
    split3((unsigned)f, axis == 1 ? v : u, axis == 2 ? v : u, i, j, k);
}

// ---------------------------------------------------------------------------------------------
// heat prep: T2 = T (previous temperature); Tw = blend(T); fixed; capped source
// ---------------------------------------------------------------------------------------------
// With a source, each workgroup also stores the fp64 sum of its cells' source rises (the new
// minus the old Tw, exact in fp32) in src_part[lb]; the heat solve's first launch sums them.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_heat_prep(Src src, float dt) {
    HFM_ENTRY(src);
    const long t = (long)xcd_tile(lb, nblocks(a.N)) * kBlock + threadIdx.x;
    double rise[1] = {0.0};
    if (t < a.N) {
        const float Tc = a.T[t];
        a.T2[t] = Tc;
        float tw = Tc;
        if (a.P > 0) {
            float al = a.alpha_p[t];
            if (al > 0.05f) tw = a.tsum[t] / al;
        }
        const uint8_t fl = a.flags[t];
        if (fl & FLAG_FIXED) tw = a.tfix[t];
        if (a.rate_on && (fl & FLAG_SOURCE)) {
            float up = tw + a.rate * dt;
            float lo = tw > a.t_cap ? tw : a.t_cap;
            const float tw0 = tw;
            tw = up < lo ? up : lo;
            rise[0] = (double)(tw - tw0);
        }
        a.Tw[t] = tw;
    }
    if (a.rate_on) {   // uniform over the workgroup: every thread reaches the reduction's barrier
        __shared__ double lds[1][kWaves];
        block_sum<1>(rise, lds);
        if (threadIdx.x == 0) a.src_part[lb] = rise[0];
    }
}

// ---------------------------------------------------------------------------------------------
// T advection (into T) + conductivity + heat rhs/diag
// ---------------------------------------------------------------------------------------------
// The rest of the T pass once the advected value t2 of cell t is known: conductivity
// (heat.py:44-47) and the heat system's rhs and Jacobi diagonal (heat.py:116-144).
template <class A>
__device__ __forceinline__ void heat_tail(const A& a, long t, int i, int j, int k, float t2, float ac, float dt) {
    const int n = a.n;
    const long nn = (long)n * n;
    const float h = a.h;
    a.T[t] = t2;
    const float kc = conductivity(ac, a.ka, a.km);
    a.kap[t] = kc;
    if (a.flags[t] & FLAG_FIXED) {
        a.hrhs[t] = a.tfix[t];
        a.hdiag[t] = 1.0f;
        return;
    }
    const float dt_hh = dt / (h * h);
    const long nb[6] = {t - nn, t + nn, t - n, t + n, t - 1, t + 1};
    const bool in[6] = {i > 0, i < n - 1, j > 0, j < n - 1, k > 0, k < n - 1};
    float ntf = 0.0f, sd = 0.0f;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        if (!in[s]) continue;
        float kf = face_kappa(kc, conductivity(a.alpha_s[nb[s]], a.ka, a.km));
        sd += kf;
        if (a.flags[nb[s]] & FLAG_FIXED) ntf += kf * a.tfix[nb[s]];
    }
    a.hrhs[t] = t2 + dt_hh * ntf;
    a.hdiag[t] = 1.0f / (1.0f + dt_hh * sd);
}

// heat.py:31-41 (T advection) fused with the conductivity and heat rhs/diagonal.  With
// HFM_FLAG_MACCORMACK this pass only stores the semi-Lagrangian value; k_mc_T finishes the cell.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_advect_T(Src src, float dt) {
    HFM_ENTRY(src);
    const CellSize hs(a.h, a.rh, a.fast_h);
    const long t = (long)xcd_tile(lb, nblocks(a.N)) * kBlock + threadIdx.x;
    if (t >= a.N) return;
    const int n = a.n;
    int i, j, k;
    split3((unsigned)t, (unsigned)n, (unsigned)n, i, j, k);
    const float h = a.h;
    const float ac = a.alpha_s[t];
    float t2;
    if (ac > 0.5f) {
        t2 = a.Tw[t];
    } else {
        float x = ((float)i + 0.5f) * h, y = ((float)j + 0.5f) * h, z = ((float)k + 0.5f) * h;
        backtrace(a.u[0], a.u[1], a.u[2], x, y, z, hs, n, dt);
        t2 = sample(a.Tw, x, y, z, 0.5f, 0.5f, 0.5f, hs, n, n, n);
    }
    if (a.maccormack) {
        a.mc_T[t] = t2;
        return;
    }
    heat_tail(a, t, i, j, k, t2, ac, dt);
}

// MacCormack step of T (HFM_FLAG_MACCORMACK): with T^ the semi-Lagrangian field of k_advect_T,
// T~ = T^ sampled at the point traced forward (RK2 with -dt) and d the departure point,
//   T = clamp(T^ + (Tw - T~) / 2, value_range(Tw, d)),
// then the same rest of the pass.  Cells kept by the solid fraction keep Tw.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_mc_T(Src src, float dt) {
    HFM_ENTRY(src);
    const CellSize hs(a.h, a.rh, a.fast_h);
    if (!a.maccormack) return;
    const long t = (long)xcd_tile(lb, nblocks(a.N)) * kBlock + threadIdx.x;
    if (t >= a.N) return;
    const int n = a.n;
    int i, j, k;
    split3((unsigned)t, (unsigned)n, (unsigned)n, i, j, k);
    const float h = a.h;
    const float ac = a.alpha_s[t];
    float t2 = a.mc_T[t];
    if (!(ac > 0.5f)) {
        const float xc = ((float)i + 0.5f) * h, yc = ((float)j + 0.5f) * h, zc = ((float)k + 0.5f) * h;
        float dx = xc, dy = yc, dz = zc, fx = xc, fy = yc, fz = zc;
        backtrace(a.u[0], a.u[1], a.u[2], dx, dy, dz, hs, n, dt);
        backtrace(a.u[0], a.u[1], a.u[2], fx, fy, fz, hs, n, -dt);
        const float back = sample(a.mc_T, fx, fy, fz, 0.5f, 0.5f, 0.5f, hs, n, n, n);
        float v = t2 + 0.5f * (a.Tw[t] - back);
        float lo, hi;
        value_range(a.Tw, dx, dy, dz, 0.5f, 0.5f, 0.5f, hs, n, n, n, lo, hi);
        t2 = v < lo ? lo : (v > hi ? hi : v);
    }
    heat_tail(a, t, i, j, k, t2, ac, dt);
}

// ---------------------------------------------------------------------------------------------
// particle update (G2P): grid temperature before diffusion = Tw, after = T
// ---------------------------------------------------------------------------------------------
template <class Src>
__global__ __launch_bounds__(kBlock) void k_g2p(Src src) {
    HFM_ENTRY(src);
    const CellSize hs(a.h, a.rh, a.fast_h);
    if (lb == 0 && threadIdx.x == 0) a.ctl[0].geom_dirty = 0;  // this substep's gather rebuilt the geometry
    const long p = (long)lb * kBlock + threadIdx.x;
    if (p >= a.P) return;
    const float x = a.pos[3 * p], y = a.pos[3 * p + 1], z = a.pos[3 * p + 2];
    const int n = a.n;
    const float dT = sample(a.T, x, y, z, 0.5f, 0.5f, 0.5f, hs, n, n, n) -
                     sample(a.Tw, x, y, z, 0.5f, 0.5f, 0.5f, hs, n, n, n);
    float tp = a.temp[p] + dT;
    tp = tp < 0.0f ? 0.0f : (tp > a.t_max ? a.t_max : tp);
    a.temp[p] = tp;
    float f = (tp - a.t_sol) / (a.t_liq - a.t_sol);
    f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
    a.melt[p] = f;
    a.yield[p] = a.ys * (1.0f - f);
    a.visc[p] = a.mus * (1.0f - f) + a.mul * f;
}

// ---------------------------------------------------------------------------------------------
// air face pass: advect (old u, v, w) -> buoyancy (new T) -> relax, into us; sub-range = axis
// ---------------------------------------------------------------------------------------------
// Face position, lattice offsets and dims of face f of an axis (fields.py:75-108).  Written with selects
// (no assignment through a branch on the axis, which the compiler kept as a per-thread array in scratch
// memory); (float)i + 0 is (float)i exactly, so the positions are the reference's.
__device__ __forceinline__ void face_geometry(int axis, int i, int j, int k, float h, int n, float& x, float& y,
                                              float& z, float& ox, float& oy, float& oz, int& dx, int& dy, int& dz) {
    ox = axis == 0 ? 0.0f : 0.5f;
    oy = axis == 1 ? 0.0f : 0.5f;
    oz = axis == 2 ? 0.0f : 0.5f;
    x = ((float)i + ox) * h;
    y = ((float)j + oy) * h;
    z = ((float)k + oz) * h;
    dx = n + (axis == 0 ? 1 : 0);
    dy = n + (axis == 1 ? 1 : 0);
    dz = n + (axis == 2 ? 1 : 0);
}

// The rest of the face pass once the advected value of an interior face is known: buoyancy
// (grid.py:36-43, new T) and implicit relaxation (grid.py:70-79).
template <class A>
__device__ __forceinline__ float face_tail(const A& a, int axis, long f, int i, int j, int k, float val, float dt) {
    const int n = a.n;
    const float h = a.h;
    const long c = ((long)i * n + j) * n + k;
    if (axis == 2 && a.buoyancy) {
        float tf = 0.5f * (a.T[c - 1] + a.T[c]);
        val = val + dt * a.g * (tf - a.t_amb) / (a.t_amb + 273.15f);
    }
    const long cl = axis == 0 ? c - (long)n * n : (axis == 1 ? c - n : c - 1);
    const float af = face_fraction(a.alpha_s[cl], a.alpha_s[c], a.amin);
    const float am = af > a.amin ? af : a.amin;
    const float cc = 1.0f / (a.rho * am * h * h * h);
    return (val + dt * cc * a.mf[axis][f]) / (1.0f + dt * cc * a.kf[axis][f]);
}

// grid.py:25-33 (face advection of u, v, w from the old fields) fused with buoyancy and
// relaxation.  With HFM_FLAG_MACCORMACK this pass only stores the semi-Lagrangian value;
// k_mc_faces finishes the face.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_air_faces(Src src, float dt) {
    HFM_ENTRY(src);
    const CellSize hs(a.h, a.rh, a.fast_h);
    const int axis = sub;
    const long f = (long)xcd_tile(lb, nblocks(a.F)) * kBlock + threadIdx.x;
    if (f >= a.F) return;
    const int n = a.n;
    int i, j, k;
    face_ijk(axis, f, n, i, j, k);
    const int idx = axis == 0 ? i : (axis == 1 ? j : k);
    float* out = a.maccormack ? a.mc_u[axis] : a.us[axis];
    if (idx == 0 || idx == n) {
        out[f] = 0.0f;
        return;
    }
    const float h = a.h;
    float x, y, z, ox, oy, oz;
    int dx, dy, dz;
    face_geometry(axis, i, j, k, h, n, x, y, z, ox, oy, oz, dx, dy, dz);
    backtrace(a.u[0], a.u[1], a.u[2], x, y, z, hs, n, dt);
    const float val = sample(a.u[axis], x, y, z, ox, oy, oz, hs, dx, dy, dz);
    out[f] = a.maccormack ? val : face_tail(a, axis, f, i, j, k, val, dt);
}

// MacCormack step of the faces (HFM_FLAG_MACCORMACK), as k_mc_T with the old component u[axis]
// as the start field and both traces through the old u, v, w; then buoyancy and relaxation.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_mc_faces(Src src, float dt) {
    HFM_ENTRY(src);
    const CellSize hs(a.h, a.rh, a.fast_h);
    if (!a.maccormack) return;
    const int axis = sub;
    const long f = (long)xcd_tile(lb, nblocks(a.F)) * kBlock + threadIdx.x;
    if (f >= a.F) return;
    const int n = a.n;
    int i, j, k;
    face_ijk(axis, f, n, i, j, k);
    const int idx = axis == 0 ? i : (axis == 1 ? j : k);
    float* out = a.us[axis];
    if (idx == 0 || idx == n) {
        out[f] = 0.0f;
        return;
    }
    const float h = a.h;
    float x, y, z, ox, oy, oz;
    int dx, dy, dz;
    face_geometry(axis, i, j, k, h, n, x, y, z, ox, oy, oz, dx, dy, dz);
    float px = x, py = y, pz = z, fx = x, fy = y, fz = z;
    backtrace(a.u[0], a.u[1], a.u[2], px, py, pz, hs, n, dt);
    backtrace(a.u[0], a.u[1], a.u[2], fx, fy, fz, hs, n, -dt);
    const float fwd = a.mc_u[axis][f];
    const float back = sample(a.mc_u[axis], fx, fy, fz, ox, oy, oz, hs, dx, dy, dz);
    float v = fwd + 0.5f * (a.u[axis][f] - back);
    float lo, hi;
    value_range(a.u[axis], px, py, pz, ox, oy, oz, hs, dx, dy, dz, lo, hi);
    v = v < lo ? lo : (v > hi ? hi : v);
    out[f] = face_tail(a, axis, f, i, j, k, v, dt);
}

// ---------------------------------------------------------------------------------------------
// divergence rhs (+ Jacobi diagonal)
// ---------------------------------------------------------------------------------------------
template <class Scene_, class Vel, class Out>
__device__ __forceinline__ void divergence_cell(const Scene_& a, const Vel& vel, long t, float scale, Out rhs,
                                                Out pdiag) {
    const int n = a.n;
    const long nn = (long)n * n;
    int i, j, k;
    split3((unsigned)t, (unsigned)n, (unsigned)n, i, j, k);
    const auto as = a.alpha_s;
    const float amin = a.amin;
    const float ac = as[t];
    const float ax0 = i > 0 ? face_fraction(as[t - nn], ac, amin) : 0.0f;
    const float ax1 = i < n - 1 ? face_fraction(ac, as[t + nn], amin) : 0.0f;
    const float ay0 = j > 0 ? face_fraction(as[t - n], ac, amin) : 0.0f;
    const float ay1 = j < n - 1 ? face_fraction(ac, as[t + n], amin) : 0.0f;
    const float az0 = k > 0 ? face_fraction(as[t - 1], ac, amin) : 0.0f;
    const float az1 = k < n - 1 ? face_fraction(ac, as[t + 1], amin) : 0.0f;
    const long fx0 = ((long)i * n + j) * n + k, fx1 = fx0 + nn;
    const long fy0 = ((long)i * (n + 1) + j) * n + k, fy1 = fy0 + n;
    const long fz0 = ((long)i * n + j) * (n + 1) + k, fz1 = fz0 + 1;
    const auto u = vel[0], v = vel[1], w = vel[2];
    const auto qx = a.q[0], qy = a.q[1], qz = a.q[2];
    float div = ax1 * u[fx1] + qx[fx1] - ax0 * u[fx0] - qx[fx0] + ay1 * v[fy1] + qy[fy1] - ay0 * v[fy0] - qy[fy0] +
                az1 * w[fz1] + qz[fz1] - az0 * w[fz0] - qz[fz0];
    div = div / a.h;
    rhs[t] = -scale * div;
    if (pdiag) {
        float s = 0.0f;
        if (i > 0) s += ax0;
        if (i < n - 1) s += ax1;
        if (j > 0) s += ay0;
        if (j < n - 1) s += ay1;
        if (k > 0) s += az0;
        if (k < n - 1) s += az1;
        const float m = s > 1.0e-12f ? s : 1.0e-12f;
        pdiag[t] = (a.h * a.h) / m;
    }
}

template <class Src>
__global__ __launch_bounds__(kBlock) void k_divergence(Src src, double dt) {
    HFM_ENTRY(src);
    const long t = (long)xcd_tile(lb, nblocks(a.N)) * kBlock + threadIdx.x;
    if (t >= a.N) return;
    const float scale = (float)(a.rho_d / dt);  // rho/dt formed in double (grid.py:237)
    divergence_cell(a, a.us, t, scale, a.rhs, a.pdiag);
}

__global__ __launch_bounds__(kBlock) void k_divergence1(SceneDev a, float* out) {
    const long t = (long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= a.N) return;
    divergence_cell(a, a.u, t, 1.0f, out, (float*)nullptr);
}

// ---------------------------------------------------------------------------------------------
// numpy float32 mean of the rhs: buffer chunks of 8192, pairwise sum per chunk (leaves of at most
// 128 elements with 8 interleaved accumulators), chunk sums added in order to a float32 total.
// One workgroup per chunk; the chunk's tree is given as a leaf table and a postfix program.
// ---------------------------------------------------------------------------------------------
constexpr int kChunk = 8192;
constexpr int kMaxLeaves = 256;
constexpr int kMaxChunksLds = 2048;

__device__ float leaf_sum(const float* a, int len) {
    if (len < 8) {
        float res = 0.0f;
        for (int q = 0; q < len; ++q) res += a[q];
        return res;
    }
    float r[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) r[q] = a[q];
    int i = 8;
    for (; i < len - (len % 8); i += 8) {
#pragma unroll
        for (int q = 0; q < 8; ++q) r[q] += a[i + q];
    }
    float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < len; ++i) res += a[i];
    return res;
}

template <class Src>
__global__ __launch_bounds__(kBlock) void k_rhs_mean(Src src) {
    HFM_ENTRY(src);
    __shared__ float leaves[kMaxLeaves];
    __shared__ int is_last;
    const int c = lb;
    const bool tail = (c == a.nchunks - 1) && a.mean_tail >= 0;
    const int* prog = tail ? a.mean_prog + a.mean_tail : a.mean_prog;
    const int nleaves = prog[0], nops = prog[1];
    const int* lv = prog + 2;
    const int* ops = prog + 2 + 2 * nleaves;
    const float* base = a.rhs + (long)c * kChunk;
    CgCtl* ctl = a.ctl + 1;
    if (!tail && a.N >= kChunk) {
        // a full chunk: 64 leaves of 128 in order, and numpy's tree over them is balanced, so wave 0
        // sums aligned pairs, then aligned quads, ... (fp32 addition commutes: the butterfly gives
        // the bits of the tree's left + right at every node)
        // the chunk is staged in LDS with coalesced loads, each 128-element leaf padded to 129
        // floats so that the 64 lanes summing one leaf each read different banks
        // (all 32 loads of a thread are issued before the first LDS store, so their latencies overlap)
        __shared__ float buf[64 * 129];
        constexpr int kPer = kChunk / kBlock;
        float v[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) v[u] = base[threadIdx.x + u * kBlock];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = threadIdx.x + u * kBlock;
            buf[(e >> 7) * 129 + (e & 127)] = v[u];
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            float x = leaf_sum(buf + 129 * threadIdx.x, 128);
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) x += __shfl_xor(x, off);
            if (threadIdx.x == 0) {
                __hip_atomic_store(&a.chunk_sums[c], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                is_last = (take_ticket(&ctl->ticket2) == (unsigned)(a.nchunks - 1));
            }
        }
        __syncthreads();
    } else {
        for (int l = threadIdx.x; l < nleaves; l += kBlock) leaves[l] = leaf_sum(base + lv[2 * l], lv[2 * l + 1]);
        __syncthreads();
        if (threadIdx.x == 0) {
            float stk[16];
            int sp = 0;
            for (int o = 0; o < nops; ++o) {
                int op = ops[o];
                if (op >= 0) {
                    stk[sp++] = leaves[op];
                } else {
                    float r = stk[--sp];
                    float l = stk[--sp];
                    stk[sp++] = l + r;
                }
            }
            __hip_atomic_store(&a.chunk_sums[c], stk[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            is_last = (take_ticket(&ctl->ticket2) == (unsigned)(a.nchunks - 1));
        }
        __syncthreads();
    }
    if (!is_last) return;
    // the chunk sums are loaded by all threads at once, then added in chunk order by one thread
    __shared__ float cs[kMaxChunksLds];
    for (int q0 = 0; q0 < a.nchunks; q0 += kMaxChunksLds) {
        const int m = a.nchunks - q0 < kMaxChunksLds ? a.nchunks - q0 : kMaxChunksLds;
        for (int q = threadIdx.x; q < m; q += kBlock)
            cs[q] = __hip_atomic_load(&a.chunk_sums[q0 + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (threadIdx.x == 0) {
            float s = q0 == 0 ? 0.0f : leaves[0];
            for (int q = 0; q < m; ++q) s += cs[q];
            leaves[0] = s;
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const float s = leaves[0];
    const float mean = (float)((double)s / (double)a.N);
    ctl->mean = mean;
    __hip_atomic_store(&a.mirror[1].mean, (double)mean, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&ctl->ticket2, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------------------------
// projection: u = us - (dt/rho) grad p / h on interior faces, walls 0; sub-range = axis
// ---------------------------------------------------------------------------------------------
template <class Src>
__global__ __launch_bounds__(kBlock) void k_project(Src src, double dt) {
    HFM_ENTRY(src);
    const int axis = sub;
    const long f = (long)xcd_tile(lb, nblocks(a.F)) * kBlock + threadIdx.x;
    if (f >= a.F) return;
    const int n = a.n;
    int i, j, k;
    face_ijk(axis, f, n, i, j, k);
    const int idx = axis == 0 ? i : (axis == 1 ? j : k);
    if (idx == 0 || idx == n) {
        a.u[axis][f] = 0.0f;
        return;
    }
    const float dor = (float)(dt / a.rho_d);  // dt/rho formed in double (grid.py:251)
    const long c = ((long)i * n + j) * n + k;
    const long cl = axis == 0 ? c - (long)n * n : (axis == 1 ? c - n : c - 1);
    const float gp = a.p[c] - a.p[cl];
    const float val = a.us[axis][f] - dor * gp / a.h;
    a.u[axis][f] = val;
}

// ---------------------------------------------------------------------------------------------
// single-scene accessors and diagnostics
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_cell_velocity(int n, long N, const float* u, const float* v,
                                                          const float* w, float* out) {
    const long t = (long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= N) return;
    const long nn = (long)n * n;
    int i, j, k;
    split3((unsigned)t, (unsigned)n, (unsigned)n, i, j, k);
    const long fx0 = ((long)i * n + j) * n + k;
    const long fy0 = ((long)i * (n + 1) + j) * n + k;
    const long fz0 = ((long)i * n + j) * (n + 1) + k;
    out[3 * t + 0] = 0.5f * (u[fx0] + u[fx0 + nn]);
    out[3 * t + 1] = 0.5f * (v[fy0] + v[fy0 + n]);
    out[3 * t + 2] = 0.5f * (w[fz0] + w[fz0 + 1]);
}

__global__ __launch_bounds__(kBlock) void k_face_fraction(int n, long F, int axis, float amin, const float* alpha_s,
                                                          float* out) {
    const long f = (long)blockIdx.x * kBlock + threadIdx.x;
    if (f >= F) return;
    int i, j, k;
    face_ijk(axis, f, n, i, j, k);
    const int idx = axis == 0 ? i : (axis == 1 ? j : k);
    if (idx == 0 || idx == n) {
        out[f] = 0.0f;
        return;
    }
    const long c = ((long)i * n + j) * n + k;
    const long cl = axis == 0 ? c - (long)n * n : (axis == 1 ? c - n : c - 1);
    out[f] = face_fraction(alpha_s[cl], alpha_s[c], amin);
}

__global__ __launch_bounds__(kBlock) void k_abs_max(const float* a, long count, unsigned* slot) {
    __shared__ unsigned m;
    if (threadIdx.x == 0) m = 0u;
    __syncthreads();
    unsigned mine = 0u;
    for (long t = (long)blockIdx.x * kBlock + threadIdx.x; t < count; t += (long)gridDim.x * kBlock) {
        unsigned b = __float_as_uint(fabsf(a[t]));
        if (b > 0x7f800000u) b = 0x7f800000u;  // NaN reported as inf
        mine = b > mine ? b : mine;
    }
    atomicMax(&m, mine);
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(slot, m);
}

__global__ __launch_bounds__(kBlock) void k_apply_fixed(long N, const uint8_t* flags, const float* tfix, float* T) {
    const long t = (long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= N) return;
    if (flags[t] & FLAG_FIXED) T[t] = tfix[t];
}

}  // namespace

void launch_heat_prep(const Launch& L, float dt) { HFM_LAUNCH(k_heat_prep, L, kBlock, nullptr, nullptr, dt); }
void launch_advect_T(const Launch& L, float dt) { HFM_LAUNCH(k_advect_T, L, kBlock, nullptr, nullptr, dt); }
void launch_g2p(const Launch& L) { HFM_LAUNCH(k_g2p, L, kBlock, nullptr, nullptr); }
void launch_air_faces(const Launch& L, float dt) { HFM_LAUNCH(k_air_faces, L, kBlock, nullptr, nullptr, dt); }

void launch_mc_T(const Launch& L, float dt) { HFM_LAUNCH(k_mc_T, L, kBlock, nullptr, nullptr, dt); }

void launch_mc_faces(const Launch& L, float dt) { HFM_LAUNCH(k_mc_faces, L, kBlock, nullptr, nullptr, dt); }
void launch_divergence(const Launch& L, double dt) { HFM_LAUNCH(k_divergence, L, kBlock, nullptr, nullptr, dt); }
void launch_rhs_mean(const Launch& L) { HFM_LAUNCH(k_rhs_mean, L, kBlock, nullptr, nullptr); }
void launch_project(const Launch& L, double dt) { HFM_LAUNCH(k_project, L, kBlock, nullptr, nullptr, dt); }

void launch_divergence_scale1(const Scene* s, hipStream_t st, float* out) {
    hipLaunchKernelGGL(k_divergence1, dim3(nblocks(s->N)), dim3(kBlock), 0, st, s->dev, out);
}

void launch_cell_velocity(const Scene* s, hipStream_t st, float* out) {
    hipLaunchKernelGGL(k_cell_velocity, dim3(nblocks(s->N)), dim3(kBlock), 0, st, s->n, s->N, s->dev.u[0],
                       s->dev.u[1], s->dev.u[2], out);
}

void launch_face_fraction(const Scene* s, hipStream_t st, int axis, float* out) {
    hipLaunchKernelGGL(k_face_fraction, dim3(nblocks(s->F)), dim3(kBlock), 0, st, s->n, s->F, axis, s->dev.amin,
                       s->dev.alpha_s, out);
}

void launch_abs_max(hipStream_t st, const float* a, long count, unsigned* slot) {
    int nb = nblocks(count);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_abs_max, dim3(nb), dim3(kBlock), 0, st, a, count, slot);
}

void launch_apply_fixed(const Scene* s, hipStream_t st, float* T) {
    hipLaunchKernelGGL(k_apply_fixed, dim3(nblocks(s->N)), dim3(kBlock), 0, st, s->N, s->dev.flags, s->dev.tfix, T);
}

}  // namespace hfm
