// hfm_fields.hip — per-cell and per-face passes of the substep (gfx950).
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
#include "hfm_device.h"
#include "hfm_scene.h"

namespace hfm {

namespace {

inline int nblocks(long count) { return (int)((count + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------------------------------------
// heat prep: Tw = blend(T) ; fixed ; capped source
// ---------------------------------------------------------------------------------------------
struct HeatPrepArgs {
    long N;
    const float* T;
    const float* alpha_p;  // null without particles
    const float* tsum;
    const uint8_t* flags;
    const float* tfix;
    float* Tw;
    float rate, t_cap, dt;
    int rate_on;
};

__global__ __launch_bounds__(kBlock) void k_heat_prep(HeatPrepArgs a) {
    long t = (long)xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (t >= a.N) return;
    float tw = a.T[t];
    if (a.alpha_p) {
        float al = a.alpha_p[t];
        if (al > 0.05f) tw = a.tsum[t] / al;
    }
    uint8_t fl = a.flags[t];
    if (fl & FLAG_FIXED) tw = a.tfix[t];
    if (a.rate_on && (fl & FLAG_SOURCE)) {
        float up = tw + a.rate * a.dt;
        float lo = tw > a.t_cap ? tw : a.t_cap;
        tw = up < lo ? up : lo;
    }
    a.Tw[t] = tw;
}

// ---------------------------------------------------------------------------------------------
// T advection + conductivity + heat rhs/diag
// ---------------------------------------------------------------------------------------------
struct AdvectTArgs {
    int n;
    long N;
    float h, dt, dt_hh, ka, km;
    const float *u, *v, *w;
    const float* Tw;
    const float* alpha_s;
    const uint8_t* flags;
    const float* tfix;
    float *T2, *kap, *hrhs, *hdiag;
};

__global__ __launch_bounds__(kBlock) void k_advect_T(AdvectTArgs a) {
    long t = (long)xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (t >= a.N) return;
    const int n = a.n;
    const long nn = (long)n * n;
    int i = (int)(t / nn), j = (int)((t / n) % n), k = (int)(t % n);
    float ac = a.alpha_s[t];
    float t2;
    if (ac > 0.5f) {
        t2 = a.Tw[t];
    } else {
        float x = ((float)i + 0.5f) * a.h, y = ((float)j + 0.5f) * a.h, z = ((float)k + 0.5f) * a.h;
        backtrace(a.u, a.v, a.w, x, y, z, a.h, n, a.dt);
        t2 = sample(a.Tw, x, y, z, 0.5f, 0.5f, 0.5f, a.h, n, n, n);
    }
    a.T2[t] = t2;
    float kc = conductivity(ac, a.ka, a.km);
    a.kap[t] = kc;
    if (a.flags[t] & FLAG_FIXED) {
        a.hrhs[t] = a.tfix[t];
        a.hdiag[t] = 1.0f;
        return;
    }
    long nb[6] = {t - nn, t + nn, t - n, t + n, t - 1, t + 1};
    bool in[6] = {i > 0, i < n - 1, j > 0, j < n - 1, k > 0, k < n - 1};
    float ntf = 0.0f, sd = 0.0f;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        if (!in[s]) continue;
        float kf = face_kappa(kc, conductivity(a.alpha_s[nb[s]], a.ka, a.km));
        sd += kf;
        if (a.flags[nb[s]] & FLAG_FIXED) ntf += kf * a.tfix[nb[s]];
    }
    a.hrhs[t] = t2 + a.dt_hh * ntf;
    a.hdiag[t] = 1.0f / (1.0f + a.dt_hh * sd);
}

// ---------------------------------------------------------------------------------------------
// particle update (G2P)
// ---------------------------------------------------------------------------------------------
struct G2PArgs {
    long P;
    int n;
    float h;
    const float* pos;
    const float *T_before, *T_after;
    float *temp, *melt, *yield, *visc;
    float t_sol, t_liq, ys, mus, mul, t_max;
};

__global__ __launch_bounds__(kBlock) void k_g2p(G2PArgs a) {
    long p = (long)blockIdx.x * kBlock + threadIdx.x;
    if (p >= a.P) return;
    float x = a.pos[3 * p], y = a.pos[3 * p + 1], z = a.pos[3 * p + 2];
    const int n = a.n;
    float dT = sample(a.T_after, x, y, z, 0.5f, 0.5f, 0.5f, a.h, n, n, n) -
               sample(a.T_before, x, y, z, 0.5f, 0.5f, 0.5f, a.h, n, n, n);
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
// air face pass: advect (old u, v, w) -> buoyancy (new T) -> relax; gridDim.y = axis
// ---------------------------------------------------------------------------------------------
struct FaceArgs {
    int n;
    long F;
    float h, dt, g, t_amb, rho, amin;
    int buoyancy;
    const float* u[3];
    float* out[3];
    const float* kf[3];
    const float* mf[3];
    const float* T;
    const float* alpha_s;
};

__device__ __forceinline__ void face_ijk(int axis, long f, int n, int& i, int& j, int& k) {
    if (axis == 0) {
        i = (int)(f / ((long)n * n));
        j = (int)((f / n) % n);
        k = (int)(f % n);
    } else if (axis == 1) {
        i = (int)(f / ((long)(n + 1) * n));
        j = (int)((f / n) % (n + 1));
        k = (int)(f % n);
    } else {
        i = (int)(f / ((long)n * (n + 1)));
        j = (int)((f / (n + 1)) % n);
        k = (int)(f % (n + 1));
    }
}

__global__ __launch_bounds__(kBlock) void k_air_faces(FaceArgs a) {
    const int axis = blockIdx.y;
    long f = (long)xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (f >= a.F) return;
    const int n = a.n;
    int i, j, k;
    face_ijk(axis, f, n, i, j, k);
    int idx = axis == 0 ? i : (axis == 1 ? j : k);
    float* out = a.out[axis];
    if (idx == 0 || idx == n) {
        out[f] = 0.0f;
        return;
    }
    const float h = a.h;
    float x, y, z, ox, oy, oz;
    int dx = n, dy = n, dz = n;
    if (axis == 0) {
        x = (float)i * h; y = ((float)j + 0.5f) * h; z = ((float)k + 0.5f) * h;
        ox = 0.0f; oy = 0.5f; oz = 0.5f; dx = n + 1;
    } else if (axis == 1) {
        x = ((float)i + 0.5f) * h; y = (float)j * h; z = ((float)k + 0.5f) * h;
        ox = 0.5f; oy = 0.0f; oz = 0.5f; dy = n + 1;
    } else {
        x = ((float)i + 0.5f) * h; y = ((float)j + 0.5f) * h; z = (float)k * h;
        ox = 0.5f; oy = 0.5f; oz = 0.0f; dz = n + 1;
    }
    backtrace(a.u[0], a.u[1], a.u[2], x, y, z, h, n, a.dt);
    float val = sample(a.u[axis], x, y, z, ox, oy, oz, h, dx, dy, dz);
    long c = ((long)i * n + j) * n + k;
    if (axis == 2 && a.buoyancy) {
        float tf = 0.5f * (a.T[c - 1] + a.T[c]);
        val = val + a.dt * a.g * (tf - a.t_amb) / (a.t_amb + 273.15f);
    }
    long cl = axis == 0 ? c - (long)n * n : (axis == 1 ? c - n : c - 1);
    float af = face_fraction(a.alpha_s[cl], a.alpha_s[c], a.amin);
    float am = af > a.amin ? af : a.amin;
    float cc = 1.0f / (a.rho * am * h * h * h);
    val = (val + a.dt * cc * a.mf[axis][f]) / (1.0f + a.dt * cc * a.kf[axis][f]);
    out[f] = val;
}

// ---------------------------------------------------------------------------------------------
// divergence rhs (+ Jacobi diagonal)
// ---------------------------------------------------------------------------------------------
struct DivArgs {
    int n;
    long N;
    float h, scale, amin;
    const float *u, *v, *w, *qx, *qy, *qz;
    const float* alpha_s;
    float* rhs;
    float* pdiag;  // null: no diagonal
};

__global__ __launch_bounds__(kBlock) void k_divergence(DivArgs a) {
    long t = (long)xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (t >= a.N) return;
    const int n = a.n;
    const long nn = (long)n * n;
    int i = (int)(t / nn), j = (int)((t / n) % n), k = (int)(t % n);
    const float* as = a.alpha_s;
    float ac = as[t];
    float ax0 = i > 0 ? face_fraction(as[t - nn], ac, a.amin) : 0.0f;
    float ax1 = i < n - 1 ? face_fraction(ac, as[t + nn], a.amin) : 0.0f;
    float ay0 = j > 0 ? face_fraction(as[t - n], ac, a.amin) : 0.0f;
    float ay1 = j < n - 1 ? face_fraction(ac, as[t + n], a.amin) : 0.0f;
    float az0 = k > 0 ? face_fraction(as[t - 1], ac, a.amin) : 0.0f;
    float az1 = k < n - 1 ? face_fraction(ac, as[t + 1], a.amin) : 0.0f;
    long fx0 = ((long)i * n + j) * n + k, fx1 = fx0 + nn;
    long fy0 = ((long)i * (n + 1) + j) * n + k, fy1 = fy0 + n;
    long fz0 = ((long)i * n + j) * (n + 1) + k, fz1 = fz0 + 1;
    float div = ax1 * a.u[fx1] + a.qx[fx1] - ax0 * a.u[fx0] - a.qx[fx0] + ay1 * a.v[fy1] + a.qy[fy1] -
                ay0 * a.v[fy0] - a.qy[fy0] + az1 * a.w[fz1] + a.qz[fz1] - az0 * a.w[fz0] - a.qz[fz0];
    div = div / a.h;
    a.rhs[t] = -a.scale * div;
    if (a.pdiag) {
        float s = 0.0f;
        if (i > 0) s += ax0;
        if (i < n - 1) s += ax1;
        if (j > 0) s += ay0;
        if (j < n - 1) s += ay1;
        if (k > 0) s += az0;
        if (k < n - 1) s += az1;
        float m = s > 1.0e-12f ? s : 1.0e-12f;
        a.pdiag[t] = (a.h * a.h) / m;
    }
}

// ---------------------------------------------------------------------------------------------
// numpy float32 mean of the rhs: buffer chunks of 8192, pairwise sum per chunk (leaves of at most
// 128 elements with 8 interleaved accumulators), chunk sums added in order to a float32 total.
// One workgroup per chunk; the chunk's tree is given as a leaf table and a postfix program.
// ---------------------------------------------------------------------------------------------
struct MeanArgs {
    long N;
    int nchunks;
    const float* rhs;
    float* chunk_sums;
    const int* prog_full;   // [nleaves, nops, leaves(off,len)..., ops...] for a full chunk
    const int* prog_tail;   // same for the last, shorter chunk (null if N % 8192 == 0)
    CgCtl* ctl;
    CgMirror* mirror;
};

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

__global__ __launch_bounds__(kBlock) void k_rhs_mean(MeanArgs a) {
    __shared__ float leaves[kMaxLeaves];
    __shared__ int is_last;
    const int c = blockIdx.x;
    const bool tail = (c == a.nchunks - 1) && a.prog_tail;
    const int* prog = tail ? a.prog_tail : a.prog_full;
    const int nleaves = prog[0], nops = prog[1];
    const int* lv = prog + 2;
    const int* ops = prog + 2 + 2 * nleaves;
    const float* base = a.rhs + (long)c * kChunk;
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
        is_last = (take_ticket(&a.ctl->ticket2) == (unsigned)(a.nchunks - 1));
    }
    __syncthreads();
    if (!is_last) return;
    // the chunk sums are loaded by all threads at once, then added in chunk order by one thread
    __shared__ float cs[kMaxChunksLds];
    for (int q0 = 0; q0 < a.nchunks; q0 += kMaxChunksLds) {
        int m = a.nchunks - q0 < kMaxChunksLds ? a.nchunks - q0 : kMaxChunksLds;
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
    float s = leaves[0];
    float mean = (float)((double)s / (double)a.N);
    a.ctl->mean = mean;
    __hip_atomic_store(&a.mirror->mean, (double)mean, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&a.ctl->ticket2, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------------------------
// projection: q -= (dt/rho) grad p / h on interior faces, walls 0; gridDim.y = axis
// ---------------------------------------------------------------------------------------------
struct ProjArgs {
    int n;
    long F;
    float h, dor;
    const float* p;
    float* q[3];
};

__global__ __launch_bounds__(kBlock) void k_project(ProjArgs a) {
    const int axis = blockIdx.y;
    long f = (long)xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (f >= a.F) return;
    const int n = a.n;
    int i, j, k;
    face_ijk(axis, f, n, i, j, k);
    int idx = axis == 0 ? i : (axis == 1 ? j : k);
    float* q = a.q[axis];
    if (idx == 0 || idx == n) {
        q[f] = 0.0f;
        return;
    }
    long c = ((long)i * n + j) * n + k;
    long cl = axis == 0 ? c - (long)n * n : (axis == 1 ? c - n : c - 1);
    float gp = a.p[c] - a.p[cl];
    q[f] = q[f] - a.dor * gp / a.h;
}

// ---------------------------------------------------------------------------------------------
// accessors and diagnostics
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_cell_velocity(int n, long N, const float* u, const float* v,
                                                          const float* w, float* out) {
    long t = (long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= N) return;
    const long nn = (long)n * n;
    int i = (int)(t / nn), j = (int)((t / n) % n), k = (int)(t % n);
    long fx0 = ((long)i * n + j) * n + k;
    long fy0 = ((long)i * (n + 1) + j) * n + k;
    long fz0 = ((long)i * n + j) * (n + 1) + k;
    out[3 * t + 0] = 0.5f * (u[fx0] + u[fx0 + nn]);
    out[3 * t + 1] = 0.5f * (v[fy0] + v[fy0 + n]);
    out[3 * t + 2] = 0.5f * (w[fz0] + w[fz0 + 1]);
}

__global__ __launch_bounds__(kBlock) void k_face_fraction(int n, long F, int axis, float amin, const float* alpha_s,
                                                          float* out) {
    long f = (long)blockIdx.x * kBlock + threadIdx.x;
    if (f >= F) return;
    int i, j, k;
    face_ijk(axis, f, n, i, j, k);
    int idx = axis == 0 ? i : (axis == 1 ? j : k);
    if (idx == 0 || idx == n) {
        out[f] = 0.0f;
        return;
    }
    long c = ((long)i * n + j) * n + k;
    long cl = axis == 0 ? c - (long)n * n : (axis == 1 ? c - n : c - 1);
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
    long t = (long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= N) return;
    if (flags[t] & FLAG_FIXED) T[t] = tfix[t];
}

}  // namespace

void launch_heat_prep(Scene* s, float dt) {
    HeatPrepArgs a;
    a.N = s->g.N;
    a.T = s->T;
    a.alpha_p = s->P > 0 ? s->alpha_p : nullptr;
    a.tsum = s->tsum;
    a.flags = s->flags;
    a.tfix = s->tfix;
    a.Tw = s->Tw;
    a.rate = (float)s->rate;
    a.t_cap = (float)s->t_cap;
    a.dt = dt;
    a.rate_on = s->rate > 0;
    hipLaunchKernelGGL(k_heat_prep, dim3(nblocks(a.N)), dim3(kBlock), 0, s->stream, a);
}

void launch_advect_T(Scene* s, float dt) {
    AdvectTArgs a;
    a.n = s->g.n;
    a.N = s->g.N;
    a.h = s->g.h;
    a.dt = dt;
    a.dt_hh = dt / (a.h * a.h);
    a.ka = (float)s->prm.kappa_air;
    a.km = (float)s->prm.kappa_mat;
    a.u = s->u[0];
    a.v = s->u[1];
    a.w = s->u[2];
    a.Tw = s->Tw;
    a.alpha_s = s->alpha_s;
    a.flags = s->flags;
    a.tfix = s->tfix;
    a.T2 = s->T2;
    a.kap = s->kap;
    a.hrhs = s->hrhs;
    a.hdiag = s->hdiag;
    hipLaunchKernelGGL(k_advect_T, dim3(nblocks(a.N)), dim3(kBlock), 0, s->stream, a);
}

void launch_g2p(Scene* s) {
    G2PArgs a;
    a.P = s->P;
    a.n = s->g.n;
    a.h = s->g.h;
    a.pos = s->pos;
    a.T_before = s->Tw;
    a.T_after = s->T2;
    a.temp = s->temp;
    a.melt = s->melt;
    a.yield = s->yield;
    a.visc = s->visc;
    a.t_sol = (float)s->prm.t_sol;
    a.t_liq = (float)s->prm.t_liq;
    a.ys = (float)s->prm.yield_solid;
    a.mus = (float)s->prm.mu_solid;
    a.mul = (float)s->prm.mu_liquid;
    a.t_max = (float)s->prm.t_max;
    hipLaunchKernelGGL(k_g2p, dim3(nblocks(a.P)), dim3(kBlock), 0, s->stream, a);
}

void launch_air_faces(Scene* s, float dt) {
    FaceArgs a;
    a.n = s->g.n;
    a.F = s->g.F;
    a.h = s->g.h;
    a.dt = dt;
    a.g = (float)s->prm.gravity;
    a.t_amb = (float)s->prm.t_amb;
    a.rho = (float)s->prm.rho;
    a.amin = (float)s->prm.alpha_f_min;
    a.buoyancy = s->prm.buoyancy;
    for (int ax = 0; ax < 3; ++ax) {
        a.u[ax] = s->u[ax];
        a.out[ax] = s->us[ax];
        a.kf[ax] = s->kf[ax];
        a.mf[ax] = s->mf[ax];
    }
    a.T = s->T;
    a.alpha_s = s->alpha_s;
    hipLaunchKernelGGL(k_air_faces, dim3(nblocks(a.F), 3), dim3(kBlock), 0, s->stream, a);
}

void launch_divergence(Scene* s, const float* const* vel, float scale, float* rhs_out, int with_diag) {
    DivArgs a;
    a.n = s->g.n;
    a.N = s->g.N;
    a.h = s->g.h;
    a.scale = scale;
    a.amin = (float)s->prm.alpha_f_min;
    a.u = vel[0];
    a.v = vel[1];
    a.w = vel[2];
    a.qx = s->q[0];
    a.qy = s->q[1];
    a.qz = s->q[2];
    a.alpha_s = s->alpha_s;
    a.rhs = rhs_out;
    a.pdiag = with_diag ? s->pdiag : nullptr;
    hipLaunchKernelGGL(k_divergence, dim3(nblocks(a.N)), dim3(kBlock), 0, s->stream, a);
}

void launch_rhs_mean(Scene* s, const int* prog_full, const int* prog_tail) {
    MeanArgs a;
    a.N = s->g.N;
    a.nchunks = (int)((s->g.N + kChunk - 1) / kChunk);
    a.rhs = s->rhs;
    a.chunk_sums = s->chunk_sums;
    a.prog_full = prog_full;
    a.prog_tail = prog_tail;
    a.ctl = &s->ctl[1];
    a.mirror = &s->mirror_dev[1];
    hipLaunchKernelGGL(k_rhs_mean, dim3(a.nchunks), dim3(kBlock), 0, s->stream, a);
}

void launch_project(Scene* s, float dt_over_rho) {
    ProjArgs a;
    a.n = s->g.n;
    a.F = s->g.F;
    a.h = s->g.h;
    a.dor = dt_over_rho;
    a.p = s->p;
    for (int ax = 0; ax < 3; ++ax) a.q[ax] = s->us[ax];
    hipLaunchKernelGGL(k_project, dim3(nblocks(a.F), 3), dim3(kBlock), 0, s->stream, a);
}

void launch_cell_velocity(Scene* s, float* out) {
    hipLaunchKernelGGL(k_cell_velocity, dim3(nblocks(s->g.N)), dim3(kBlock), 0, s->stream, s->g.n, s->g.N, s->u[0],
                       s->u[1], s->u[2], out);
}

void launch_face_fraction(Scene* s, int axis, float* out) {
    hipLaunchKernelGGL(k_face_fraction, dim3(nblocks(s->g.F)), dim3(kBlock), 0, s->stream, s->g.n, s->g.F, axis,
                       (float)s->prm.alpha_f_min, s->alpha_s, out);
}

void launch_abs_max(Scene* s, const float* a, long count, unsigned* slot) {
    int nb = nblocks(count);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_abs_max, dim3(nb), dim3(kBlock), 0, s->stream, a, count, slot);
}

void launch_apply_fixed(Scene* s, float* T) {
    hipLaunchKernelGGL(k_apply_fixed, dim3(nblocks(s->g.N)), dim3(kBlock), 0, s->stream, s->g.N, s->flags, s->tfix,
                       T);
}

}  // namespace hfm
