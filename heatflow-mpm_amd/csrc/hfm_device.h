// hfm_device.h — device-side building blocks shared by the substep kernels (gfx950).
//
// Arithmetic rule for the whole library: every expression keeps the operation order of the
// reference's Warp kernels and the library is compiled with -ffp-contract=off, so each fp32
// multiply and add is rounded on its own exactly as written there.  Reductions accumulate fp32
// products in double in a fixed tree (independent of scheduling) and round the result to fp32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hfm {

constexpr int kBlock = 256;      // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kXcds = 8;         // MI355X: 8 XCDs, workgroups dispatched round-robin over them
constexpr int kMaxPcgBlocks = 1024;     // pass B / init workgroups (4 per CU); bounds their partial sums
constexpr int kMaxDirBlocks = 2048;     // pass A workgroups

__host__ __device__ inline int nblocks(long count) { return (int)((count + kBlock - 1) / kBlock); }

// cell flags (1 byte per cell)
constexpr uint8_t FLAG_FIXED = 1;
constexpr uint8_t FLAG_SOURCE = 2;

// Map the hardware workgroup id to a logical tile id so that each XCD (which receives every
// 8th workgroup) works on one contiguous range of tiles and keeps its halo reads in its own L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    int per = nb / kXcds, rem = nb % kXcds;
    int x = b % kXcds, q = b / kXcds;
    int base = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
    return base + q;
}

__device__ __forceinline__ int clampi(int i, int lo, int hi) {
    int a = i > lo ? i : lo;
    return a < hi ? a : hi;
}

// fields.py:5-10 — g = pos/h - off; i = floor(g); f = g - i (true IEEE divide)
__device__ __forceinline__ void base_index1(float pos, float off, float h, int& i, float& f) {
    float g = pos / h - off;
    i = (int)floorf(g);
    f = g - (float)i;
}

// fields.py:25-36 — trilinear sample on a lattice of dims (nx, ny, nz), clamped indices, weights
// from the unclamped fractions, accumulation order dx, dy, dz.
__device__ __forceinline__ float sample(const float* __restrict__ a, float px, float py, float pz, float ox,
                                        float oy, float oz, float h, int nx, int ny, int nz) {
    int bi, bj, bk;
    float fx, fy, fz;
    base_index1(px, ox, h, bi, fx);
    base_index1(py, oy, h, bj, fy);
    base_index1(pz, oz, h, bk, fz);
    float s = 0.0f;
#pragma unroll
    for (int dx = 0; dx < 2; ++dx) {
        int ix = clampi(bi + dx, 0, nx - 1);
        float wx = dx ? fx : 1.0f - fx;
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
            int iy = clampi(bj + dy, 0, ny - 1);
            float wy = dy ? fy : 1.0f - fy;
            const float* row = a + ((long)ix * ny + iy) * nz;
#pragma unroll
            for (int dz = 0; dz < 2; ++dz) {
                int iz = clampi(bk + dz, 0, nz - 1);
                float wz = dz ? fz : 1.0f - fz;
                float w = wx * wy;
                w = w * wz;
                s += w * row[iz];
            }
        }
    }
    return s;
}

// fields.py:68-72
__device__ __forceinline__ void sample_velocity(const float* __restrict__ u, const float* __restrict__ v,
                                                const float* __restrict__ w, float px, float py, float pz, float h,
                                                int n, float& vx, float& vy, float& vz) {
    vx = sample(u, px, py, pz, 0.0f, 0.5f, 0.5f, h, n + 1, n, n);
    vy = sample(v, px, py, pz, 0.5f, 0.0f, 0.5f, h, n, n + 1, n);
    vz = sample(w, px, py, pz, 0.5f, 0.5f, 0.0f, h, n, n, n + 1);
}

// fields.py:111-116 — midpoint RK2 departure point
__device__ __forceinline__ void backtrace(const float* __restrict__ u, const float* __restrict__ v,
                                          const float* __restrict__ w, float& x, float& y, float& z, float h, int n,
                                          float dt) {
    float vx, vy, vz;
    sample_velocity(u, v, w, x, y, z, h, n, vx, vy, vz);
    float hd = 0.5f * dt;
    float x1 = x - hd * vx, y1 = y - hd * vy, z1 = z - hd * vz;
    sample_velocity(u, v, w, x1, y1, z1, h, n, vx, vy, vz);
    x = x - dt * vx;
    y = y - dt * vy;
    z = z - dt * vz;
}

// grid.py:46-59 — air fraction of an interior face between cells with solid fractions a, b
__device__ __forceinline__ float face_fraction(float a_left, float a_right, float amin) {
    float a = 0.5f * (a_left + a_right);
    float v = 1.0f - a;
    return v < amin ? amin : (v > 1.0f ? 1.0f : v);
}

// heat.py:44-47
__device__ __forceinline__ float conductivity(float alpha, float ka, float km) {
    float c = alpha / 0.6f;
    c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
    return ka + (km - ka) * c;
}

// heat.py:50-52
__device__ __forceinline__ float face_kappa(float ka, float kb) { return 2.0f * ka * kb / (ka + kb + 1.0e-12f); }

// ---------------------------------------------------------------------------------------------
// Deterministic grid reductions
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum of one double per thread over a workgroup of BS threads; the result is valid in thread 0.
// The tree is fixed (lane butterfly, then waves in order), so the value does not depend on
// timing.  lds must hold NV x (BS / 64) doubles.
template <int NV, int BS = kBlock>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*lds)[BS / 64]) {
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = wave_sum(v[q]);
        if (lane == 0) lds[q][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = lds[q][0];
#pragma unroll
            for (int w = 1; w < BS / 64; ++w) s += lds[q][w];
            v[q] = s;
        }
    }
    __syncthreads();
}

// Cross-workgroup traffic of the reductions uses relaxed agent-scope atomics only: they are
// coherent across the XCDs without the L2 write-back / invalidate that an agent-scope release /
// acquire fence costs.  Ordering between a partial's store and the arrival ticket comes from
// waiting for the store to complete (s_waitcnt 0) before the ticket is taken.
__device__ __forceinline__ void store_agent(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_agent(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void wait_stores() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
// arrival ticket: returns the number of workgroups that arrived before this one
__device__ __forceinline__ unsigned take_ticket(unsigned* ticket) {
    wait_stores();
    return __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace hfm
