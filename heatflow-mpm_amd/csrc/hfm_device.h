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
constexpr int kMaxDirBlocks = 4096;     // pass A / one-launch iteration workgroups

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

// The index helpers below (clampi through trilinear) are also compiled for the host, so that a
// host build can drive them with any input under the address and undefined-behaviour sanitizers.
__host__ __device__ __forceinline__ int clampi(int i, int lo, int hi) {
    int a = i > lo ? i : lo;
    return a < hi ? a : hi;
}

// fields.py:5-10 — the lattice index i = int(floor(g)) and fraction f = g - i of g = pos / h - off.
// The index is converted from floor(g) limited to [-2^30, 2^30] (NaN takes the lower limit) and
// f = g - floor(g).  For |g| < 2^30 that is int(floor(g)) and g - float(i) bit for bit.  For any
// other g it keeps the conversion defined and i + 1 free of signed overflow, so every clamped
// lattice index stays in range, including for non-finite or huge departure points.  An unlimited
// conversion of +inf (or of g >= 2^31) saturates to INT_MAX; the compiler forms
// clampi(i + 1, 0, n - 1) as min(max(i, -1) + 1, n - 1), where that +1 wraps to INT_MIN and
// survives the min: a gather 2^31 elements out of bounds.  A NaN g keeps a NaN fraction (NaN
// weights), as in the unlimited form.
__host__ __device__ __forceinline__ void split_floor(float g, int& i, float& f) {
    constexpr float kLim = 1073741824.0f;   // 2^30
    const float fl = floorf(g);
    i = (int)fminf(fmaxf(fl, -kLim), kLim);
    f = g - fl;
}

// x / b rounded correctly, for a scene constant b with rb = RN(1/b): q = RN(x rb), then one
// residual correction q + RN(x - q b) rb (Markstein's theorem: with rb the correctly rounded
// reciprocal and q within an ulp, the corrected quotient is the correctly rounded one).  The
// host enables it (fast = 1) only after checking every significand of one binade of x against
// IEEE division for this b; all intermediate values stay normal for 2^-100 <= |x| <= 2^100, so
// that check covers the whole range, and anything outside it (or fast = 0), NaN and infinities
// included, divides.
__host__ __device__ __forceinline__ float div_const(float x, float b, float rb, int fast) {
    const float ax = fabsf(x);
    if (fast && ((ax >= 0x1p-100f && ax <= 0x1p100f) || x == 0.0f)) {
        const float q = x * rb;
        const float e = __builtin_fmaf(-q, b, x);
        return __builtin_fmaf(e, rb, q);
    }
    return x / b;
}

// The cell size h as the index helpers divide by it: pos / h correctly rounded, through div_const's
// corrected product where the host has checked it for this h (fast), else by division.  A plain
// float converts to the division form.
struct CellSize {
    float h, rh;
    int fast;
    __host__ __device__ CellSize(float h_) : h(h_), rh(1.0f), fast(0) {}
    __host__ __device__ CellSize(float h_, float rh_, int fast_) : h(h_), rh(rh_), fast(fast_) {}
    __host__ __device__ float quotient(float pos) const { return div_const(pos, h, rh, fast); }
};

// a * b for a, b < 2^24 with a product below 2^32 (lattice row offsets): the full-rate 24-bit multiply
__host__ __device__ __forceinline__ unsigned umul24(unsigned a, unsigned b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24(a, b);
#else
    return a * b;
#endif
}

// fields.py:5-10 — g = pos/h - off; i = floor(g); f = g - i (pos/h correctly rounded)
__host__ __device__ __forceinline__ void base_index1(float pos, float off, CellSize hs, int& i, float& f) {
    const float g = hs.quotient(pos) - off;
    split_floor(g, i, f);
}

// (i, j, k) of flat index t of a lattice with ny x nz elements per i-slab, in 32-bit arithmetic
// (every lattice holds at most (1024 + 1) * 1024^2 < 2^32 elements): two divisions instead of the
// three 64-bit divisions of i = t / (ny nz), j = (t / nz) % ny, k = t % nz.
__device__ __forceinline__ void split3(unsigned t, unsigned ny, unsigned nz, int& i, int& j, int& k) {
    const unsigned row = t / nz;   // i ny + j
    const unsigned a = row / ny;
    i = (int)a;
    j = (int)(row - a * ny);
    k = (int)(t - row * nz);
}

// fields.py:5-10 on a quotient q = pos / h already formed: base_index1's operations, so a point's
// quotient serves the lattices of every offset.
__host__ __device__ __forceinline__ void split_index(float q, float off, int& i, float& f) {
    const float g = q - off;
    split_floor(g, i, f);
}

// The values at clamp(bk) and clamp(bk + 1) of a lattice row of nz >= 2 elements, from one load
// of the adjacent pair at b = min(clamp(bk), nz - 2) (in the row for every bk): each index is b
// or b + 1, so each value is the pair's low or high element.
__host__ __device__ __forceinline__ void row_pair(const float* __restrict__ row, int bk, int nz, float& v0, float& v1) {
    const int i0 = clampi(bk, 0, nz - 1), i1 = clampi(bk + 1, 0, nz - 1);
    const int b = i0 < nz - 2 ? i0 : nz - 2;
    unsigned p[2];
    __builtin_memcpy(p, row + b, 8);
    const float lo = __builtin_bit_cast(float, p[0]);
    const float hi = __builtin_bit_cast(float, p[1]);
    v0 = i0 == b ? lo : hi;
    v1 = i1 == b ? lo : hi;
}

// fields.py:25-36 — trilinear sum on a lattice of dims (nx, ny, nz) from the base indices and
// fractions of a point: clamped indices, weights from the unclamped fractions, accumulation order
// dx, dy, dz.  Index arithmetic is 32-bit (a lattice holds at most (1024 + 1) * 1024^2 < 2^32
// elements); the two dz values of a row come from one load (row_pair).
__host__ __device__ __forceinline__ float trilinear(const float* __restrict__ a, int bi, float fx, int bj, float fy, int bk,
                                           float fz, int nx, int ny, int nz) {
    float s = 0.0f;
#pragma unroll
    for (int dx = 0; dx < 2; ++dx) {
        const unsigned ix = (unsigned)clampi(bi + dx, 0, nx - 1);
        const float wx = dx ? fx : 1.0f - fx;
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
            const unsigned iy = (unsigned)clampi(bj + dy, 0, ny - 1);
            const float wy = dy ? fy : 1.0f - fy;
            const unsigned row = umul24(umul24(ix, (unsigned)ny) + iy, (unsigned)nz);
            float v[2];
            row_pair(a + row, bk, nz, v[0], v[1]);
#pragma unroll
            for (int dz = 0; dz < 2; ++dz) {
                const float wz = dz ? fz : 1.0f - fz;
                float w = wx * wy;
                w = w * wz;
                s += w * v[dz];
            }
        }
    }
    return s;
}

// fields.py:25-36 — trilinear sample of a lattice with offsets (ox, oy, oz) at a point
__host__ __device__ __forceinline__ float sample(const float* __restrict__ a, float px, float py, float pz, float ox,
                                        float oy, float oz, CellSize h, int nx, int ny, int nz) {
    int bi, bj, bk;
    float fx, fy, fz;
    base_index1(px, ox, h, bi, fx);
    base_index1(py, oy, h, bj, fy);
    base_index1(pz, oz, h, bk, fz);
    return trilinear(a, bi, fx, bj, fy, bk, fz, nx, ny, nz);
}

// fields.py:51-65 — min and max of the eight lattice values around a point (the MacCormack
// limiter), in the loop order of sample.
__device__ __forceinline__ void value_range(const float* __restrict__ a, float px, float py, float pz, float ox,
                                            float oy, float oz, CellSize h, int nx, int ny, int nz, float& lo,
                                            float& hi) {
    int bi, bj, bk;
    float fx, fy, fz;
    base_index1(px, ox, h, bi, fx);
    base_index1(py, oy, h, bj, fy);
    base_index1(pz, oz, h, bk, fz);
    lo = 1.0e30f;
    hi = -1.0e30f;
#pragma unroll
    for (int dx = 0; dx < 2; ++dx) {
        int ix = clampi(bi + dx, 0, nx - 1);
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
            int iy = clampi(bj + dy, 0, ny - 1);
            const float* row = a + ((long)ix * ny + iy) * nz;
#pragma unroll
            for (int dz = 0; dz < 2; ++dz) {
                const float v = row[clampi(bk + dz, 0, nz - 1)];
                lo = v < lo ? v : lo;
                hi = v > hi ? v : hi;
            }
        }
    }
}

// fields.py:68-72
__host__ __device__ __forceinline__ void sample_velocity(const float* __restrict__ u, const float* __restrict__ v,
                                                const float* __restrict__ w, float px, float py, float pz, CellSize h,
                                                int n, float& vx, float& vy, float& vz) {
    // the three face lattices take offsets 0 or 0.5 per axis: one quotient per axis, two splits
    const float qx = h.quotient(px), qy = h.quotient(py), qz = h.quotient(pz);
    int x0, x5, y0, y5, z0, z5;
    float fx0, fx5, fy0, fy5, fz0, fz5;
    split_index(qx, 0.0f, x0, fx0);
    split_index(qx, 0.5f, x5, fx5);
    split_index(qy, 0.0f, y0, fy0);
    split_index(qy, 0.5f, y5, fy5);
    split_index(qz, 0.0f, z0, fz0);
    split_index(qz, 0.5f, z5, fz5);
    vx = trilinear(u, x0, fx0, y5, fy5, z5, fz5, n + 1, n, n);
    vy = trilinear(v, x5, fx5, y0, fy0, z5, fz5, n, n + 1, n);
    vz = trilinear(w, x5, fx5, y5, fy5, z0, fz0, n, n, n + 1);
}

// fields.py:111-116 — midpoint RK2 departure point
__host__ __device__ __forceinline__ void backtrace(const float* __restrict__ u, const float* __restrict__ v,
                                          const float* __restrict__ w, float& x, float& y, float& z, CellSize h, int n,
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

// grid.py:46-59 — air fraction of an interior face between cells with solid fractions a, b:
// clamp(1 - 0.5 (a + b), amin, 1).  0.5 x is exact, so RN(1 - RN(0.5 x)) = RN(1 - 0.5 x) is one
// fused multiply-add, and for finite values the clamp is v_med3_f32: the same bits in 3
// instructions instead of 5.
__device__ __forceinline__ float face_fraction(float a_left, float a_right, float amin) {
    return __builtin_amdgcn_fmed3f(__builtin_fmaf(-0.5f, a_left + a_right, 1.0f), amin, 1.0f);
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

// One DPP step on a double: each lane of the rows in ROW_MASK receives the value of the lane CTRL
// selects; lanes of the other rows receive an unspecified value.  The trees below read only lanes
// that every step wrote (the quad and mirror steps write all rows; after the row broadcasts only
// lane 63 is read, whose chain passes through rows 1 and 3 alone), so no step needs its
// destination cleared first (a DPP move with a defined old value costs a v_mov per half).
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_double(double x) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave in a fixed tree with DPP moves (no LDS round trips): pairs and
// quads (quad_perm), 8- and 16-lane rows (half mirror, mirror), then row 0 into row 1 and row 2
// into row 3 (row_bcast:15), and rows 0-1 into row 3 (row_bcast:31).  Lane 63 holds the total,
// which is returned to every lane.
__device__ __forceinline__ double wave_sum(double x) {
    x += dpp_double<0xB1>(x);        // quad_perm [1,0,3,2]
    x += dpp_double<0x4E>(x);        // quad_perm [2,3,0,1]
    x += dpp_double<0x141>(x);       // row_half_mirror
    x += dpp_double<0x140>(x);       // row_mirror
    x += dpp_double<0x142, 0xA>(x);  // row_bcast:15
    x += dpp_double<0x143, 0xC>(x);  // row_bcast:31
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 63);
    return __hiloint2double(hi, lo);
}

// wave_sum of NV values at once: the same tree for each value, the values' steps interleaved (no value
// waits on its own previous step while another's can issue)
template <int NV>
__device__ __forceinline__ void wave_sums(double (&x)[NV]) {
#pragma unroll
    for (int q = 0; q < NV; ++q) x[q] += dpp_double<0xB1>(x[q]);
#pragma unroll
    for (int q = 0; q < NV; ++q) x[q] += dpp_double<0x4E>(x[q]);
#pragma unroll
    for (int q = 0; q < NV; ++q) x[q] += dpp_double<0x141>(x[q]);
#pragma unroll
    for (int q = 0; q < NV; ++q) x[q] += dpp_double<0x140>(x[q]);
#pragma unroll
    for (int q = 0; q < NV; ++q) x[q] += dpp_double<0x142, 0xA>(x[q]);
#pragma unroll
    for (int q = 0; q < NV; ++q) x[q] += dpp_double<0x143, 0xC>(x[q]);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const int lo = __builtin_amdgcn_readlane(__double2loint(x[q]), 63);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(x[q]), 63);
        x[q] = __hiloint2double(hi, lo);
    }
}

// Sum of the per-wave values w[0..NW) (NW <= 16, in LDS) in a fixed tree, returned to every lane:
// lane l < NW reads w[l] and the first 16-lane row is reduced with DPP moves.
template <int NW>
__device__ __forceinline__ double sum_wave_values(const double* w) {
    static_assert(NW <= 16, "at most 16 waves per workgroup");
    const int lane = threadIdx.x & 63;
    double x = lane < NW ? w[lane] : 0.0;
    x += dpp_double<0xB1>(x);   // quad_perm [1,0,3,2]
    x += dpp_double<0x4E>(x);   // quad_perm [2,3,0,1]
    x += dpp_double<0x141>(x);  // row_half_mirror
    x += dpp_double<0x140>(x);  // row_mirror: every lane of row 0 holds the sum
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), 0);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 0);
    return __hiloint2double(hi, lo);
}

// sum_wave_values of rows w[0..NV) at once: the same tree for each row, the rows' steps interleaved
template <int NW, int NV, int STRIDE>
__device__ __forceinline__ void sum_wave_values_n(const double (*w)[STRIDE], double (&out)[NV]) {
    static_assert(NW <= 16 && NW <= STRIDE, "at most 16 waves per workgroup");
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = lane < NW ? w[q][lane] : 0.0;
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] += dpp_double<0xB1>(out[q]);
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] += dpp_double<0x4E>(out[q]);
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] += dpp_double<0x141>(out[q]);
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] += dpp_double<0x140>(out[q]);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const int lo = __builtin_amdgcn_readlane(__double2loint(out[q]), 0);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(out[q]), 0);
        out[q] = __hiloint2double(hi, lo);
    }
}

// Sum of one double per thread over a workgroup of BS threads; the result is valid in thread 0.
// The tree is fixed (wave tree, then waves in order), so the value does not depend on timing.
// lds must hold NV x (BS / 64) doubles and is not reused afterwards (the caller ends the kernel).
template <int NV, int BS = kBlock>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*lds)[BS / 64]) {
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double s[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q] = v[q];
    wave_sums<NV>(s);
#pragma unroll
    for (int q = 0; q < NV; ++q)
        if (lane == 0) lds[q][wv] = s[q];
    __syncthreads();
    if (threadIdx.x < 64) sum_wave_values_n<BS / 64, NV>(lds, v);
}

// Element i of an array through a 32-bit byte offset (i * sizeof(T) < 2^32: arrays of at most
// 2^30 four-byte elements), so the access is a global load/store with a scalar base and a 32-bit
// vector offset instead of a 64-bit address computed per lane.
template <class T>
__device__ __forceinline__ T ld32(const T* p, unsigned i) {
    return *(const T*)((const char*)p + i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st32(T* p, unsigned i, T v) {
    *(T*)((char*)p + i * (unsigned)sizeof(T)) = v;
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
