// hfm_mesh.hip — inside test of points against a closed triangle mesh, for seeding material
// particles in a mesh (shapes.py:13-17 `_inside`, launched by shapes.py:35-44 `sample_interior`).
//
// The reference asks Warp for the sign of each point by the mesh's winding number
// (mesh_query_point_sign_winding_number, accuracy 2.0, threshold 0.5: inside where the winding
// number exceeds 0.5).  Here the generalized winding number is computed exactly,
//   w(p) = (1 / 4 pi) sum over triangles of the solid angle the triangle subtends at p,
// with each solid angle from the formula of Van Oosterom and Strackee (IEEE Trans. Biomed. Eng.
// 30, 1983): for a, b, c the vertices relative to p,
//   tan(Omega / 2) = a . (b x c) / (|a||b||c| + (a . b)|c| + (b . c)|a| + (c . a)|b|),
// so Omega = 2 atan2(numerator, denominator).  w is 1 inside and 0 outside a closed mesh whose
// triangles wind counter-clockwise seen from outside (the fast-winding-number approximation of
// Warp's query differs from it only by its approximation error, near the surface).
//
// One thread per point; each workgroup stages 256 triangles (9 floats each) at a time in LDS and
// every thread sweeps them with a broadcast read; a chunk's solid angles are summed in fp32 and
// the chunks in double.  Compute bound (~60 VALU per point-triangle pair with the atan2).
#include <hip/hip_runtime.h>

#include <mutex>

#include "hfm_device.h"
#include "hfm_scene.h"

namespace hfm {

namespace {

constexpr int kTriChunk = 256;

__device__ __forceinline__ float solid_angle(float ax, float ay, float az, float bx, float by, float bz, float cx,
                                             float cy, float cz) {
    const float la = sqrtf(ax * ax + ay * ay + az * az);
    const float lb = sqrtf(bx * bx + by * by + bz * bz);
    const float lc = sqrtf(cx * cx + cy * cy + cz * cz);
    const float num = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    const float ab = ax * bx + ay * by + az * bz;
    const float bc = bx * cx + by * cy + bz * cz;
    const float ca = cx * ax + cy * ay + cz * az;
    const float den = la * lb * lc + ab * lc + bc * la + ca * lb;
    return 2.0f * atan2f(num, den);
}

__global__ __launch_bounds__(kBlock) void k_winding(const float* __restrict__ tri, long nt, const float* __restrict__ pts,
                                                    long np, float threshold, int* __restrict__ flags,
                                                    float* __restrict__ wn) {
    __shared__ float T[kTriChunk * 9];
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    const long ic = i < np ? i : np - 1;   // threads past the end sweep the last point, unstored
    const float px = pts[3 * ic], py = pts[3 * ic + 1], pz = pts[3 * ic + 2];
    double acc = 0.0;
    for (long t0 = 0; t0 < nt; t0 += kTriChunk) {
        const int m = (int)(nt - t0 < kTriChunk ? nt - t0 : kTriChunk);
        __syncthreads();   // the previous chunk has been read by every thread
        for (int q = threadIdx.x; q < 9 * m; q += kBlock) T[q] = tri[9 * t0 + q];
        __syncthreads();
        float part = 0.0f;
        for (int j = 0; j < m; ++j) {
            const float* v = T + 9 * j;
            part += solid_angle(v[0] - px, v[1] - py, v[2] - pz, v[3] - px, v[4] - py, v[5] - pz, v[6] - px,
                                v[7] - py, v[8] - pz);
        }
        acc += (double)part;
    }
    const double w = acc * (0.25 / 3.14159265358979323846);
    if (i < np) {
        flags[i] = w > (double)threshold ? 1 : 0;
        if (wn) wn[i] = (float)w;
    }
}

}  // namespace

void launch_winding(const float* tri, long nt, const float* pts, long np, float threshold, int* flags, float* wn,
                    hipStream_t st) {
    if (np <= 0) return;
    hipLaunchKernelGGL(k_winding, dim3((unsigned)((np + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, tri, nt, pts, np,
                       threshold, flags, wn);
}

// ---------------------------------------------------------------------------------------------
// The inside test without device allocations (hfm_mesh_inside_noalloc).  Its device copies of the
// triangles, points, flags and winding numbers live in `scratch`, one 8 MiB array per GPU held for
// the life of the process.  Host code never passes the array's name to a kernel or a copy (in host
// code it is a stand-in, not the device address): the address of each GPU's copy comes from
// hipGetSymbolAddress, taken once per device with that device current, and is kept in
// g_scratch_dev.  Every call on a GPU shares that GPU's array and leaves its contents there; nothing
// in it needs to start at zero (the kernel writes every flag and winding number it is asked for).
// Calls that overlap on one GPU (from several host threads) would race on the array, so the calls
// of a process are serialised by one mutex.
// ---------------------------------------------------------------------------------------------
constexpr long kScratchWords = 1L << 21;      // 8 MiB: 9 nt + 5 np + 3 * 64 words at most
static __device__ __align__(256) float scratch[kScratchWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxDevices] = {};

long round64(long w) { return (w + 63) & ~63L; }
}  // namespace

}  // namespace hfm

extern "C" int hfm_mesh_inside_noalloc(const float* tri, int64_t nt, const float* points, int64_t np, float threshold,
                                       int32_t* flags, float* winding, int device) {
    using namespace hfm;
    if (nt < 0 || np < 0 || (nt > 0 && !tri) || (np > 0 && (!points || !flags)) || !(threshold == threshold)) {
        set_error("bad mesh-inside arguments (triangles %lld, points %lld)", (long long)nt, (long long)np);
        return HFM_ERR_ARG;
    }
    if (nt > kScratchWords || np > kScratchWords ||
        round64(9 * nt) + round64(3 * np) + round64(np) + round64(np) > kScratchWords) {
        set_error("hfm_mesh_inside_noalloc: 9 x %lld triangle + 5 x %lld point words exceed its %ld-word scratch "
                  "(use hfm_mesh_inside)", (long long)nt, (long long)np, kScratchWords);
        return HFM_ERR_ARG;
    }
    if (np == 0) return HFM_OK;
    int rc = check_finite_host(tri, 9 * nt, "triangles");
    if (!rc) rc = check_finite_host(points, 3 * np, "points");
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;
    if (device < 0 || device >= ndev || device >= kMaxDevices) {
        set_error("device %d not available (%d HIP devices)", device, ndev);
        return HFM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess) e = hipSetDevice(device);
    if (e == hipSuccess && !g_scratch_dev[device]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e == hipSuccess) g_scratch_dev[device] = (float*)ptr;
    }
    float* base = g_scratch_dev[device];
    float* d_tri = base;
    float* d_pts = d_tri + round64(9 * nt);
    int* d_flags = (int*)(d_pts + round64(3 * np));
    float* d_wn = (float*)d_flags + round64(np);
    if (e == hipSuccess && nt > 0) e = hipMemcpy(d_tri, tri, 9 * nt * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_pts, points, 3 * np * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_winding(d_tri, nt, d_pts, np, threshold, d_flags, winding ? d_wn : nullptr, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(flags, d_flags, np * sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess && winding) e = hipMemcpy(winding, d_wn, np * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipSetDevice(prev);
    if (e != hipSuccess) {
        set_error("hfm_mesh_inside_noalloc: %s", hipGetErrorString(e));
        return HFM_ERR_HIP;
    }
    return HFM_OK;
}
