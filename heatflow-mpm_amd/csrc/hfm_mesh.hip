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

}  // namespace hfm
