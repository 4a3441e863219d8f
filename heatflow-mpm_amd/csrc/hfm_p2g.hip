// hfm_p2g.hip — particle-to-grid transfer without float atomics (gfx950).
//
// The reference scatters every particle into up to ten lattices with float atomics
// (solver.py:19-38, heat.py:12-18 via fields.py:39-48), so its sums depend on timing.  Here the
// particles are binned by cell (counting sort: integer counts, exclusive scan, then the ids of a
// bin sorted ascending), and every lattice node gathers its own contributions from the bins that
// can reach it, in bin order then particle order.  The result is the same sum of the same terms
// in a fixed order: deterministic run to run.
//
// Particle geometry (solid fraction, flux, relaxation weights and momenta) depends on positions
// and velocities only, so after a re-binning the gather rebuilds every lattice once; while the
// particles stay put it recomputes only the heat sum, whose particle temperatures change every
// substep.  Positions divided by h are stored in bin order at binning time (the same fp32
// divide base_index performs), so the gather reads them contiguously and divides nothing.
//
// A node at index a on an axis where the lattice offset is 1/2 is reached by particles in bins
// a-1, a, a+1; where the offset is 0 (the face axis) by bins a-1, a.  Particles outside the box
// are binned into the nearest edge bin, which is also where their clamped weights land.
#include "hfm_device.h"
#include "hfm_scene.h"

namespace hfm {

namespace {

__device__ __forceinline__ int bin_of(float x, float h, int n) {
    int i;
    float f;
    split_floor(x / h, i, f);
    return clampi(i, 0, n - 1);
}

__global__ __launch_bounds__(kBlock) void k_bin_count(long P, int n, float h, const float* pos, int* cnt, int* key,
                                                      int* slot) {
    long p = (long)blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    int bx = bin_of(pos[3 * p], h, n), by = bin_of(pos[3 * p + 1], h, n), bz = bin_of(pos[3 * p + 2], h, n);
    int c = (bx * n + by) * n + bz;
    key[p] = c;
    slot[p] = atomicAdd(&cnt[c], 1);
}

// Exclusive scan of cnt[0..count) into start[0..count], three passes: per-workgroup totals,
// a scan of those totals, then per-workgroup scans with their offsets.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ int block_exclusive_scan(int v, int* lds, int& total) {
    // Hillis-Steele over the workgroup in LDS (256 entries)
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        int add = threadIdx.x >= off ? lds[threadIdx.x - off] : 0;
        __syncthreads();
        lds[threadIdx.x] += add;
        __syncthreads();
    }
    int incl = lds[threadIdx.x];
    total = lds[kBlock - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_tiles(const int* in, long count, int* tile_sums) {
    __shared__ int lds[kBlock];
    long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q)
        if (base + q < count) s += in[base + q];
    int total;
    block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_scan_sums(int* tile_sums, int ntiles) {
    __shared__ int lds[kBlock];
    int carry = 0;
    for (int b0 = 0; b0 < ntiles; b0 += kBlock) {
        int idx = b0 + threadIdx.x;
        int v = idx < ntiles ? tile_sums[idx] : 0;
        int total;
        int ex = block_exclusive_scan(v, lds, total);
        if (idx < ntiles) tile_sums[idx] = carry + ex;
        carry += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_final(const int* in, long count, const int* tile_offsets, int* out) {
    __shared__ int lds[kBlock];
    long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanItems;
    int vals[kScanItems];
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        vals[q] = base + q < count ? in[base + q] : 0;
        s += vals[q];
    }
    int total;
    int ex = block_exclusive_scan(s, lds, total) + tile_offsets[blockIdx.x];
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        if (base + q <= count) out[base + q] = ex;  // out has count + 1 entries
        ex += vals[q];
    }
}

__global__ __launch_bounds__(kBlock) void k_bin_fill(long P, const int* key, const int* slot, const int* start,
                                                     int* sorted) {
    long p = (long)blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    sorted[start[key[p]] + slot[p]] = (int)p;
}

// Ascending particle ids inside each bin (bins hold a handful of particles).
__global__ __launch_bounds__(kBlock) void k_bin_sort(long N, const int* start, int* sorted) {
    long c = (long)blockIdx.x * kBlock + threadIdx.x;
    if (c >= N) return;
    int b = start[c], e = start[c + 1];
    for (int x = b + 1; x < e; ++x) {
        int v = sorted[x];
        int y = x - 1;
        while (y >= b && sorted[y] > v) {
            sorted[y + 1] = sorted[y];
            --y;
        }
        sorted[y + 1] = v;
    }
}

// pos / h in bin order; marks the particle geometry as changed for the next gather and the
// cell lattice's contribution list as stale
__global__ __launch_bounds__(kBlock) void k_bin_pack(long P, float h, const float* pos, const int* sorted, float* qpos,
                                                     CgCtl* ctl0) {
    long s = (long)blockIdx.x * kBlock + threadIdx.x;
    if (s >= P) return;
    const long p = sorted[s];
#pragma unroll
    for (int d = 0; d < 3; ++d) qpos[3 * s + d] = pos[3 * p + d] / h;
    if (s == 0) {
        ctl0->geom_dirty = 1;
        ctl0->csr_ready = 0;
    }
}

// ---------------------------------------------------------------------------------------------
// gather: sub-range 0 = cell lattice (alpha_s, alpha_p, tsum), 1..3 = face lattices (q, K, m)
// ---------------------------------------------------------------------------------------------
// fields.py:5-10 with q = pos / h already formed: g = q - off; i = floor(g); f = g - i
__device__ __forceinline__ void base_index_q(float q, float off, int& i, float& f) { split_index(q, off, i, f); }

// Weight with which the particle at (qx,qy,qz) = pos / h reaches node (ni,nj,nk) of a lattice
// with offsets (ox,oy,oz) and dims (dx,dy,dz): the sum over the corners whose clamped index is
// the node, each added separately in the corner order of fields.py:41-48.
__device__ __forceinline__ void node_terms(float qx, float qy, float qz, float ox, float oy, float oz, int dx, int dy,
                                           int dz, int ni, int nj, int nk, float* w_out, int& count) {
    int bi, bj, bk;
    float fx, fy, fz;
    base_index_q(qx, ox, bi, fx);
    base_index_q(qy, oy, bj, fy);
    base_index_q(qz, oz, bk, fz);
    count = 0;
#pragma unroll
    for (int cx = 0; cx < 2; ++cx) {
        if (clampi(bi + cx, 0, dx - 1) != ni) continue;
        float wx = cx ? fx : 1.0f - fx;
#pragma unroll
        for (int cy = 0; cy < 2; ++cy) {
            if (clampi(bj + cy, 0, dy - 1) != nj) continue;
            float wy = cy ? fy : 1.0f - fy;
#pragma unroll
            for (int cz = 0; cz < 2; ++cz) {
                if (clampi(bk + cz, 0, dz - 1) != nk) continue;
                float wz = cz ? fz : 1.0f - fz;
                float w = wx * wy;
                w_out[count++] = w * wz;
            }
        }
    }
}

// The particles that can reach node (ni, nj, nk) of a lattice (offsets ox, oy, oz; dims dx, dy,
// dz), in bin order: f(slot, w, cnt) with the node's corner terms w[0..cnt) of each particle that
// has any (bins of the 3 x 3 x 3 neighbourhood, 2 along a face axis; fields.py:41-48).
template <class F>
__device__ __forceinline__ void walk_node(const int* bin_start, const float* qpos, int n, float ox, float oy, float oz,
                                          int dx, int dy, int dz, int ni, int nj, int nk, F&& f) {
    int lo_x = ni - 1, hi_x = ox == 0.0f ? ni : ni + 1;
    int lo_y = nj - 1, hi_y = oy == 0.0f ? nj : nj + 1;
    int lo_z = nk - 1, hi_z = oz == 0.0f ? nk : nk + 1;
    lo_x = lo_x < 0 ? 0 : lo_x; hi_x = hi_x > n - 1 ? n - 1 : hi_x;
    lo_y = lo_y < 0 ? 0 : lo_y; hi_y = hi_y > n - 1 ? n - 1 : hi_y;
    lo_z = lo_z < 0 ? 0 : lo_z; hi_z = hi_z > n - 1 ? n - 1 : hi_z;
    for (int bx = lo_x; bx <= hi_x; ++bx)
        for (int by = lo_y; by <= hi_y; ++by) {
            long row = ((long)bx * n + by) * n;
            int b = bin_start[row + lo_z], e = bin_start[row + hi_z + 1];
            for (int s = b; s < e; ++s) {
                const float* qp = qpos + 3 * (long)s;
                float w[8];
                int cnt;
                node_terms(qp[0], qp[1], qp[2], ox, oy, oz, dx, dy, dz, ni, nj, nk, w, cnt);
                if (cnt == 0) continue;
                f(s, w, cnt);
            }
        }
}

__device__ __forceinline__ void cell_of(long node, int n, int& i, int& j, int& k) {
    i = (int)(node / ((long)n * n));
    j = (int)((node / n) % n);
    k = (int)(node % n);
}

template <class Src>
__global__ __launch_bounds__(kBlock) void k_p2g_gather(Src src) {
    HFM_ENTRY(src);
    const int lat = sub;  // 0 cell, 1 + axis for faces
    const bool full = a.ctl[0].geom_dirty != 0;
    if (!full && lat != 0) return;  // face lattices unchanged since the last gather
    // the cell lattice of unchanged particles: summed from its contribution list once that exists;
    // in the first substep without it, this gather also counts each node's terms for the list
    if (!full && a.ctl[0].csr_ready) return;
    const bool count_terms = !full && lat == 0;
    const int n = a.n;
    const long count = lat == 0 ? a.N : a.F;
    if (lb >= nblocks(count)) return;  // a batch of one launches the face lattices' workgroup count for all four
    const long node = (long)xcd_tile(lb, nblocks(count)) * kBlock + threadIdx.x;
    if (node >= count) return;
    int dx = n, dy = n, dz = n;
    float ox = 0.5f, oy = 0.5f, oz = 0.5f;
    const int axis = lat - 1;
    if (axis == 0) { dx = n + 1; ox = 0.0f; }
    if (axis == 1) { dy = n + 1; oy = 0.0f; }
    if (axis == 2) { dz = n + 1; oz = 0.0f; }
    int ni = (int)(node / ((long)dy * dz)), nj = (int)((node / dz) % dy), nk = (int)(node % dz);
    float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f;  // cell: alpha, tsum ; face: q, K, m
    int terms = 0;
    walk_node(a.bin_start, a.qpos, n, ox, oy, oz, dx, dy, dz, ni, nj, nk, [&](int s, const float* w, int cnt) {
        const int p = a.sorted[s];
        terms += cnt;
        if (lat == 0) {
            const float vt = a.frac * a.temp[p];
            if (full)
                for (int q = 0; q < cnt; ++q) acc0 += w[q] * a.frac;
            for (int q = 0; q < cnt; ++q) acc1 += w[q] * vt;
        } else {
            float vq = a.frac * a.vel[3 * p + axis];
            float vm = a.k0 * a.vel[3 * p + axis];
            for (int q = 0; q < cnt; ++q) {
                acc0 += w[q] * vq;
                acc1 += w[q] * a.k0;
                acc2 += w[q] * vm;
            }
        }
    });
    if (lat == 0) {
        a.tsum[node] = acc1;
        if (count_terms) a.csr_start[node] = terms;  // scanned in place into the list's starts
        if (full) {
            a.alpha_p[node] = acc0;
            a.alpha_s[node] = acc0 + a.st_alpha[node];
        }
    } else {
        int idx = axis == 0 ? ni : (axis == 1 ? nj : nk);
        a.q[axis][node] = (idx == 0 || idx == n) ? 0.0f : acc0;
        a.kf[axis][node] = acc1 + a.st_k[axis][node];
        a.mf[axis][node] = acc2 + a.st_m[axis][node];
    }
}

// The contribution list of the cell lattice: each node's terms in the gather's order (bins,
// particles, corners) at the node's start; then the list is marked ready.
// (Every particle has 8 corner terms, each on one node within a bin of its own: 8 P terms in all;
// the bound below only keeps a stray count from writing past the list.)
__global__ __launch_bounds__(kBlock) void k_csr_fill(long N, int n, const int* bin_start, const float* qpos,
                                                     const int* sorted, const int* start, int* csr_p, float* csr_w,
                                                     long cap, CgCtl* ctl0) {
    const long node = (long)blockIdx.x * kBlock + threadIdx.x;
    if (node == 0 && threadIdx.x == 0) ctl0->csr_ready = 1;   // read by later launches only
    if (node >= N) return;
    int ni, nj, nk;
    cell_of(node, n, ni, nj, nk);
    int e = start[node];
    walk_node(bin_start, qpos, n, 0.5f, 0.5f, 0.5f, n, n, n, ni, nj, nk, [&](int s, const float* w, int cnt) {
        const int p = sorted[s];
        for (int q = 0; q < cnt; ++q) {
            if (e < cap) {
                csr_p[e] = p;
                csr_w[e] = w[q];
            }
            ++e;
        }
    });
}

// The cell lattice's particle temperature sum from the contribution list: the gather's terms
// w * (frac * temp) in the gather's order, so the same bits.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_p2g_listed(Src src) {
    HFM_ENTRY(src);
    if (a.P == 0 || !a.ctl[0].csr_ready || a.ctl[0].geom_dirty) return;
    const long N = a.N;
    const long node = (long)xcd_tile(lb, nblocks(N)) * kBlock + threadIdx.x;
    if (node >= N) return;
    const long cap = 8L * a.P;
    const int b = a.csr_start[node];
    const int e = a.csr_start[node + 1] < cap ? a.csr_start[node + 1] : (int)cap;
    // the loads of kLoads terms are issued together (the list is read twice-indirect); the sum
    // still adds the terms one by one in list order
    constexpr int kLoads = 8;
    float acc1 = 0.0f;
    int i = b;
    for (; i + kLoads <= e; i += kLoads) {
        float vt[kLoads], w[kLoads];
#pragma unroll
        for (int u = 0; u < kLoads; ++u) {
            vt[u] = a.temp[a.csr_p[i + u]];
            w[u] = a.csr_w[i + u];
        }
#pragma unroll
        for (int u = 0; u < kLoads; ++u) acc1 += w[u] * (a.frac * vt[u]);
    }
    for (; i < e; ++i) {
        const float vt = a.frac * a.temp[a.csr_p[i]];
        acc1 += a.csr_w[i] * vt;
    }
    a.tsum[node] = acc1;
}

// Debug view of the binning (hfm_get_particle_bins): for the particle at bin-order slot s, its bin
// and, from the stored pos / h exactly as the gather reads it, the base index and fraction on the
// cell, x-, y- and z-face lattices (fields.py:5-10, offsets of fields.py:84-90), written at the
// particle's own index.
__global__ __launch_bounds__(kBlock) void k_particle_bins(long P, const int* sorted, const float* qpos, const int* key,
                                                          int* bin, int* base, float* frac) {
    const long s = (long)blockIdx.x * kBlock + threadIdx.x;
    if (s >= P) return;
    const long p = sorted[s];
    bin[p] = key[p];
    const float off[4][3] = {{0.5f, 0.5f, 0.5f}, {0.0f, 0.5f, 0.5f}, {0.5f, 0.0f, 0.5f}, {0.5f, 0.5f, 0.0f}};
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int i;
            float f;
            split_index(qpos[3 * s + d], off[l][d], i, f);
            base[12 * p + 3 * l + d] = i;
            frac[12 * p + 3 * l + d] = f;
        }
}

}  // namespace

void launch_particle_bins(const Scene* s, hipStream_t st, int* bin, int* base, float* frac) {
    const long P = s->dev.P;
    if (P > 0)
        hipLaunchKernelGGL(k_particle_bins, dim3(nblocks(P)), dim3(kBlock), 0, st, P, s->dev.sorted, s->dev.qpos,
                           s->pkey, bin, base, frac);
}

void launch_bin_particles(Scene* s, hipStream_t st) {
    const long N = s->N, P = s->dev.P;
    (void)hipMemsetAsync(s->bin_cnt, 0, (N + 1) * sizeof(int), st);
    if (P > 0)
        hipLaunchKernelGGL(k_bin_count, dim3(nblocks(P)), dim3(kBlock), 0, st, P, s->n, s->dev.h, s->dev.pos,
                           s->bin_cnt, s->pkey, s->pslot);
    const long count = N;  // scan N counts into N + 1 starts
    const int ntiles = (int)((count + 1 + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kBlock), 0, st, s->bin_cnt, count, s->scan_tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, st, s->scan_tmp, ntiles);
    hipLaunchKernelGGL(k_scan_final, dim3(ntiles), dim3(kBlock), 0, st, s->bin_cnt, count, s->scan_tmp,
                       s->dev.bin_start);
    if (P > 0) {
        hipLaunchKernelGGL(k_bin_fill, dim3(nblocks(P)), dim3(kBlock), 0, st, P, s->pkey, s->pslot,
                           s->dev.bin_start, s->dev.sorted);
        hipLaunchKernelGGL(k_bin_sort, dim3(nblocks(N)), dim3(kBlock), 0, st, N, s->dev.bin_start, s->dev.sorted);
        hipLaunchKernelGGL(k_bin_pack, dim3(nblocks(P)), dim3(kBlock), 0, st, P, s->dev.h, s->dev.pos,
                           s->dev.sorted, s->dev.qpos, &s->dev.ctl[0]);
    }
    s->bins_dirty = 0;
}

void launch_p2g(const Launch& L) { HFM_LAUNCH(k_p2g_gather, L, kBlock, nullptr, nullptr); }

// After the gather that counted each node's terms into csr_start: scan the counts in place into
// starts (N + 1), then fill the list.
void launch_csr_build(Scene* s, hipStream_t st) {
    const long N = s->N;
    const int ntiles = (int)((N + 1 + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kBlock), 0, st, s->dev.csr_start, N, s->scan_tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, st, s->scan_tmp, ntiles);
    hipLaunchKernelGGL(k_scan_final, dim3(ntiles), dim3(kBlock), 0, st, s->dev.csr_start, N, s->scan_tmp,
                       s->dev.csr_start);
    hipLaunchKernelGGL(k_csr_fill, dim3(nblocks(N)), dim3(kBlock), 0, st, N, s->n, s->dev.bin_start, s->dev.qpos,
                       s->dev.sorted, s->dev.csr_start, s->dev.csr_p, s->dev.csr_w, 8L * s->dev.P, &s->dev.ctl[0]);
}

void launch_p2g_listed(const Launch& L) { HFM_LAUNCH(k_p2g_listed, L, kBlock, nullptr, nullptr); }

}  // namespace hfm
